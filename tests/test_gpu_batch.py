"""Batched cases (one launch per step over many grids, hipGraph replays) against the
same cases run one at a time through lbm_step_db: populations bit-identical, forces
to FP64 summation order. Covers graph replays plus a remainder, the persistent grid,
3-D batches with KBC / TRT / RLBM + LES, replaced states, the autotuner and the API's errors."""
import numpy as np
import pytest

from tests import _util as U

pytestmark = pytest.mark.gpu


def _small_cylinders(n=5, seed=7):
    from mojo_lbm_amd.cylinder import CylinderCase

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(64, 161))
        ny = int(rng.integers(24, 49))
        d = float(rng.uniform(6.0, 0.4 * ny))
        out.append(CylinderCase(nx=nx, ny=ny, diameter=d, cx=nx / 5.0, cy=ny / 2.0 + rng.uniform(-1.0, 1.0),
                                re=float(rng.uniform(20.0, 100.0))))
    return out


def _batch(cases, lattice, config, tile=None):
    from mojo_lbm_amd.batch import CaseBatch

    b = CaseBatch(lattice, [(c.nx, c.ny, 1) for c in cases], [c.tau for c in cases], config, tile=tile)
    for i, c in enumerate(cases):
        flags, bc, links = c.arrays(lattice)
        b.set_case(i, flags, bc, links)
    return b


def _single(b, c, f0, tau, steps):
    """Case c of the batch alone through lbm_step_db (same padded grid and tile)."""
    d, flags, bc, nodes, qd = b.export_case(c)
    run = U.DeviceRun(d, f0, bc, flags)
    if len(nodes):
        run.set_links(nodes, qd)
    for _ in range(steps):
        run.step_db(tau)
    return run.f_host(), (run.link_force() if len(nodes) else None)


def _config(**kw):
    from mojo_lbm_amd import DoubleBufferConfig, Flags

    return DoubleBufferConfig(BCs={Flags.EQUILIBRIUM}, **kw)


@pytest.mark.parametrize("grid_blocks,graph_steps,steps", [(0, 50, 123), (16, 0, 37), (24, 20, 64)])
def test_batch_matches_single_runs(gpu, grid_blocks, graph_steps, steps):
    from mojo_lbm_amd import get_D2Q9

    print(f"case batch grid={grid_blocks} graph={graph_steps} steps={steps}", flush=True)
    cases = _small_cylinders()
    b = _batch(cases, get_D2Q9(), _config())
    b.set_option(grid_blocks=grid_blocks, graph_steps=graph_steps)
    b.init_equilibrium()
    f0 = [b.get_f(i).cpu().numpy() for i in range(len(cases))]
    b.run(steps)
    forces = b.force()
    for i, c in enumerate(cases):
        got = b.get_f(i).cpu().numpy()
        ref, fref = _single(b, i, f0[i], c.tau, steps)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), f"case {i} differs"
        assert np.allclose(forces[i], fref, rtol=1e-12, atol=1e-15)


def test_batch_links_added_after_graph_run(gpu):
    """A graph captured while no case had links must not be replayed once set_case adds them:
    the replays would run the kernel without links and report a force nobody computed."""
    from mojo_lbm_amd import get_D2Q9

    print("case batch: links added after a graph run", flush=True)
    lat = get_D2Q9()
    cases = _small_cylinders(3, seed=11)
    b = _batch(cases, lat, _config())
    for i, c in enumerate(cases):
        flags, bc, _ = c.arrays(lat)
        b.set_case(i, flags, bc, None)  # halfway walls only
    b.set_option(graph_steps=10)
    b.init_equilibrium()
    b.run(20)  # two graph replays without links
    for i, c in enumerate(cases):
        flags, bc, links = c.arrays(lat)
        b.set_case(i, flags, bc, links)
    f0 = [b.get_f(i).cpu().numpy() for i in range(len(cases))]
    b.run(20)
    forces = b.force()
    for i, c in enumerate(cases):
        ref, fref = _single(b, i, f0[i], c.tau, 20)
        assert fref is not None
        assert np.array_equal(b.get_f(i).cpu().numpy().view(np.uint8), ref.view(np.uint8)), f"case {i} differs"
        assert np.allclose(forces[i], fref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("collision", ["KBC", "TRT", "RLBM"])
def test_batch_3d_les_from_perturbed_states(gpu, collision):
    import torch

    from mojo_lbm_amd import Flags, get_D3Q19
    from mojo_lbm_amd.batch import CaseBatch
    from mojo_lbm_amd.geometry import sphere_links_rowmajor, sphere_mask

    lat = get_D3Q19()
    sizes = [(24, 16, 16), (16, 16, 8), (40, 24, 16)]
    taus = [0.52, 0.6, 0.55]
    print(f"case batch 3-D {collision} + LES", flush=True)
    cfg = _config(collision_op=collision, LES=True, DDF_shift=True)
    b = CaseBatch(lat, sizes, taus, cfg, tile=(8, 8, 8))
    rng = np.random.default_rng(5)
    for i, n in enumerate(sizes):
        flags = np.zeros(n, np.uint8)
        bc = np.zeros(n + (4,), np.float64)
        bc[..., 3] = 1.0
        flags[:, 0] = Flags.SOLID
        flags[:, -1] = Flags.SOLID
        bc[:, -1, :, 0] = 0.05  # moving lid
        c = (n[0] / 3 + 0.3, n[1] / 2 - 0.2, n[2] / 2 + 0.1)
        inside = sphere_mask(n, c, 3.1, 3)
        flags[inside] = Flags.SOLID
        links = sphere_links_rowmajor(lat.directions, flags, c, 3.1, inside)
        b.set_case(i, flags, bc, links)
    b.init_equilibrium()
    f0 = []
    for i in range(len(sizes)):
        f = b.get_f(i)
        f = f + (1e-3 * torch.from_numpy(rng.standard_normal(f.numel())).to(f.device, f.dtype))
        b.set_f(i, f)
        f0.append(f.cpu().numpy())
    b.run(9)
    forces = b.force()
    for i in range(len(sizes)):
        ref, fref = _single(b, i, f0[i], taus[i], 9)
        assert np.array_equal(b.get_f(i).cpu().numpy().view(np.uint8), ref.view(np.uint8)), f"case {i} differs"
        assert np.allclose(forces[i], fref, rtol=1e-12, atol=1e-15)


def test_batch_autotune_then_matches(gpu):
    from mojo_lbm_amd import get_D2Q9

    cases = _small_cylinders(4, seed=9)
    b = _batch(cases, get_D2Q9(), _config())
    tile, grid = b.autotune(steps=20)
    info = b.info()
    assert (info["tx"], info["ty"], info["tz"]) == tile and info["grid_blocks"] == grid
    f0 = [b.get_f(i).cpu().numpy() for i in range(len(cases))]
    b.run(11)
    for i, c in enumerate(cases):
        ref, _ = _single(b, i, f0[i], c.tau, 11)
        assert np.array_equal(b.get_f(i).cpu().numpy().view(np.uint8), ref.view(np.uint8))


def test_batch_errors(gpu):
    from mojo_lbm_amd import _lib, get_D2Q9

    cases = _small_cylinders(2)
    b = _batch(cases, get_D2Q9(), _config())
    with pytest.raises(_lib.LbmError):
        b.run(1)  # state not initialised
    with pytest.raises(_lib.LbmError):
        b.set_option(graph_steps=3)  # odd
    with pytest.raises(_lib.LbmError):
        b.set_tile((12, 16, 1))  # not a power of two
    with pytest.raises(IndexError):
        b.set_case(5)
    b.init_equilibrium()
    with pytest.raises(_lib.LbmError):
        b.force()  # no step yet
    b.run(2)
    assert np.all(np.isfinite(b.force()))
