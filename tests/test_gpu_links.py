"""GPU parity of interpolated bounce-back links and the momentum-exchange force.

A channel with an immersed circle / sphere (the reference cylinder script's boundary
order) runs through lbm_set_links + lbm_step_db and through the oracle; populations
must be bit-identical after every step, the force of the last step must agree to
rtol 1e-12 (FP64 sums in different orders), and the macroscopic fields / totals
(which re-apply the links) must agree as in the plain parity tests.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U

pytestmark = pytest.mark.gpu

CASES = [("D2Q9", 0, 0), ("D2Q9", 0, 1), ("D2Q9", 1, 1), ("D2Q9", 2, 0),
         ("D3Q19", 0, 0), ("D3Q19", 1, 1), ("D3Q27", 0, 1), ("D3Q27", 2, 1)]


@pytest.mark.parametrize("lat,storage,ddf", CASES)
def test_links_match_oracle(gpu, lat, storage, ddf):
    print(f"case links {lat} storage={storage} ddf={ddf}", flush=True)
    desc, f, bc, flags, nodes, qdist = U.obstacle_case(lat, storage=storage, ddf=ddf, seed=storage * 7 + ddf)
    tau = 0.6
    run = U.DeviceRun(desc, f, bc, flags)
    run.set_links(nodes, qdist)
    a, b = f.copy(), f.copy()
    for step in range(4):
        run.step_db(tau)
        force = O.step_db_links(desc, b, a, bc, flags, tau, nodes, qdist)
        a, b = b, a
        got = run.f_host()
        assert np.array_equal(got.view(np.uint8), a.view(np.uint8)), f"populations differ after step {step + 1}"
    gforce = run.link_force()
    scale = np.max(np.abs(force)) + 1e-30
    assert np.all(np.abs(gforce - force) <= 1e-12 * scale), (gforce, force)
    rho, u = run.macroscopic()
    orho, ou = O.macroscopic(desc, a, bc, flags, links=(nodes, qdist))
    assert np.max(np.abs(rho - orho)) <= 1e-10 and np.max(np.abs(u - ou)) <= 1e-10
    tot = run.reduce()
    otot = O.reduce_moments(desc, a, bc, flags, links=(nodes, qdist))
    assert np.all(np.abs(tot - otot) <= 1e-10 * np.maximum(1.0, np.abs(otot)))


def test_links_kbc_les(gpu):
    desc, f, bc, flags, nodes, qdist = U.obstacle_case("D3Q19", storage=0, ddf=1, seed=11, collision=2, les=1)
    run = U.DeviceRun(desc, f, bc, flags)
    run.set_links(nodes, qdist)
    a, b = f.copy(), f.copy()
    for _ in range(3):
        run.step_db(0.52)
        force = O.step_db_links(desc, b, a, bc, flags, 0.52, nodes, qdist)
        a, b = b, a
    assert np.array_equal(run.f_host().view(np.uint8), a.view(np.uint8))
    assert np.allclose(run.link_force(), force, rtol=1e-12, atol=1e-15)


def test_links_clear_and_errors(gpu):
    from mojo_lbm_amd import _lib

    desc, f, bc, flags, nodes, qdist = U.obstacle_case("D2Q9")
    run = U.DeviceRun(desc, f, bc, flags)
    with pytest.raises(_lib.LbmError):
        run.link_force()  # no step with links yet
    with pytest.raises(_lib.LbmError):
        run.set_links(nodes[::-1].copy(), qdist)  # unsorted
    run.set_links(nodes, qdist)
    run.step_db(0.6)
    run.link_force()
    # clearing the links gives the plain step back
    _lib.check(run.L.lbm_set_links(run.h.ptr, None, None, 0))
    run2 = U.DeviceRun(desc, run.f_host(), bc, flags)
    run.step_db(0.6)
    run2.step_db(0.6)
    assert np.array_equal(run.f_host(), run2.f_host())


def _porous_case(lat, storage, ddf, seed):
    """Periodic box with about a quarter of the nodes SOLID: every fluid node next to a SOLID node is
    FLUID_SOLID_BOUNDARY with a link record whose fractions are random in [0, 1] (a few -1), so
    one-node gaps, corners and opposite links on one node all occur."""
    D, Q = U.LATTICES[lat]
    n = (32, 16, 1) if D == 2 else (16, 16, 8)
    desc = U.make_desc(lat, n, storage=storage, ddf=ddf)
    f, bc, flags = U.perturbed_state(desc, seed=seed, amp=2e-3, walls=False)
    rng = np.random.default_rng(seed)
    c, _, _ = O.lattice(D, Q)
    solid = rng.random(n) < 0.25
    ids = U.index_map(desc, 1)[..., 0]
    flags[ids[solid]] = 1
    bc.reshape(-1, D + 1)[:, D] = 1.0
    nodes, qd = [], []
    for x, y, z in zip(*np.nonzero(~solid)):
        row = np.full(Q - 1, -1.0, np.float32)
        for q in range(1, Q):
            p = ((x - c[q][0]) % n[0], (y - c[q][1]) % n[1], (z - c[q][2]) % n[2] if D == 3 else 0)
            if solid[p] and rng.random() < 0.9:
                row[q - 1] = rng.random()
        if (row >= 0).any():
            nodes.append(ids[x, y, z])
            qd.append(row)
    order = np.argsort(nodes)
    nodes = np.array(nodes, np.uint32)[order]
    qd = np.array(qd, np.float32)[order]
    flags[nodes] = 3
    return desc, f, bc, flags, nodes, qd


@pytest.mark.parametrize("q_clamp", [None, 0.3])
@pytest.mark.parametrize("lat,storage,ddf", [("D2Q9", 0, 0), ("D2Q9", 2, 1), ("D3Q19", 0, 1), ("D3Q27", 1, 1)])
def test_links_porous_gaps_match_oracle(gpu, lat, storage, ddf, q_clamp):
    print(f"case porous links {lat} storage={storage} ddf={ddf} q_clamp={q_clamp}", flush=True)
    desc, f, bc, flags, nodes, qdist = _porous_case(lat, storage, ddf, seed=3 + storage)
    run = U.DeviceRun(desc, f, bc, flags)
    run.set_links(nodes, qdist, q_clamp=q_clamp)
    O.set_q_clamp(q_clamp or 0.0)
    try:
        a, b = f.copy(), f.copy()
        for step in range(3):
            run.step_db(0.7)
            force = O.step_db_links(desc, b, a, bc, flags, 0.7, nodes, qdist)
            a, b = b, a
            assert np.array_equal(run.f_host().view(np.uint8), a.view(np.uint8)), f"differs after step {step + 1}"
    finally:
        O.set_q_clamp(0.0)
    assert np.allclose(run.link_force(), force, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("lat,storage,ddf", [("D2Q9", 0, 0), ("D3Q19", 1, 1), ("D3Q27", 2, 0)])
def test_node_forces_match_oracle(gpu, lat, storage, ddf):
    """lbm_set_link_node_force: the force of every link record's node after each step equals the
    oracle's (the same FP64 sum in the same order), and the records sum to lbm_link_force."""
    import torch

    from mojo_lbm_amd import _lib

    print(f"case node forces {lat} storage={storage} ddf={ddf}", flush=True)
    desc, f, bc, flags, nodes, qdist = _porous_case(lat, storage, ddf, seed=11)
    D = desc.D
    run = U.DeviceRun(desc, f, bc, flags)
    run.set_links(nodes, qdist)
    nf = torch.full((len(nodes) * D,), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(run.L.lbm_set_link_node_force(run.h.ptr, nf.data_ptr()))
    a, b = f.copy(), f.copy()
    onf = np.zeros(len(nodes) * D)
    for step in range(2):
        run.step_db(0.7)
        O.step_db_links(desc, b, a, bc, flags, 0.7, nodes, qdist, node_force=onf)
        a, b = b, a
        torch.cuda.synchronize()
        got = nf.cpu().numpy()
        assert np.array_equal(got, onf), f"node forces differ after step {step + 1}"
    assert np.allclose(got.reshape(-1, D).sum(axis=0), run.link_force()[:D], rtol=1e-12, atol=1e-12)
    # new links clear the output; stopping it leaves the buffer alone
    run.set_links(nodes, qdist)
    nf.fill_(7.0)
    run.step_db(0.7)
    torch.cuda.synchronize()
    assert torch.all(nf == 7.0)
    _lib.check(run.L.lbm_set_links(run.h.ptr, None, None, 0))
    with pytest.raises(_lib.LbmError):
        _lib.check(run.L.lbm_set_link_node_force(run.h.ptr, nf.data_ptr()))
