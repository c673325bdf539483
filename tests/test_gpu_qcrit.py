"""GPU parity of lbm_q_criterion (calculate_Q_criterion, src/lbm/output/Q_criterion.mojo:90-215).

Each case runs a few steps of a lid-driven box with seeded noise on the GPU, takes the
velocity field from lbm_macroscopic, and computes Q on the GPU and in the oracle from the
same populations and velocity field: the results must be bit-identical (both evaluate the
reference's operations in its order with no contraction). The Q = |Omega|^2 closed form of
a solid-body rotation is checked on the GPU too, and the interpolated-link variant against
the oracle's.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U

pytestmark = pytest.mark.gpu

CASES = [("D2Q9", 0, 0, 0), ("D2Q9", 1, 1, 1), ("D2Q9", 2, 0, 1), ("D2Q9", 0, 1, 1),
         ("D3Q19", 0, 0, 0), ("D3Q19", 1, 1, 1), ("D3Q19", 2, 1, 0), ("D3Q19", 0, 0, 1),
         ("D3Q27", 0, 1, 0), ("D3Q27", 1, 0, 1), ("D3Q27", 2, 0, 1)]


def _steps(run, desc, tau, k):
    for i in range(k):
        if desc.method == 0:
            run.step_db(tau)
        else:
            run.step_ep(tau, i % 2 == 0)


@pytest.mark.parametrize("lat,storage,ddf,method", CASES)
@pytest.mark.parametrize("u_layout", [0, 1])
def test_q_criterion_matches_oracle(gpu, lat, storage, ddf, method, u_layout):
    print(f"case qcrit {lat} storage={storage} ddf={ddf} method={method} u_layout={u_layout}", flush=True)
    n = (64, 32, 1) if lat == "D2Q9" else (16, 16, 16)
    desc = U.make_desc(lat, n, storage=storage, ddf=ddf, method=method)
    f, bc, flags = U.perturbed_state(desc, seed=11 * storage + 3 * ddf + method, amp=2e-3)
    tau = 0.58
    run = U.DeviceRun(desc, f, bc, flags)
    k = 5
    _steps(run, desc, tau, k)
    after_odd = (k % 2 == 1)
    _, u = run.macroscopic(after_odd=after_odd, u_layout=u_layout)
    q = run.q_criterion(u, tau, after_odd=after_odd, u_layout=u_layout)
    oq = O.q_criterion(desc, u, run.f_host(), bc, flags, tau, after_odd=after_odd, u_layout=u_layout)
    assert np.all(np.isfinite(q))
    assert np.array_equal(q.view(np.uint8), oq.view(np.uint8)), float(np.max(np.abs(q - oq)))
    assert np.max(np.abs(q)) > 0


@pytest.mark.parametrize("lat", ["D2Q9", "D3Q19"])
def test_rotation_closed_form(gpu, lat):
    n = (32, 32, 1) if lat == "D2Q9" else (16, 16, 16)
    desc = U.make_desc(lat, n, storage=2)
    f, bc, flags = O.alloc(desc)
    O.init_equilibrium(desc, f)
    omega = np.array([0.0, 0.0, 1e-3]) if desc.D == 2 else np.array([3e-4, -7e-4, 5e-4])
    x, y, z = np.meshgrid(np.arange(desc.nx), np.arange(desc.ny), np.arange(desc.nz), indexing="ij")
    r = np.stack([x - 10.3, y - 12.6, z - 7.2], axis=-1).astype(np.float64)
    u = np.cross(np.broadcast_to(omega, r.shape), r)[..., : desc.D].reshape(-1)
    run = U.DeviceRun(desc, f, bc, flags)
    q = run.q_criterion(u, 0.7).reshape(desc.nx, desc.ny, desc.nz)
    inner = tuple(slice(1, s - 1) if s > 1 else slice(None) for s in (desc.nx, desc.ny, desc.nz))
    np.testing.assert_allclose(q[inner], float(omega @ omega), rtol=1e-9)


@pytest.mark.parametrize("lat,storage", [("D2Q9", 0), ("D3Q19", 1)])
def test_q_criterion_with_links(gpu, lat, storage):
    desc, f, bc, flags, nodes, qdist = U.obstacle_case(lat, storage=storage, ddf=1, seed=5)
    tau = 0.6
    run = U.DeviceRun(desc, f, bc, flags)
    run.set_links(nodes, qdist)
    for _ in range(3):
        run.step_db(tau)
    _, u = run.macroscopic()
    q = run.q_criterion(u, tau)
    oq = O.q_criterion(desc, u, run.f_host(), bc, flags, tau, links=(nodes, qdist))
    assert np.array_equal(q.view(np.uint8), oq.view(np.uint8)), float(np.max(np.abs(q - oq)))


def test_q_criterion_errors(gpu):
    from mojo_lbm_amd import _lib

    desc = U.make_desc("D2Q9", (32, 32, 1))
    f, bc, flags = U.perturbed_state(desc)
    run = U.DeviceRun(desc, f, bc, flags)
    u = np.zeros(32 * 32 * 2, np.float32)
    with pytest.raises(RuntimeError, match="u_layout"):
        run.q_criterion(u, 0.6, u_layout=7)
    with pytest.raises(RuntimeError, match="tau"):
        run.q_criterion(u, 0.0)
    assert _lib.lib().lbm_q_criterion(run.h.ptr, None, None, None, None, None, 0.6, 0, 0, None) != 0


@pytest.mark.parametrize("method_name", ["DOUBLE_BUFFER", "ESOTERIC_PULL"])
def test_output_request(gpu, method_name):
    """OutputRequest through the host API on a D3Q19 lid-driven cube: the numpy fields match
    calculate_rho_and_velocity and the oracle's Q of the same state."""
    import mojo_lbm_amd as lbm
    from mojo_lbm_amd.solver import make_desc

    grid = lbm.LBM_Grid(lbm.get_D3Q19(), 16, 16, 16, 8)
    config = lbm.LBM_Config(getattr(lbm, method_name))
    solver = lbm.Solver(grid, config)
    asm = lbm.Assembly(grid, config)
    lbm.initialize_fluid_at_rest(grid, config, asm.f, solver)
    if asm.f2 is not None:
        lbm.initialize_fluid_at_rest(grid, config, asm.f2, solver)
    lbm.set_exterior_walls(grid, config, asm.flags, asm.bc, "+Y", lbm.Flags.SOLID, [0.1, 0.0, 0.0], 1.0)
    for side in ("-Y", "+X", "-X", "-Z", "+Z"):
        lbm.set_exterior_walls(grid, config, asm.flags, asm.bc, side, lbm.Flags.SOLID, [0.0] * 3, 1.0)
    tau = 0.56
    for k in range(6):
        (solver.even_step if k % 2 == 0 else solver.odd_step)(asm, tau)
    units = lbm.UnitSystem(1.0, 0.1, 2.0, 16.0, 1000.0, 1e-3)
    out = lbm.OutputRequest(solver, units, Q_criterion=True)
    out.velocity_frame(asm, True)
    out.Q_criterion_frame(asm, True, tau)
    rho, u = lbm.calculate_rho_and_velocity(solver, asm, True)
    assert np.array_equal(out.velocity_as_numpy(False), u)
    assert np.array_equal(out.density_as_numpy(), rho)
    np.testing.assert_allclose(out.velocity_as_numpy(True), u * (1.0 / 0.1), rtol=1e-6)
    q = out.Q_criterion_as_numpy(False)
    desc = make_desc(grid, config)
    oq = O.q_criterion(desc, u.reshape(-1), asm.f.cpu(), asm.bc.cpu(), asm.flags.cpu(), tau, after_odd=True)
    assert np.array_equal(q.reshape(-1), oq)
    assert np.array_equal(lbm.calculate_Q_criterion(solver, asm, u, tau, True), q)
    with pytest.raises(ValueError):
        lbm.OutputRequest(solver, velocity=False, Q_criterion=True)


COLUMN_CASES = [("D3Q19", 0, 0, 0, (32, 32, 32), False), ("D3Q19", 1, 1, 1, (32, 16, 64), False),
                ("D3Q27", 2, 0, 0, (16, 16, 32), False), ("D3Q19", 0, 0, 1, (16, 24, 32), True),
                ("D3Q27", 1, 1, 0, (16, 16, 32), True)]


@pytest.mark.parametrize("lat,storage,ddf,method,n,walls", COLUMN_CASES)
@pytest.mark.parametrize("u_layout", [0, 1])
def test_q_criterion_tall_grids(gpu, lat, storage, ddf, method, n, walls, u_layout):
    """Q on grids of a multiple of 4 tiles in z (the four-tile column kernel; an output off a
    16-byte boundary takes the tile kernel): bit-identical to the oracle in periodic boxes (with the
    domain-edge one-sided differences) and in the walled box, both u layouts."""
    print(f"case qcrit tall grid {lat} storage={storage} ddf={ddf} method={method} n={n} walls={walls} "
          f"u_layout={u_layout}", flush=True)
    desc = U.make_desc(lat, n, storage=storage, ddf=ddf, method=method)
    f, bc, flags = U.perturbed_state(desc, seed=5 * storage + ddf + 2 * method, amp=2e-3, walls=walls)
    tau = 0.58
    run = U.DeviceRun(desc, f, bc, flags)
    k = 3
    _steps(run, desc, tau, k)
    after_odd = (k % 2 == 1)
    _, u = run.macroscopic(after_odd=after_odd, u_layout=u_layout)
    oq = O.q_criterion(desc, u, run.f_host(), bc, flags, tau, after_odd=after_odd, u_layout=u_layout)
    for misalign in (False, True):
        q = run.q_criterion(u, tau, after_odd=after_odd, u_layout=u_layout, misalign=misalign)
        assert np.array_equal(q.view(np.uint8), oq.view(np.uint8)), (misalign, float(np.max(np.abs(q - oq))))
    assert np.max(np.abs(q)) > 0
