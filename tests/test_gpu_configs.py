"""The benchmark configurations run through the host API mirror (LBM_Config,
LBM_Grid, Assembly, Solver.even_step / odd_step, set_exterior_walls,
initialize_fluid_at_rest, calculate_rho_and_velocity) on the GPU, compared with
the oracle on the same setup.

BL-cfg1: D2Q9 SRT lid-driven cavity 128^2, FP32 pull, u_lid = 0.1, Re = 100,
10,000 steps: rho, u, v equal to the oracle's, and centrelines near Ghia et al.
BL-cfg2: D3Q19 SRT lid-driven cube, FP32 two-buffer pull, 8x8x8 tiles: the full
1,000 steps at 64^3 and the first steps at the full 256^3; velocity relative L2
against the oracle (contract 1e-5; the runs are bit-identical).
BL-cfg3 storage: Float16C + DDF shift, double buffer and single-buffer esoteric pull.
"""
import os

import numpy as np
import pytest

import mojo_lbm_amd as lbm
from mojo_lbm_amd.solver import make_desc
from oracle import oracle as O
from tests import _util as U

pytestmark = pytest.mark.gpu


def _setup(lattice, n, tile, config, u_lid):
    grid = lbm.LBM_Grid(lattice, *n, tile)
    solver = lbm.Solver(grid, config)
    asm = lbm.Assembly(grid, config)
    lbm.initialize_fluid_at_rest(grid, config, asm.f, solver)
    if asm.f2 is not None:
        lbm.initialize_fluid_at_rest(grid, config, asm.f2, solver)
    D = grid.D
    lbm.set_exterior_walls(grid, config, asm.flags, asm.bc, "+Y", lbm.Flags.SOLID, [u_lid] + [0.0] * (D - 1), 1.0)
    for side in ["-Y", "+X", "-X"] + (["-Z", "+Z"] if D == 3 else []):
        lbm.set_exterior_walls(grid, config, asm.flags, asm.bc, side, lbm.Flags.SOLID, [0.0] * D, 1.0)
    desc = make_desc(grid, config)
    # the same setup on the oracle side
    f, bc, flags = O.alloc(desc)
    O.init_equilibrium(desc, f)
    if config.lbm_method == lbm.ESOTERIC_PULL:
        f = U.db_to_ep_odd(desc, f)  # uniform state: the odd store pattern leaves it unchanged
    U.ldc_walls(desc, flags, bc, u_lid)
    assert np.array_equal(asm.flags.cpu(), flags)
    assert np.array_equal(asm.bc.cpu().view(np.uint8), bc.view(np.uint8))
    assert np.array_equal(asm.f.cpu().view(np.uint8), f.view(np.uint8))
    return grid, solver, asm, desc, f, bc, flags


def _run_pairs(solver, asm, tau, pairs):
    for _ in range(pairs):
        solver.even_step(asm, tau)
        solver.odd_step(asm, tau)


def _rel_l2(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b.astype(np.float64)), 1e-300))


def test_cfg1_d2q9_cavity_128_re100(gpu):
    N, u_lid, Re = 128, 0.1, 100.0
    tau = lbm.tau_from_re(u_lid, N, Re)
    cfg = lbm.DoubleBufferConfig()
    grid, solver, asm, desc, f, bc, flags = _setup(lbm.get_D2Q9(), (N, N, 1), 16, cfg, u_lid)
    print("cfg1: D2Q9 128^2, 10000 steps", flush=True)
    _run_pairs(solver, asm, tau, 5000)
    rho, u = lbm.calculate_rho_and_velocity(solver, asm, after_odd_step=True)
    ref_f = U.oracle_run_db(desc, f, bc, flags, tau, 10000)
    rr, ur = O.macroscopic(desc, ref_f, bc, flags)
    np.testing.assert_array_equal(rho.ravel(), rr)
    np.testing.assert_array_equal(u.ravel(), ur)
    here = os.path.dirname(__file__)
    gu = np.loadtxt(os.path.join(here, "golden", "ghia_u_centerline.txt"), skiprows=1)
    uu = u[:, :, 0, :] / u_lid
    pos = (np.arange(N) - 0.5) / (N - 2)
    uc = 0.5 * (uu[N // 2 - 1, :, 0] + uu[N // 2, :, 0])
    iu = np.interp(gu[1:-1, 0], pos[1:-1], uc[1:-1])
    assert np.max(np.abs(iu - gu[1:-1, 1])) < 0.05


@pytest.mark.parametrize("n,pairs", [(64, 500), (256, 1)])
def test_cfg2_d3q19_cube_fp32(gpu, n, pairs):
    u_lid, Re = 0.1, 100.0
    tau = lbm.tau_from_re(u_lid, n, Re)
    cfg = lbm.DoubleBufferConfig()
    grid, solver, asm, desc, f, bc, flags = _setup(lbm.get_D3Q19(), (n, n, n), 8, cfg, u_lid)
    print(f"cfg2: D3Q19 {n}^3, {2 * pairs} steps", flush=True)
    _run_pairs(solver, asm, tau, pairs)
    rho, u = lbm.calculate_rho_and_velocity(solver, asm, after_odd_step=True)
    ref_f = U.oracle_run_db(desc, f, bc, flags, tau, 2 * pairs)
    rr, ur = O.macroscopic(desc, ref_f, bc, flags)
    assert _rel_l2(u.ravel(), ur) <= 1e-5
    assert np.max(np.abs(asm.f.cpu() - ref_f)) <= 1e-6
    np.testing.assert_array_equal(u.ravel(), ur)  # bit-identical


@pytest.mark.parametrize("method", [lbm.DOUBLE_BUFFER, lbm.ESOTERIC_PULL])
def test_cfg3_f16c_ddf_cube(gpu, method):
    n, u_lid, Re = 64, 0.1, 1000.0
    tau = lbm.tau_from_re(u_lid, n, Re)
    cfg = lbm.LBM_Config(method, use_float16c=True, DDF_shift=True)
    grid, solver, asm, desc, f, bc, flags = _setup(lbm.get_D3Q19(), (n, n, n), 8, cfg, u_lid)
    print(f"cfg3: D3Q19 {n}^3 Float16C method={method}, 200 steps", flush=True)
    _run_pairs(solver, asm, tau, 100)
    got = asm.f.cpu()
    if method == lbm.DOUBLE_BUFFER:
        ref = U.oracle_run_db(desc, f, bc, flags, tau, 200)
    else:
        ref = U.oracle_run_ep(desc, f, bc, flags, tau, 200)
    np.testing.assert_array_equal(got, ref)
    rho, u = lbm.calculate_rho_and_velocity(solver, asm, after_odd_step=True)
    rr, ur = O.macroscopic(desc, ref, bc, flags, after_odd=True)
    np.testing.assert_array_equal(u.ravel(), ur)
    tot = lbm.reduce_moments(solver, asm, after_odd_step=True)
    # FP64 totals: the GPU sums in a tree, the oracle sequentially; near-zero components need an absolute bound
    np.testing.assert_allclose(tot, O.reduce_moments(desc, ref, bc, flags, after_odd=True), rtol=1e-12, atol=1e-10)


def test_dfg_re20_cylinder_drag(gpu):
    """Schaefer-Turek (DFG) 2-D cylinder at Re = 20, the reference's cylinder benchmark set-up
    (benchmarks/Cylinder2D/CylinderFlow_Re20.mojo:96-150: N = ny nodes over H = 0.41, nx = 5 N,
    centre node floor(0.2 / dx), parabolic inlet, EQUILIBRIUM outlet), here at N = 128 with
    interpolated bounce-back on the cylinder. Published Cd = 5.57953523384, Cl = 0.010618948146
    (:59-60); the reference reports 1.2 % / 0.17 % / 0.35 % drag error at N = 256 / 512 / 1024
    (:68-91). The drag must be within 1 % and steady; the lift positive and of the published size."""
    from mojo_lbm_amd.batch import CaseBatch
    from mojo_lbm_amd.cylinder import CylinderCase

    lat = lbm.get_D2Q9()
    N = 128
    dx = 0.41 / (N - 1)
    cen = float(np.floor(0.2 / dx))
    case = CylinderCase(nx=5 * N, ny=N, diameter=0.1 / dx, cx=cen, cy=cen, re=20.0, u_max=0.06)
    b = CaseBatch(lat, [(case.nx, case.ny, 1)], [case.tau], lbm.DoubleBufferConfig(BCs={lbm.Flags.EQUILIBRIUM}))
    flags, bc, links = case.arrays(lat)
    b.set_case(0, flags, bc, links)
    b.init_equilibrium()
    b.run(50000)
    cd0, _ = case.coefficients(b.force()[0])
    b.run(10000)
    cd, cl = case.coefficients(b.force()[0])
    assert abs(cd - cd0) < 1e-4 * cd
    assert abs(cd - 5.57953523384) < 0.01 * 5.57953523384, cd
    assert 0.0 < cl < 0.05, cl
