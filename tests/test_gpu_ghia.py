"""Lid-driven cavity against the Ghia, Ghia & Shin (1982) centrelines the reference ships
(tests/golden, from benchmarks/LDC/*): D2Q9 256^2 at Re = 400 and 1000 run to a steady state on the
GPU through the host API; u(y) on x = 1/2 and v(x) on y = 1/2 must follow the published values.
The cavity spans the N - 2 fluid nodes between the halfway walls, so Re = u_lid (N - 2) / nu."""
import os

import numpy as np
import pytest

import mojo_lbm_amd as lbm

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(__file__)


def _ghia(re):
    gu = np.loadtxt(os.path.join(HERE, "golden", "ghia_u_centerline.txt"), skiprows=1)
    gv = np.loadtxt(os.path.join(HERE, "golden", "ghia_v_centerline.csv"), skiprows=1, delimiter=",")
    col = [100, 400, 1000, 3200, 5000, 7500, 10000].index(re) + 1
    return gu[:, [0, col]], gv[:, [0, col]]


@pytest.mark.parametrize("re,steps,method,f16c,tol", [
    (400, 60000, "DOUBLE_BUFFER", False, 0.01),
    (1000, 120000, "DOUBLE_BUFFER", False, 0.02),
    (1000, 120000, "ESOTERIC_PULL", True, 0.02),
])
def test_cavity_matches_ghia(gpu, re, steps, method, f16c, tol):
    print(f"case ghia Re={re} {method} f16c={f16c}", flush=True)
    N, u_lid = 256, 0.1
    tau = 0.5 + 3.0 * u_lid * (N - 2) / re
    grid = lbm.LBM_Grid(lbm.get_D2Q9(), N, N, 1, 16)
    cfg = lbm.LBM_Config(getattr(lbm, method), use_float16c=f16c, DDF_shift=f16c)
    solver = lbm.Solver(grid, cfg)
    asm = lbm.Assembly(grid, cfg)
    lbm.initialize_fluid_at_rest(grid, cfg, asm.f, solver)
    if asm.f2 is not None:
        lbm.initialize_fluid_at_rest(grid, cfg, asm.f2, solver)
    lbm.set_exterior_walls(grid, cfg, asm.flags, asm.bc, "+Y", lbm.Flags.SOLID, [u_lid, 0.0], 1.0)
    for side in ("-Y", "+X", "-X"):
        lbm.set_exterior_walls(grid, cfg, asm.flags, asm.bc, side, lbm.Flags.SOLID, [0.0, 0.0], 1.0)
    for _ in range(steps // 2):
        solver.even_step(asm, tau)
        solver.odd_step(asm, tau)
    rho, u = lbm.calculate_rho_and_velocity(solver, asm, after_odd_step=True)
    assert np.all(np.isfinite(u))
    uu = u[:, :, 0, :] / u_lid
    pos = (np.arange(N) - 0.5) / (N - 2)
    gu, gv = _ghia(re)
    ucl = 0.5 * (uu[N // 2 - 1, :, 0] + uu[N // 2, :, 0])  # u(y) at x = 1/2
    vcl = 0.5 * (uu[:, N // 2 - 1, 1] + uu[:, N // 2, 1])  # v(x) at y = 1/2
    if re == 400:
        # Table II of Ghia et al. lists v = -0.23827 at x = 0.9063 for Re = 400, a value off the
        # smooth profile of its neighbours (-0.22847 at 0.9453, -0.44993 at 0.8594): a known
        # misprint, left out here (this run gives -0.388 there).
        gv = gv[np.abs(gv[:, 0] - 0.9063) > 1e-6]
    du = np.max(np.abs(np.interp(gu[1:-1, 0], pos[1:-1], ucl[1:-1]) - gu[1:-1, 1]))
    dv = np.max(np.abs(np.interp(gv[1:-1, 0], pos[1:-1], vcl[1:-1]) - gv[1:-1, 1]))
    print(f"Re={re}: max |u - Ghia| = {du:.4f}, max |v - Ghia| = {dv:.4f}", flush=True)
    assert du < tol and dv < tol
