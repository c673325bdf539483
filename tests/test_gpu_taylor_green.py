"""Taylor-Green vortex decay on the GPU: a periodic 2-D vortex set up with initialize_f_from_func
(equilibrium plus the non-equilibrium gradient term) decays as exp(-2 nu k^2 t) with the lattice
viscosity nu = (tau - 1/2) / 3 of SRT, TRT, RLBM and KBC (and in the x-y plane of a D3Q19 box); the
measured decay rate of the velocity amplitude must match to 0.1 % (measured: within 3e-4)."""
import numpy as np
import pytest

import mojo_lbm_amd as lbm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lat,collision,f16c", [("D2Q9", "SRT", False), ("D2Q9", "TRT", False),
                                                 ("D2Q9", "RLBM", False), ("D2Q9", "KBC", False),
                                                 ("D2Q9", "SRT", True),
                                                 ("D3Q19", "SRT", False)])
def test_taylor_green_decay(gpu, lat, collision, f16c):
    print(f"case taylor-green {lat} {collision} f16c={f16c}", flush=True)
    N, u0, tau = 64, 0.02, 0.8
    D3 = lat == "D3Q19"
    grid = lbm.LBM_Grid(lbm.get_D3Q19() if D3 else lbm.get_D2Q9(), N, N, 8 if D3 else 1, 8 if D3 else 16)
    cfg = lbm.LBM_Config(lbm.DOUBLE_BUFFER, collision_op=collision, use_float16c=f16c, DDF_shift=f16c)
    solver = lbm.Solver(grid, cfg)
    asm = lbm.Assembly(grid, cfg)
    k = 2 * np.pi / N

    def u(x, y, z):
        ux = u0 * np.sin(k * x) * np.cos(k * y)
        uy = -u0 * np.cos(k * x) * np.sin(k * y)
        return [ux, uy] + ([np.zeros_like(x)] if D3 else [])

    def du(x, y, z, vel):
        a = u0 * k
        g = [[a * np.cos(k * x) * np.cos(k * y), -a * np.sin(k * x) * np.sin(k * y)],
             [a * np.sin(k * x) * np.sin(k * y), -a * np.cos(k * x) * np.cos(k * y)]]
        if D3:
            z0 = np.zeros_like(x)
            g = [row + [z0] for row in g] + [[z0, z0, z0]]
        return g

    lbm.initialize_f_from_func(grid, cfg, asm.f, u, 1.0, tau, None, du, solver)
    nu = (tau - 0.5) / 3.0
    amps = []
    steps = 1000
    for _ in range(2):
        for _ in range(steps // 2):
            solver.even_step(asm, tau)
            solver.odd_step(asm, tau)
        _, vel = lbm.calculate_rho_and_velocity(solver, asm, after_odd_step=True)
        amps.append(np.sqrt(np.mean(vel[..., 0] ** 2 + vel[..., 1] ** 2)))
    rate = np.log(amps[0] / amps[1]) / steps
    expect = 2 * nu * k * k
    print(f"decay rate {rate:.6e}, 2 nu k^2 = {expect:.6e}, relative error {rate / expect - 1:+.2e}", flush=True)
    assert abs(rate / expect - 1) < 1e-3, (rate, expect)
