"""GPU parity of lbm_init_f_from_fields (initialize_f_from_func, src/lbm/preprocess/initial_condition.mojo:86-255)
and the host mirror initialize_f_from_func: bit-identical to the oracle on the same fields for
every lattice, storage, method and DDF setting, with and without the gradient term, whether
the grid is filled in one call or in x-slabs."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import _util as U

pytestmark = pytest.mark.gpu

CASES = [(lat, s, m, d) for lat in ("D2Q9", "D3Q19", "D3Q27") for s in (0, 1, 2) for m in (0, 1) for d in (0, 1)
         if (s + m + d + len(lat)) % 2 == 0 or lat == "D3Q19"]


def _fields(desc, seed):
    rng = np.random.default_rng(seed)
    n = desc.nx * desc.ny * desc.nz
    dt = O.r_dtype(desc)
    u = (0.05 * rng.standard_normal((n, desc.D))).astype(dt)
    g = (2e-3 * rng.standard_normal((n, desc.D, desc.D))).astype(dt)
    return u, g


@pytest.mark.parametrize("lat,storage,method,ddf", CASES)
@pytest.mark.parametrize("with_grad", [False, True])
def test_init_from_fields_matches_oracle(gpu, lat, storage, method, ddf, with_grad):
    from mojo_lbm_amd import _lib

    print(f"case init {lat} storage={storage} method={method} ddf={ddf} grad={with_grad}", flush=True)
    n = (32, 32, 1) if lat == "D2Q9" else (16, 8, 16)
    desc = U.make_desc(lat, n, storage=storage, method=method, ddf=ddf)
    u, g = _fields(desc, 7 + storage)
    tau = 0.73
    of, _, _ = O.alloc(desc)
    O.init_f_from_fields(desc, of, 1.02, u, g if with_grad else None, tau)
    h = _lib.Handle(desc)
    L = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream
    fd = torch.zeros(of.nbytes, dtype=torch.uint8, device="cuda")
    ud = torch.from_numpy(u).cuda()
    gd = torch.from_numpy(g).cuda()
    _lib.check(L.lbm_init_f_from_fields(h.ptr, fd.data_ptr(), 1.02, ud.data_ptr(),
                                        gd.data_ptr() if with_grad else None, tau, 0, desc.nx, st))
    torch.cuda.synchronize()
    assert np.array_equal(fd.cpu().numpy(), of.view(np.uint8))
    # the same fill in three x-slabs
    fd2 = torch.zeros_like(fd)
    per = desc.ny * desc.nz
    for x0, nxc in ((0, 5), (5, 7), (12, desc.nx - 12)):
        _lib.check(L.lbm_init_f_from_fields(h.ptr, fd2.data_ptr(), 1.02, ud[x0 * per:].data_ptr(),
                                            gd[x0 * per:].data_ptr() if with_grad else None, tau, x0, nxc, st))
    torch.cuda.synchronize()
    assert torch.equal(fd, fd2)


def test_init_from_fields_errors(gpu):
    from mojo_lbm_amd import _lib

    desc = U.make_desc("D2Q9", (32, 32, 1))
    h = _lib.Handle(desc)
    L = _lib.lib()
    buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    assert L.lbm_init_f_from_fields(h.ptr, buf.data_ptr(), 1.0, buf.data_ptr(), None, 1.0, 30, 3, None) != 0
    assert "x-slab" in L.lbm_last_error().decode()
    assert L.lbm_init_f_from_fields(h.ptr, buf.data_ptr(), 1.0, buf.data_ptr(), buf.data_ptr(), 0.0, 0, 1, None) != 0


@pytest.mark.parametrize("method_name", ["DOUBLE_BUFFER", "ESOTERIC_PULL"])
def test_initialize_f_from_func_taylor_green(gpu, method_name):
    """A 2-D Taylor-Green vortex in physical units through the host API: the device fill equals the
    oracle's on the fields the functions give, and the macroscopic velocity returns u(x)."""
    import mojo_lbm_amd as lbm
    from mojo_lbm_amd.solver import make_desc

    grid = lbm.LBM_Grid(lbm.get_D2Q9(), 64, 64, 1, 16, dx=0.25, origin=(-1.0, 0.5, 0.0))
    config = lbm.LBM_Config(getattr(lbm, method_name), use_float16c=True, DDF_shift=True)
    solver = lbm.Solver(grid, config)
    asm = lbm.Assembly(grid, config)
    units = lbm.UnitSystem(2.0, 0.05, 16.0, 64.0, 1000.0, 1e-3, D=2)
    k = 2 * np.pi / 16.0

    def u(x, y, z):
        return [2.0 * np.sin(k * x) * np.cos(k * y), -2.0 * np.cos(k * x) * np.sin(k * y)]

    def du(x, y, z, vel):
        a = 2.0 * k
        return [[a * np.cos(k * x) * np.cos(k * y), -a * np.sin(k * x) * np.sin(k * y)],
                [a * np.sin(k * x) * np.sin(k * y), -a * np.cos(k * x) * np.cos(k * y)]]

    tau = 0.8
    lbm.initialize_f_from_func(grid, config, asm.f, u, 1.0, tau, units, du, solver, slab_nodes=700)
    got = asm.f.cpu().copy()
    # the same fields on the host, in the grid float type, through the oracle
    dt = np.float32
    xs, ys, zs = np.meshgrid(np.arange(64), np.arange(64), np.arange(1), indexing="ij")
    X = xs.astype(dt) * dt(0.25) + dt(-1.0)
    Y = ys.astype(dt) * dt(0.25) + dt(0.5)
    Z = zs.astype(dt) * dt(0.25) + dt(0.0)
    us = dt(dt(0.05) / dt(2.0))
    gs = dt(us / dt(dt(64.0) / dt(16.0)))
    vel = np.stack([np.asarray(c, dt) for c in u(X, Y, Z)], axis=-1) * us
    grad = np.stack([np.stack([np.asarray(c, dt) for c in row], axis=-1) for row in du(X, Y, Z, vel)], axis=-2) * gs
    desc = make_desc(grid, config)
    of, _, _ = O.alloc(desc)
    O.init_f_from_fields(desc, of, 1.0, vel.reshape(-1), grad.reshape(-1), tau)
    assert np.array_equal(got, of)
    # periodic box (one wavelength each way), no walls: the re-streamed velocity of the stored
    # state is u(x) in lattice units up to the streaming error O(|u|^2 k) ~ 3e-4
    _, uu = lbm.calculate_rho_and_velocity(solver, asm, after_odd_step=True)
    np.testing.assert_allclose(uu.reshape(vel.shape), vel, atol=1e-3)
