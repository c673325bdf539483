"""Full-size runs checked through size-independent properties (the oracle is too slow at these
sizes): the FP64 256^3 lid-driven cube (64-bit addressing: 2.5 GB per buffer) conserves mass to
rounding under halfway bounce-back, the Float16C + DDF esoteric 256^3 cube and the 512^3 Float16C
cube of the closure's training runs (5.1 GB, 64-bit addressing) stay finite and conserve mass to
the codec's rounding (and there the column output kernels equal the tile kernel bit for bit), and
the FP32 and FP64 256^3 runs agree on the velocity field."""
import numpy as np
import pytest
import torch

from mojo_lbm_amd import _lib
from tests import _util as U

pytestmark = pytest.mark.gpu


def _cube(n, storage, method, ddf):
    desc = U.make_desc("D3Q19", (n, n, n), storage=storage, method=method, ddf=ddf)
    L = _lib.lib()
    h = _lib.Handle(desc)
    fb, bb, gb = (L.lbm_f_bytes(U.C.byref(desc)), L.lbm_bc_bytes(U.C.byref(desc)), L.lbm_flags_bytes(U.C.byref(desc)))
    f = torch.empty(fb, dtype=torch.uint8, device="cuda")
    f2 = torch.empty(fb, dtype=torch.uint8, device="cuda") if method == 0 else None
    bc = torch.zeros(bb, dtype=torch.uint8, device="cuda")
    flags = torch.zeros(gb, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    for buf in (f, f2):
        if buf is not None:
            _lib.check(L.lbm_init_equilibrium(h.ptr, buf.data_ptr(), 1.0, None, st))
    _lib.check(L.lbm_set_exterior_wall(h.ptr, flags.data_ptr(), bc.data_ptr(), b"+Y", 1, _lib.dvec([0.1, 0, 0]), 1.0, st))
    for side in (b"-Y", b"+X", b"-X", b"-Z", b"+Z"):
        _lib.check(L.lbm_set_exterior_wall(h.ptr, flags.data_ptr(), bc.data_ptr(), side, 1, _lib.dvec([0, 0, 0]), 1.0, st))
    return dict(desc=desc, L=L, h=h, f=f, f2=f2, bc=bc, flags=flags, st=st)


def _steps(c, tau, k):
    L, h, st = c["L"], c["h"], c["st"]
    for s in range(k):
        if c["desc"].method == 0:
            _lib.check(L.lbm_step_db(h.ptr, c["f2"].data_ptr(), c["f"].data_ptr(), c["bc"].data_ptr(),
                                     c["flags"].data_ptr(), tau, st))
            c["f"], c["f2"] = c["f2"], c["f"]
        else:
            _lib.check(L.lbm_step_ep(h.ptr, c["f"].data_ptr(), c["bc"].data_ptr(), c["flags"].data_ptr(), tau,
                                     s % 2 == 0, st))
    c["after_odd"] = k > 0 and (k - 1) % 2 == 1


def _totals(c):
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    _lib.check(c["L"].lbm_reduce_moments_f64(c["h"].ptr, out.data_ptr(), c["f"].data_ptr(), c["bc"].data_ptr(),
                                             c["flags"].data_ptr(), int(c.get("after_odd", True)), c["st"]))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _velocity(c):
    n = c["desc"].nx
    dt = torch.float64 if c["desc"].storage == 2 else torch.float32
    rho = torch.empty(n ** 3, dtype=dt, device="cuda")
    u = torch.empty(n ** 3 * 3, dtype=dt, device="cuda")
    _lib.check(c["L"].lbm_macroscopic(c["h"].ptr, rho.data_ptr(), u.data_ptr(), c["f"].data_ptr(), c["bc"].data_ptr(),
                                      c["flags"].data_ptr(), int(c["after_odd"]), _lib.U_XYZD, c["st"]))
    torch.cuda.synchronize()
    return u.double().cpu().numpy()


def test_fp64_256_conserves_mass_and_matches_fp32(gpu):
    print("case full-size FP64 / FP32 256^3 DB", flush=True)
    tau = 0.5 + 3 * 0.1 * 256 / 100
    c64 = _cube(256, 2, 0, 0)
    m0 = _totals(c64)[0]
    _steps(c64, tau, 60)
    m1 = _totals(c64)
    assert np.all(np.isfinite(m1)) and abs(m1[0] - m0) <= 1e-12 * m0
    u64 = _velocity(c64)
    del c64
    torch.cuda.empty_cache()
    c32 = _cube(256, 0, 0, 0)
    _steps(c32, tau, 60)
    u32 = _velocity(c32)
    del c32
    torch.cuda.empty_cache()
    assert np.max(np.abs(u64)) > 1e-2
    assert np.linalg.norm(u32 - u64) / np.linalg.norm(u64) < 1e-4


def test_f16c_esoteric_256_conserves_mass(gpu):
    print("case full-size F16C + DDF EP 256^3", flush=True)
    c = _cube(256, 1, 1, 1)
    c["after_odd"] = True
    m0 = _totals(c)[0]
    _steps(c, 0.5 + 3 * 0.1 * 256 / 1000, 100)
    m1 = _totals(c)
    del c
    torch.cuda.empty_cache()
    assert np.all(np.isfinite(m1)) and abs(m1[0] - m0) <= 5e-6 * m0


def test_f16c_512_cube_addr64(gpu):
    print("case full-size F16C + DDF DB 512^3 (64-bit addressing)", flush=True)
    free, _ = torch.cuda.mem_get_info()
    if free < 24 * 2 ** 30:
        pytest.skip("needs 24 GiB of free device memory")
    c = _cube(512, 1, 0, 1)
    m0 = _totals(c)[0]
    _steps(c, 0.50768, 20)
    m1 = _totals(c)
    _outputs_match_tile_kernel(c, 0.50768)
    del c
    torch.cuda.empty_cache()
    assert np.all(np.isfinite(m1)) and abs(m1[0] - m0) <= 1e-6 * m0


def _outputs_match_tile_kernel(c, tau):
    """The column output kernels (64-bit addressing here) against the tile kernel, which an output
    one element past a 16-byte boundary selects: rho, u (both layouts) and Q bit for bit."""
    L, h, st, n = c["L"], c["h"], c["st"], c["desc"].nx ** 3
    args = (c["f"].data_ptr(), c["bc"].data_ptr(), c["flags"].data_ptr())
    after_odd = int(c["after_odd"])

    def bits(t):
        return t.view(torch.int32)

    for layout in (_lib.U_XYZD, _lib.U_DXYZ):
        res = []
        for k in (0, 1):  # aligned (columns), then misaligned (tiles)
            rho = torch.empty(n + k, dtype=torch.float32, device="cuda")[k:]
            u = torch.empty(3 * n + k, dtype=torch.float32, device="cuda")[k:]
            q = torch.empty(n + k, dtype=torch.float32, device="cuda")[k:]
            _lib.check(L.lbm_macroscopic(h.ptr, rho.data_ptr(), u.data_ptr(), *args, after_odd, layout, st))
            _lib.check(L.lbm_q_criterion(h.ptr, q.data_ptr(), u.data_ptr(), *args, tau, after_odd, layout, st))
            torch.cuda.synchronize()
            res.append((rho, u, q))
        for a, b, name in zip(res[0], res[1], ("rho", "u", "Q")):
            assert torch.equal(bits(a), bits(b)), f"{name}, layout {layout}"
        assert float(res[0][1].abs().max()) > 1e-3
        del res
        torch.cuda.empty_cache()
