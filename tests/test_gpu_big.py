"""Buffers of more than 2^32 populations (64-bit element offsets in the kernels that buffers of
2 GiB and more use). A 640^3 D3Q19 Float16C + DDF channel -- walls at -Y and +Y, the +Y one moving
in x, periodic in x and z -- holds 4.98e9 populations per buffer (9.96 GB); its flow depends on y
only, so after a few steps every (x, z) column must equal, bit for bit, the column of a 64 x 640 x 64
run of the same case, whose buffers need no 64-bit offsets and which is compared with the oracle.
The column at x = z = 639 lies past element 2^32 of the big buffer (tile 505,600 of 512,000): an
offset that wrapped at 2^32 would read and write other tiles and q planes there."""
import ctypes as C

import numpy as np
import pytest

from mojo_lbm_amd import _lib
from oracle import oracle as O
from tests import _util as U

pytestmark = pytest.mark.gpu

NY, TAU, STEPS = 640, 0.6, 4


def _column_index(desc, x, z):
    """Element indices (y, q) of the populations of column (x, :, z) in the tiled f layout."""
    import torch

    tx, ty, tz, Q = desc.tx, desc.ty, desc.tz, desc.Q
    Tx, Ty = desc.nx // tx, desc.ny // ty
    y = torch.arange(desc.ny, dtype=torch.int64)
    tile = (x // tx) + Tx * ((y // ty) + Ty * (z // tz))
    loc = (x % tx) + tx * ((y % ty) + ty * (z % tz))
    q = torch.arange(Q, dtype=torch.int64)
    return ((tile[:, None] * Q + q[None, :]) * (tx * ty * tz) + loc[:, None]).cuda()


def _run(n, method):
    """The channel on the GPU through the C ABI; returns (desc, f tensor (int16 codes), extra)."""
    import torch

    desc = U.make_desc("D3Q19", n, storage=1, ddf=1, method=method)
    L = _lib.lib()
    h = _lib.Handle(desc)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    fb = L.lbm_f_bytes(C.byref(desc))
    f = torch.empty(fb // 2, dtype=torch.int16, device="cuda")
    f2 = torch.empty_like(f) if method == 0 else None
    bc = torch.zeros(L.lbm_bc_bytes(C.byref(desc)), dtype=torch.uint8, device="cuda")
    flags = torch.zeros(L.lbm_flags_bytes(C.byref(desc)), dtype=torch.uint8, device="cuda")
    for t in (f, f2):
        if t is not None:
            _lib.check(L.lbm_init_equilibrium(h.ptr, t.data_ptr(), 1.0, None, st))
    _lib.check(L.lbm_set_exterior_wall(h.ptr, flags.data_ptr(), bc.data_ptr(), b"+Y", 1, _lib.dvec([0.05, 0.0, 0.0]),
                                       1.0, st))
    _lib.check(L.lbm_set_exterior_wall(h.ptr, flags.data_ptr(), bc.data_ptr(), b"-Y", 1, _lib.dvec([0.0, 0.0, 0.0]),
                                       1.0, st))
    for k in range(STEPS):
        if method == 0:
            a, b = (f, f2) if k % 2 == 0 else (f2, f)
            _lib.check(L.lbm_step_db(h.ptr, b.data_ptr(), a.data_ptr(), bc.data_ptr(), flags.data_ptr(), TAU, st))
        else:
            _lib.check(L.lbm_step_ep(h.ptr, f.data_ptr(), bc.data_ptr(), flags.data_ptr(), TAU, int(k % 2 == 0), st))
    tot = torch.zeros(4, dtype=torch.float64, device="cuda")
    _lib.check(L.lbm_reduce_moments_f64(h.ptr, tot.data_ptr(), f.data_ptr(), bc.data_ptr(), flags.data_ptr(),
                                        int(STEPS % 2 == 0), st))
    torch.cuda.synchronize()
    return desc, f, (bc, flags, tot.cpu().numpy())


@pytest.mark.parametrize("method", [0, 1])
def test_more_than_2_32_populations(gpu, method):
    import torch

    print(f"case 640^3 Float16C channel, method={method}", flush=True)
    small_desc, small_f, (sbc, sflags, stot) = _run((64, NY, 64), method)
    ref = small_f[_column_index(small_desc, 5, 7)].cpu().numpy()
    # the small run against the oracle
    f0, bc0, fl0 = O.alloc(small_desc)
    O.init_equilibrium(small_desc, f0)
    O.set_exterior_wall(small_desc, fl0, bc0, "+Y", 1, [0.05, 0.0, 0.0], 1.0)
    O.set_exterior_wall(small_desc, fl0, bc0, "-Y", 1, [0.0, 0.0, 0.0], 1.0)
    if method == 0:
        want = U.oracle_run_db(small_desc, f0, bc0, fl0, TAU, STEPS)
    else:
        want = U.oracle_run_ep(small_desc, f0, bc0, fl0, TAU, STEPS)
    assert np.array_equal(small_f.cpu().numpy().view(np.uint16), want)
    del small_f, sbc, sflags
    big_desc, big_f, (bbc, bflags, btot) = _run((640, NY, 640), method)
    assert big_f.numel() > 2 ** 32
    for x, z in ((0, 0), (331, 17), (639, 639), (100, 600)):
        col = big_f[_column_index(big_desc, x, z)].cpu().numpy()
        assert np.array_equal(col, ref), f"column ({x}, {z}) differs from the small run"
    # the FP64 totals scale with the number of columns
    np.testing.assert_allclose(btot, stot * (640 * 640) / (64 * 64), rtol=1e-9, atol=1e-6)
    del big_f, bbc, bflags
    torch.cuda.empty_cache()
