"""The device Float16C codec (the one the F16C population storage uses) against the
oracle's codec, which is pinned to the published format in test_oracle.py:
all 65,536 codes decoded, and encoding of every code value, every midpoint between
adjacent codes (ties), values just below them, random magnitudes and specials."""
import ctypes as C

import numpy as np
import pytest

from mojo_lbm_amd import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _convert(gpu, arr, to_f16c):
    import torch

    src = torch.from_numpy(arr).to(gpu)
    out = torch.empty(arr.size, dtype=torch.uint16 if to_f16c else torch.float32, device=gpu)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(_lib.lib().lbm_f16c_convert(src.data_ptr(), out.data_ptr(), arr.size, int(to_f16c), st))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_decode_all_codes(gpu):
    codes = np.arange(65536, dtype=np.uint16)
    got = _convert(gpu, codes, False)
    np.testing.assert_array_equal(got.view(np.uint32), O.f16c_decode(codes).view(np.uint32))


def test_encode_matches_oracle(gpu):
    vals = O.f16c_decode(np.arange(65536, dtype=np.uint16)).astype(np.float64)
    pos = vals[:0x8000]
    mids = ((pos[:-1] + pos[1:]) / 2).astype(np.float32)
    rng = np.random.default_rng(7)
    rand = np.concatenate([rng.uniform(-2.5, 2.5, 200000), np.exp2(rng.uniform(-40, 2, 200000)) *
                           rng.choice([-1, 1], 200000)]).astype(np.float32)
    specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 2.0, -2.0, 1.9999, 3e38, 1e-45, -1e-45,
                         2.0 ** -26, 2.0 ** -14, 2.0 ** -14 - 2.0 ** -26], np.float32)
    x = np.concatenate([vals.astype(np.float32), mids, np.nextafter(mids, np.float32(0)),
                        np.nextafter(mids, np.float32(3)), rand, specials])
    got = _convert(gpu, x, True)
    np.testing.assert_array_equal(got, O.f16c_encode(x))


def test_encode_all_inputs(gpu):
    """Every one of the 2^32 fp32 bit patterns (denormals, infinities and NaNs included),
    encoded on the device, equals the oracle's code: in chunks of 2^27, the oracle run on
    host threads."""
    import torch
    from concurrent.futures import ThreadPoolExecutor

    chunk, parts = 1 << 27, 16
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = torch.empty(chunk, dtype=torch.uint16, device=gpu)
    ref = np.empty(chunk, np.uint16)

    def oracle_part(x, k):
        n = chunk // parts
        ref[k * n:(k + 1) * n] = O.f16c_encode(x[k * n:(k + 1) * n])

    with ThreadPoolExecutor(parts) as pool:
        for c in range((1 << 32) // chunk):
            start = c * chunk
            src = torch.arange(start, start + chunk, dtype=torch.int64, device=gpu).to(torch.int32)
            _lib.check(_lib.lib().lbm_f16c_convert(src.data_ptr(), out.data_ptr(), chunk, 1, st))
            x = np.arange(start, start + chunk, dtype=np.uint64).astype(np.uint32).view(np.float32)
            list(pool.map(lambda k: oracle_part(x, k), range(parts)))
            got = out.cpu().numpy()
            bad = np.flatnonzero(got != ref)
            assert bad.size == 0, (f"chunk {c}: {bad.size} codes differ, first input bits "
                                   f"{int(start + bad[0]):#010x}: {got[bad[0]]:#06x} vs {ref[bad[0]]:#06x}")
            del src
