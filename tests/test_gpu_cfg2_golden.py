"""BL-cfg2 at full size and full length: the 256^3 D3Q19 SRT FP32 lid-driven cube (u_lid = 0.1,
Re = 100, tau = 1.268, walls in the order of benchmarks/LDC/LDC_3D_Timed.mojo:88-93), 1,000
double-buffer pull steps through the host API (the reference's timed loop, :121-124), against the
oracle's run of the same case committed in tests/golden/cfg2_d3q19_256_1000.json (written by
tests/golden/make_cfg2_256.py): the populations, rho and u bit-identical (CRC-32 of their bytes),
the velocity on the two centrelines equal, and the FP64 totals equal to summation order. The
contract tolerance (velocity relative L2 <= 1e-5) is checked on the centrelines as well."""
import json
import zlib
from pathlib import Path

import numpy as np
import pytest

import mojo_lbm_amd as lbm

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "cfg2_d3q19_256_1000.json"


def test_cfg2_256_cube_1000_steps_matches_oracle(gpu):
    import torch

    ref = json.loads(GOLDEN.read_text())
    N, u_lid = 256, 0.1
    tau = lbm.tau_from_re(u_lid, N, 100.0)
    assert tau == ref["tau"]
    print("cfg2 golden: D3Q19 256^3 FP32, 1000 steps", flush=True)
    cfg = lbm.DoubleBufferConfig()
    grid = lbm.LBM_Grid(lbm.get_D3Q19(), N, N, N, 8)
    solver = lbm.Solver(grid, cfg)
    asm = lbm.Assembly(grid, cfg)
    lbm.initialize_fluid_at_rest(grid, cfg, asm.f, solver)
    lbm.initialize_fluid_at_rest(grid, cfg, asm.f2, solver)
    lbm.set_exterior_walls(grid, cfg, asm.flags, asm.bc, "+Y", lbm.Flags.SOLID, [u_lid, 0.0, 0.0], 1.0)
    for side in ["-Y", "+X", "-X", "-Z", "+Z"]:
        lbm.set_exterior_walls(grid, cfg, asm.flags, asm.bc, side, lbm.Flags.SOLID, [0.0, 0.0, 0.0], 1.0)
    for _ in range(ref["steps"] // 2):
        solver.even_step(asm, tau)
        solver.odd_step(asm, tau)
    rho, u = lbm.calculate_rho_and_velocity(solver, asm, after_odd_step=True)
    f = asm.f.cpu()
    assert zlib.crc32(np.ascontiguousarray(f).tobytes()) == ref["crc32_f"], "populations differ from the oracle's"
    assert zlib.crc32(np.ascontiguousarray(rho, np.float32).tobytes()) == ref["crc32_rho"]
    assert zlib.crc32(np.ascontiguousarray(u, np.float32).tobytes()) == ref["crc32_u"]
    u3 = np.asarray(u, np.float32).reshape(N, N, N, 3)
    c = N // 2
    for got, key in ((u3[c, :, c, :], "u_line_y"), (u3[:, c, c, :], "u_line_x")):
        want = np.array(ref[key], np.float32)
        rel = np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want.astype(np.float64))
        assert rel <= 1e-5, f"{key}: relative L2 {rel:.3g} above the 1e-5 contract"
        np.testing.assert_array_equal(got, want)
    from mojo_lbm_amd import _lib
    from mojo_lbm_amd.solver import _stream

    tot = torch.zeros(4, dtype=torch.float64, device="cuda")
    _lib.check(solver.L.lbm_reduce_moments_f64(solver.handle.ptr, tot.data_ptr(), asm.f.gpu().data_ptr(),
                                               asm.bc.gpu().data_ptr(), asm.flags.gpu().data_ptr(), 1, _stream()))
    torch.cuda.synchronize()
    np.testing.assert_allclose(tot.cpu().numpy(), ref["totals_f64"], rtol=1e-12, atol=1e-12)
