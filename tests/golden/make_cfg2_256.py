"""Writes cfg2_d3q19_256_1000.json: the oracle's state of the 256^3 D3Q19 SRT FP32 lid-driven cube
after 1,000 double-buffer pull steps (u_lid = 0.1, Re = 100, tau = 0.5 + 3 * 0.1 * 256 / 100 = 1.268,
started from equilibrium at rho = 1, u = 0; walls in the order of benchmarks/LDC/LDC_3D_Timed.mojo:88-93:
+Y lid, then -Y, +X, -X, -Z, +Z at rest; the timed loop of :121-124).

The fixture holds CRC-32s of the final populations and of rho and u (the bytes of the float32 arrays),
the velocity on the two centrelines through the cube's centre, and the FP64 totals of mass and
momentum. tests/test_gpu_cfg2_golden.py compares the HIP path's run against it.

Run from the repository root (about 15 minutes on 8 cores):  python -m tests.golden.make_cfg2_256
"""
from __future__ import annotations

import json
import time
import zlib
from pathlib import Path

import numpy as np

from oracle import oracle as O
from tests import _util as U

N, STEPS, U_LID, RE = 256, 1000, 0.1, 100.0
OUT = Path(__file__).resolve().parent / "cfg2_d3q19_256_1000.json"


def main() -> None:
    tau = 0.5 + 3 * U_LID * N / RE
    desc = U.make_desc("D3Q19", (N, N, N), storage=0)
    f, bc, flags = O.alloc(desc)
    O.init_equilibrium(desc, f)
    U.ldc_walls(desc, flags, bc, U_LID)
    O.set_threads(O.max_threads())
    a, b = f, f.copy()
    t0 = time.time()
    for k in range(STEPS):
        O.step_db(desc, b, a, bc, flags, tau)
        a, b = b, a
        if (k + 1) % 50 == 0:
            print(f"step {k + 1}/{STEPS}, {time.time() - t0:.0f} s", flush=True)
    rho, u = O.macroscopic(desc, a, bc, flags)
    tot = O.reduce_moments(desc, a, bc, flags)
    u3 = u.reshape(N, N, N, 3)
    c = N // 2
    doc = {
        "workload": "D3Q19 SRT FP32 lid-driven cube 256^3, double-buffer pull, 8x8x8 tiles, u_lid 0.1, Re 100, "
                    "tau 1.268, 1000 steps from rest",
        "tau": tau,
        "steps": STEPS,
        "crc32_f": zlib.crc32(a.tobytes()),
        "crc32_rho": zlib.crc32(rho.tobytes()),
        "crc32_u": zlib.crc32(u.tobytes()),
        "u_line_y": u3[c, :, c, :].astype(float).tolist(),  # u(x = 128, y, z = 128), (y, d)
        "u_line_x": u3[:, c, c, :].astype(float).tolist(),  # u(x, y = 128, z = 128), (x, d)
        "totals_f64": tot.tolist(),
        "u_max_abs": float(np.max(np.abs(u))),
        "oracle_seconds": round(time.time() - t0, 1),
    }
    OUT.write_text(json.dumps(doc) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
