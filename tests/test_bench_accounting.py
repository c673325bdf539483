"""The benchmark's accounting without a GPU: algorithmic bytes per step of the 256^3 D3Q19
lid-driven cube (N_f = 254^3 fluid nodes read and written, one flag byte per node) for each
storage, the MLUPs convention (all N_c nodes, walls included) and the multi-GPU refusal."""
import subprocess
import sys
from pathlib import Path

import bench

ROOT = Path(__file__).resolve().parents[1]


def test_algorithmic_bytes_256_cube():
    assert bench.algorithmic_bytes(19, 256, 256, 256, 4) == 2_507_610_944
    assert bench.algorithmic_bytes(19, 256, 256, 256, 2) == 1_262_194_080
    assert bench.algorithmic_bytes(19, 256, 256, 256, 8) == 4_998_444_672
    # the periodic box has no walls: every node is fluid
    assert bench.algorithmic_bytes(19, 256, 256, 256, 4, walls=False) == 256 ** 3 * (2 * 19 * 4 + 1)


def test_multi_gpu_request_is_refused():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         cwd=ROOT, timeout=120)
    assert res.returncode != 0
    assert "GPU" in (res.stdout + res.stderr)
