"""Builds liblbm_hip.so (gfx950) in-tree with hipcc.

Each lbm_inst_*.hip compilation unit holds the kernels of one (velocity set,
storage) pair; they compile in parallel and link with lbm_capi.hip into
mojo_lbm_amd/liblbm_hip.so. Run as ``python -m mojo_lbm_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "liblbm_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every multiply and add is rounded as written, matching the
# reference's operation order bit for bit. Packed FP32 instructions are disabled.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def _up_to_date(src: Path, obj: Path) -> bool:
    if not obj.exists():
        return False
    deps = [src, *CSRC.glob("*.hpp"), PKG.parent / "include" / "lbm_hip.h"]
    return obj.stat().st_mtime > max(d.stat().st_mtime for d in deps if d.exists())


def _compile(src: Path, incremental: bool = False) -> Path:
    obj = BUILD / (src.stem + ".o")
    if incremental and _up_to_date(src, obj):
        return obj
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True, incremental: bool = False) -> Path:
    """Compiles every HIP source (only out-of-date objects when incremental) and links
    liblbm_hip.so; returns its path."""
    BUILD.mkdir(exist_ok=True)
    # the instantiation units of the larger velocity sets first: they take the longest
    srcs = sorted(CSRC.glob("*.hip"), key=lambda p: (-("d3q27" in p.name) - ("d3q19" in p.name), p.name))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("-")]
    build(int(args[0]) if args else None, incremental="--incremental" in sys.argv)
