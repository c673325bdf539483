"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (host code only)
and driven through every entry point on small grids of each lattice, storage and method
(tests/sanitize/oracle_driver.cpp): no out-of-bounds access, no undefined behaviour."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_sanitized"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           str(ROOT / "oracle" / "lbm_oracle.cpp"), str(ROOT / "tests" / "sanitize" / "oracle_driver.cpp"),
           "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert res.returncode == 0, (res.stdout + res.stderr)[-3000:]
    assert "ok" in res.stdout
