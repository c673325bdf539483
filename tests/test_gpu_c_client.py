"""The library driven from plain C (tests/c_client/lbm_c_client.c: only include/lbm_hip.h and the
library's device-memory helpers, no Python or torch on the path): a D3Q19 lid-driven cube with both
methods. The FP64 totals it prints equal those of the same run driven from Python, bit for bit,
mass is conserved and the lid drives the flow."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests import _util as U

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def _client():
    exe = ROOT / "tests" / "c_client" / "lbm_c_client"
    if not exe.exists():
        import __graft_entry__

        __graft_entry__.build_c_client()
    return exe


def _python_totals(method, n, steps):
    import torch

    desc = U.make_desc("D3Q19", (n, n, n), method=method)
    f, bc, flags = U.perturbed_state(desc, amp=0.0)
    run = U.DeviceRun(desc, f, bc, flags)
    tau = 0.5 + 3.0 * 0.1 * n / 100.0
    for s in range(steps):
        if method == 0:
            run.step_db(tau)
        else:
            run.step_ep(tau, s % 2 == 0)
    torch.cuda.synchronize()
    return run.reduce(after_odd=steps > 0 and (steps - 1) % 2 == 1)


def test_c_client(gpu):
    res = subprocess.run([str(_client()), "32", "200"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = [json.loads(l) for l in res.stdout.splitlines() if l.startswith("{")]
    assert [r["method"] for r in lines] == [0, 1]
    for r in lines:
        assert abs(r["mass"] - r["mass0"]) < 5e-5 * r["mass0"]  # FP32 without DDF drifts ~1e-5 here
        assert r["u_below_lid"] > 1e-3
        py = _python_totals(r["method"], r["n"], r["steps"])
        assert py[0] == r["mass"], (py, r)
        assert np.allclose(py[1:], r["momentum"], rtol=1e-5, atol=1e-12)
