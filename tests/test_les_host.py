"""Host-side helpers of the LES closure pipeline (no GPU): the output-scale calibration
touches only the nu_t scale of the parameter block, and the `scale` command records it."""
import json

import numpy as np

from mojo_lbm_amd import les


def test_scaled_params_changes_only_the_output_scale():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(les.MLP_TOTAL).astype(np.float32)
    q = les.scaled_params(p, 4.0)
    assert q.dtype == np.float32 and q.shape == p.shape
    assert np.array_equal(np.delete(q, les.MLP_NUSCALE), np.delete(p, les.MLP_NUSCALE))
    assert q[les.MLP_NUSCALE] == np.float32(p[les.MLP_NUSCALE] * 4.0)
    assert les.MLP_NUSCALE == les.MLP_TOTAL - 1 == 1445


def test_scale_command_writes_a_calibrated_copy(tmp_path):
    p = np.linspace(-1, 1, les.MLP_TOTAL).astype(np.float32)
    src, dst = tmp_path / "in.json", tmp_path / "out.json"
    src.write_text(json.dumps({"params": [float(v) for v in p], "config": {"collision": "KBC"}}))
    les.main(["scale", "--params", str(src), "--factor", "0.5", "--note", "seed 1", "--out", str(dst)])
    out = json.loads(dst.read_text())
    q = np.asarray(out["params"], np.float32)
    assert q[les.MLP_NUSCALE] == np.float32(p[les.MLP_NUSCALE] * 0.5)
    assert out["calibration"] == [{"nu_scale_factor": 0.5, "note": "seed 1"}]
    assert out["config"] == {"collision": "KBC"}
