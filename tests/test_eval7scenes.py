"""The 7-Scenes evaluator's reader (no GPU): which frames and rows it keeps."""
import json
import os
import subprocess

import numpy as np
import pytest

from turboransac_amd import build as B
from turboransac_amd import synth


def _run(*args):
    if not os.path.exists(B.EVAL_BIN):
        pytest.skip("evaluator not built")
    return subprocess.run([B.EVAL_BIN, *args], capture_output=True, text=True, timeout=60)


def test_reader_keeps_valid_frames_and_rows(tmp_path):
    rng = np.random.default_rng(1)
    good = {f"seq-01/frame-{i:06d}": synth.seven_scenes_record(
        synth.make_frame(rng, 50 + 10 * i, 0.7, 1.0, hypotheses=64, seed=i)) for i in range(3)}
    scene = dict(good)
    scene["no-gt"] = dict(good["seq-01/frame-000000"])
    del scene["no-gt"]["gt_w2c"]
    scene["null-K"] = dict(good["seq-01/frame-000001"], **{"intrinsic K": None})
    bad_rows = dict(good["seq-01/frame-000002"])
    bad_rows["point pair"] = bad_rows["point pair"] + [[1.0, 2.0, 3.0, 4.0], "x",
                                                       [1, 2, 3, 4, 5, 6], [1, 2, "3", 4, 5]]
    bad_rows["extra"] = {"nested": [1, {"a": True}], "s": 'q"uote'}
    scene['seq-02/frame-"esc"'] = bad_rows
    (tmp_path / "s.json").write_text(json.dumps(scene, indent=1))
    r = _run("--parse-only", "--data", str(tmp_path), "--scenes", "s")
    assert r.returncode == 0, r.stderr
    rows = sum(len(v["point pair"]) for v in good.values()) + len(good["seq-01/frame-000002"]["point pair"])
    assert f"Loaded 4 valid frames ({rows} correspondences)" in r.stdout, r.stdout


def test_reader_reports_malformed_files(tmp_path):
    (tmp_path / "bad.json").write_text('{"a": {"gt_w2c": [[1, 2], ')
    r = _run("--parse-only", "--data", str(tmp_path), "--scenes", "bad,missing")
    assert r.returncode == 0
    assert "Failed to load JSON" in r.stderr and "cannot open" in r.stderr
