"""The 7-Scenes evaluator end to end on a synthetic scene file."""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import turboransac_amd as tr
from turboransac_amd import _abi
from turboransac_amd import build as B
from turboransac_amd import synth

pytestmark = pytest.mark.gpu


def _eval(data, *args):
    r = subprocess.run([B.EVAL_BIN, "--data", str(data), "--scenes", "syn", "--save-pred", *args],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Scene: syn\s+\| Avg AE = +([0-9.]+) deg, Avg TE = +([0-9.]+) cm, "
                  r"Avg Time = +([0-9.]+) ms \(\s*(\d+) frames, (\d+) failed\)", r.stdout)
    assert m, r.stdout
    with open(os.path.join(data, "syn_pred.json")) as f:
        pred = json.load(f)
    return [float(m.group(1)), float(m.group(2)), int(m.group(4)), int(m.group(5))], pred


def test_evaluator_matches_the_api(ctx, tmp_path):
    rng = np.random.default_rng(8)
    frames = {f"seq-01/frame-{i:06d}": synth.make_frame(rng, int(n), w, 1.0, seed=i)
              for i, (n, w) in enumerate([(1500, 0.6), (4000, 0.3), (2500, 0.8), (800, 0.5)])}
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    synth.write_seven_scenes(str(a / "syn.json"), frames)
    shutil.copytree(a, b)
    print("evaluator per frame", flush=True)
    (ae, te, nf, failed), pred = _eval(a)
    print("evaluator batched", flush=True)
    (ae_b, te_b, nf_b, _), pred_b = _eval(b, "--batch")
    assert nf == nf_b == len(frames) and failed == 0
    assert pred == pred_b  # the batch and per-frame paths give the same bits
    errs = []
    for name, fr in frames.items():
        # the evaluator's options: 1024 samples per round, success 0.9999, defaults
        ro = _abi.default_ransac_options(max_iterations=1024, success_prob=0.9999)
        fx, fy, cx, cy = fr.K
        cam = _abi.make_camera(fx, fy, cx, cy, width=int(cx) * 2, height=int(cy) * 2)
        res, _ = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, cam, ro,
                                            _abi.default_bundle_options())
        M = np.array(pred[name]["pred_w2c"])
        assert np.allclose(M[:3, :3], synth.quat_to_rotmat(res.q), atol=1e-15)
        assert np.array_equal(M[:3, 3], np.array(res.t[:])) and np.array_equal(M[3], [0, 0, 0, 1])
        errs.append(synth.pose_error(M[:3, :3], M[:3, 3], fr.R, fr.t))
    errs = np.array(errs)
    assert abs(errs[:, 0].mean() - ae) <= 1e-6 and abs(errs[:, 1].mean() - te) <= 1e-6
    assert ae < 0.5 and te < 2.0
