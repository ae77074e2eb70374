"""The rest of the reference's public calls through the C ABI and the C++
wrapper, on the GPU:

* tr_bundle_adjust (bundle_adjust_cuda, ba.cuh:17-24) against the oracle's
  LM on the same correspondences, start pose, options and weights, for
  every loss type, in the one-launch form (N <= 4096) and the
  one-iteration-per-launch form (N > 4096), and for N = 0; its BundleStats
  (iterations, rejected steps and final damping equal, costs and norms to
  rounding);
* tr_get_inliers (get_inliers_mx, entry.h:22-25) against the closed form
  in numpy;
* the refinement statistics a per-frame estimate returns, against the
  oracle's;
* turboransac::estimate_absolute_pose, the wrapper with the reference's
  C++ signature (include/turboransac.hpp), bit for bit against the C-ABI
  call;
* the entry points that take device addresses refuse host memory with
  INVALID_ARGUMENT (checked on the host, before any kernel).
"""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import turboransac_amd as tr
from turboransac_amd import _abi, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ba_problem(rng, n, outl=0.2, noise_px=1.0, rot_sd=0.01, t_sd=0.01):
    """Pixels of a 525 px camera, a start pose rot_sd rad and t_sd m off."""
    fr = synth.make_frame(rng, n, 1.0 - outl, noise_px, hypotheses=16)
    q_gt = Rotation.from_matrix(fr.R).as_quat()
    q0 = Rotation.from_rotvec(rng.normal(0, rot_sd, 3)) * Rotation.from_quat(q_gt)
    q0 = q0.as_quat()
    q0 = np.array([q0[3], q0[0], q0[1], q0[2]])
    t0 = fr.t + rng.normal(0, t_sd, 3)
    return fr, q0, t0


def close(a, b, rtol, floor):
    return abs(a - b) <= rtol * abs(b) + floor


# Each problem stops at a step of 1e-6 (rad, m), well above the rounding of
# the cost comparisons: a decision taken at the FP64 noise floor could go
# either way under the device's fused arithmetic.  The "far" problem starts
# 0.03 rad and 5 cm off with initial damping 1e-9 (8 points behind the
# camera) and rejects 29 of its 41 steps; every accept/reject decision on
# its way has a cost gap of at least 5e-6 relative, and starts moved by
# 1e-13 take the same 41 steps.  (Starts farther off, with hundreds of
# points behind the camera, make the 6 x 6 solves near-singular: there a
# 1e-14 change of the start already changes the oracle's own trajectory,
# so no two implementations can be expected to agree step by step.)
@pytest.mark.parametrize("loss,n,weighted,far", [
    (tr.TR_LOSS_CAUCHY, 600, False, False),
    (tr.TR_LOSS_CAUCHY, 9000, True, False),
    (tr.TR_LOSS_TRIVIAL, 1500, False, False),
    (tr.TR_LOSS_TRUNCATED, 3000, True, False),
    (tr.TR_LOSS_HUBER, 5000, False, False),
    (tr.TR_LOSS_TRUNCATED_LE_ZACH, 2500, False, False),
    (tr.TR_LOSS_TRUNCATED_LE_ZACH, 7000, True, False),
    (tr.TR_LOSS_CAUCHY, 3000, False, True),
])
def test_bundle_adjust_matches_oracle(ctx, orc, loss, n, weighted, far):
    if far:
        rng = np.random.default_rng(626)
        fr, q0, t0 = ba_problem(rng, n, rot_sd=0.03, t_sd=0.05)
    else:
        rng = np.random.default_rng(1000 + n + loss)
        fr, q0, t0 = ba_problem(rng, n)
    K = np.array(fr.K)
    w = rng.uniform(0.5, 2.0, n) if weighted else None
    bo = _abi.default_bundle_options(loss_type=loss, loss_scale=3.0, max_iterations=80,
                                     step_tol=1e-6, initial_lambda=1e-9 if far else 1e-3)
    print(f"bundle adjust loss {loss} N={n} weights={weighted}", flush=True)
    q, t, st = ctx.bundle_adjust(fr.points2D, fr.points3D, fr.camera(), bo, q0, t0, weights=w)
    if w is None:
        qr, trf, sr = orc.lm_refine(fr.points2D, fr.points3D, K, bo, q0, t0)
    else:
        qr, trf, sr = orc.lm_refine_weighted(fr.points2D, fr.points3D, w, K, bo, q0, t0)
    rot, cm = synth.pose_error(synth.quat_to_rotmat(q), t, synth.quat_to_rotmat(qr), trf)
    print(f"  vs oracle {rot:.2e} deg {cm:.2e} cm; iterations {st.iterations}/{sr.iterations}, "
          f"rejected {st.invalid_steps}/{sr.invalid_steps}", flush=True)
    assert rot <= 1e-6 and cm <= 1e-6
    assert st.iterations == sr.iterations and st.invalid_steps == sr.invalid_steps
    assert sr.invalid_steps > 0 or not far
    assert st.lambda_ == sr.lambda_
    assert close(st.initial_cost, sr.initial_cost, 1e-12, 0.0)
    assert close(st.cost, sr.cost, 1e-10, 1e-300)
    # the last norms are near zero at convergence: compare against the scale
    # of the first gradient
    g0 = np.linalg.norm(orc.lm_normal_equations(fr.points2D, fr.points3D, K, bo, q0, t0)[1])
    assert close(st.grad_norm, sr.grad_norm, 1e-6, 1e-12 * g0)
    assert close(st.step_norm, sr.step_norm, 1e-6, 1e-12)


def test_bundle_adjust_without_points_keeps_the_pose(ctx):
    cam = _abi.make_camera(525.0, 525.0, 320.0, 240.0)
    q0 = np.array([0.9, 0.1, -0.3, 0.2])
    t0 = np.array([0.5, -0.1, 2.0])
    q, t, st = ctx.bundle_adjust(np.zeros((0, 2)), np.zeros((0, 3)), cam, None, q0, t0)
    assert np.array_equal(q, q0) and np.array_equal(t, t0)
    assert st.iterations == 0 and st.initial_cost == 0.0 and st.grad_norm == 0.0


def test_bundle_adjust_rejects_bad_input(ctx):
    cam = _abi.make_camera(525.0, 525.0, 320.0, 240.0)
    x = np.zeros((4, 2))
    X = np.ones((4, 3))
    with pytest.raises(tr.TRError) as e:
        ctx.bundle_adjust(x, X, cam, _abi.default_bundle_options(loss_type=9))
    assert e.value.status == _abi.TR_ERR_INVALID_ARGUMENT
    x[1, 0] = np.inf
    with pytest.raises(tr.TRError) as e:
        ctx.bundle_adjust(x, X, cam)
    assert e.value.status == _abi.TR_ERR_NOT_FINITE


@pytest.mark.parametrize("n", [1, 255, 256, 257, 100000])
def test_get_inliers_closed_form(ctx, n):
    rng = np.random.default_rng(n)
    R = synth.random_rotation(rng)
    q = Rotation.from_matrix(R).as_quat()
    q = np.array([q[3], q[0], q[1], q[2]])
    # depths in [1.5 - 2 sqrt(3), 1.5 + 2 sqrt(3)]: points behind the camera
    # project (through the division) near their x, so only the depth test
    # rejects them
    t = np.array([0.2, -0.1, 1.5])
    X = rng.uniform(-2, 2, (n, 3))
    Z = X @ R.T + t
    x = Z[:, :2] / Z[:, 2:] + rng.normal(0, 0.02, (n, 2))
    thr = 0.02 ** 2
    mask, cnt = ctx.get_inliers(q, t, x, X, thr)
    Rq = synth.quat_to_rotmat(q)
    Zq = X @ Rq.T + t
    with np.errstate(divide="ignore", invalid="ignore"):
        r2 = np.sum((Zq[:, :2] / Zq[:, 2:] - x) ** 2, axis=1)
    ref = (r2 < thr) & (Zq[:, 2] > 0)
    bad = mask.astype(bool) != ref
    assert not np.any(bad & (np.abs(r2 - thr) > 1e-12 * thr))
    assert cnt == int(mask.sum())
    if n > 10:
        behind = Zq[:, 2] <= 0
        assert 0 < cnt < n and np.any(behind & (r2 < thr))
        assert not np.any(mask.astype(bool) & behind)


def test_estimate_returns_bundle_stats_of_the_oracle(ctx, orc):
    rng = np.random.default_rng(71)
    for loss in (tr.TR_LOSS_CAUCHY, tr.TR_LOSS_TRUNCATED_LE_ZACH):
        fr = synth.make_frame(rng, 3000, 0.4, 1.0, hypotheses=512, seed=9)
        bo = fr.bundle_options(loss_type=loss, loss_scale=2.0, max_iterations=50)
        res, _ = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(),
                                            fr.ransac_options(), bo)
        st, ref, _, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(), fr.ransac_options(),
                                     bo)
        assert st == 0 and ref.status == 0
        print(f"loss {loss}: LM {res.refinements}/{ref.refinements} iterations, rejected "
              f"{res.lm_invalid_steps}/{ref.lm_invalid_steps}", flush=True)
        assert res.refinements == ref.refinements
        assert res.lm_invalid_steps == ref.lm_invalid_steps
        assert res.lm_lambda == ref.lm_lambda
        assert close(res.lm_initial_cost, ref.lm_initial_cost, 1e-12, 0.0)
        assert close(res.lm_cost, ref.lm_cost, 1e-10, 1e-300)
        assert res.lm_grad_norm >= 0.0 and res.lm_step_norm >= 0.0
        rot, cm = synth.pose_error(synth.quat_to_rotmat(res.q), np.array(res.t[:]),
                                   synth.quat_to_rotmat(ref.q), np.array(ref.t[:]))
        assert rot <= 1e-6 and cm <= 1e-6


def test_cpp_wrapper_is_the_c_abi_call(ctx):
    """turboransac::estimate_absolute_pose (the reference's C++ signature)
    in a separate process gives the C-ABI call's bits."""
    exe = os.path.join(ROOT, "turboransac_amd", "lib", "tr_example")
    rng = np.random.default_rng(81)
    for n, ratio, samples in ((2000, 0.6, 1024), (5000, 0.3, 4096)):
        fr = synth.make_frame(rng, n, ratio, 1.2, hypotheses=samples, seed=123 + n)
        with tempfile.TemporaryDirectory() as d:
            fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
            with open(fin, "wb") as f:
                f.write(struct.pack("<q5d3Q", n, *fr.K, fr.max_reproj_error, samples, fr.seed,
                                    fr.lm_max_iterations))
                f.write(np.ascontiguousarray(fr.points2D, np.float64).tobytes())
                f.write(np.ascontiguousarray(fr.points3D, np.float64).tobytes())
            print(f"C++ wrapper on N={n}", flush=True)
            subprocess.run([exe, "--frame", fin, fout], check=True, timeout=120)
            blob = open(fout, "rb").read()
        q = np.frombuffer(blob[:32], np.float64)
        t = np.frombuffer(blob[32:56], np.float64)
        refinements, iterations, nin = struct.unpack("<3Q", blob[56:80])
        ratio_out, score = struct.unpack("<2d", blob[80:96])
        mask = np.frombuffer(blob[96:], np.uint8)
        res, m = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(),
                                            fr.ransac_options(), fr.bundle_options())
        assert list(q) == list(res.q) and list(t) == list(res.t)
        assert (refinements, iterations, nin) == (res.refinements, res.iterations,
                                                  res.num_inliers)
        assert ratio_out == res.inlier_ratio and score == res.model_score
        assert np.array_equal(mask, m)


def test_device_entry_points_refuse_host_memory(ctx):
    """The entry points that take device addresses check them on the host
    (hipPointerGetAttributes) and return INVALID_ARGUMENT for host memory,
    before any kernel could read it; device memory is still accepted."""
    from _hipmem import DeviceArray

    fr = synth.make_frame(np.random.default_rng(71), 300, 0.6, 1.0, hypotheses=128, seed=3)
    p2 = np.ascontiguousarray(fr.points2D, dtype=np.float64)
    p3 = np.ascontiguousarray(fr.points3D, dtype=np.float64)
    d2, d3 = DeviceArray.from_numpy(p2), DeviceArray.from_numpy(p3)
    try:
        for call in (ctx.estimate_absolute_pose_device, ctx.estimate_absolute_pose_device_noalloc):
            for a, b in ((p2.ctypes.data, d3.ptr), (d2.ptr, p3.ctypes.data)):
                with pytest.raises(tr.TRError) as e:
                    call(a, b, fr.n, fr.camera(), fr.ransac_options(), fr.bundle_options())
                assert e.value.status == tr._abi.TR_ERR_INVALID_ARGUMENT
            host_mask = np.zeros(fr.n, dtype=np.uint8)
            with pytest.raises(tr.TRError) as e:
                call(d2.ptr, d3.ptr, fr.n, fr.camera(), fr.ransac_options(), fr.bundle_options(),
                     d_mask=host_mask.ctypes.data)
            assert e.value.status == tr._abi.TR_ERR_INVALID_ARGUMENT
            res = call(d2.ptr, d3.ptr, fr.n, fr.camera(), fr.ransac_options(),
                       fr.bundle_options())
            assert res.status == 0 and res.num_inliers > 0
        srv = tr.Server(0, hbm_budget_bytes=1 << 30, max_batch_frames=8, max_inflight=1)
        with pytest.raises(tr.TRError) as e:
            srv.submit_device(tr.FrameSet([fr]), 0)  # host arrays
        assert e.value.status == tr._abi.TR_ERR_INVALID_ARGUMENT
        dfs = tr.DeviceFrameSet([fr])
        res, _, _ = srv.wait(srv.submit_device(dfs, 0))
        assert res.status == 0 and res.num_inliers > 0
        srv.close()
        dfs.close()
    finally:
        d2.close()
        d3.close()
