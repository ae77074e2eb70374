"""Fundamental-matrix estimation (tr_estimate_fundamental) on the GPU against
the oracle, on the same seeded inputs:

* sampler: the 7-subsets bit for bit;
* 7-point solver: the same solution count per sample and the same
  unit-norm solutions in the same order (FP64, to rounding);
* FP32 Sampson MSAC: every candidate's score within 1e-5 N tau^2 of the
  oracle's FP64 score;
* FP64 re-rank and choice: the same best hypothesis id, its score to 1e-12
  relative, the same candidate count;
* mask: equal, except points whose Sampson error lies within 1e-9 of the
  threshold;
* IRLS refinement and F in pixels: the same iteration count (one more or
  less where the stop test meets the step tolerance at rounding level) and
  F to 1e-8;
* a 64-frame stream of N in [200, 8000], N = 7, 8, 100000, noise-free
  frames (F recovered to 1e-9), a degenerate frame (no solution) and the
  argument errors;
* tr_estimate_fundamental_batch: every frame of a mixed batch (the stream,
  N = 7 and 8, the degenerate frame, other sample counts and thresholds)
  gives the bits of its one-frame call, and a frame the one-frame call
  refuses makes the batch return the same status.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import turboransac_amd as tr
from turboransac_amd import _abi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def numpy_sampson2(Fn, T1, T2, p1, p2):
    x1 = T1[0] * (p1 - T1[1:])
    x2 = T2[0] * (p2 - T2[1:])
    a = np.column_stack([x1, np.ones(len(x1))]) @ Fn.T
    b = np.column_stack([x2, np.ones(len(x2))]) @ Fn
    num = np.einsum("ij,ij->i", np.column_stack([x2, np.ones(len(x2))]), a)
    den = T2[0] ** 2 * (a[:, 0] ** 2 + a[:, 1] ** 2) + T1[0] ** 2 * (b[:, 0] ** 2 + b[:, 1] ** 2)
    return num * num / den


def check_frame(ctx, orc, fr, ro=None, bo=None, debug=0):
    """The device's result against the oracle's on one frame.  The chosen
    hypothesis is the oracle's, or (duplicate 7-subsets, near-degenerate
    frames) one whose oracle FP64 score is within 1e-12 N tau^2 of the
    oracle's best: then the mask and the refined F are checked against that
    model, through the oracle's own IRLS and denormalisation."""
    ro = ro or fr.ransac_options()
    bo = bo or fr.bundle_options()
    H = int(ro.max_iterations)
    st, ref, rmask, trc = orc.estimate_fundamental(fr.points1, fr.points2, ro, bo,
                                                   trace_samples=H)
    assert st == 0
    if debug:
        res, mask, dbg = ctx.estimate_fundamental(fr.points1, fr.points2, ro, bo,
                                                  debug_samples=debug)
    else:
        res, mask = ctx.estimate_fundamental(fr.points1, fr.points2, ro, bo)
    assert res.status == 0
    assert res.iterations == ref.iterations
    assert res.num_candidates == ref.num_candidates
    tau2 = ro.max_epipolar_error ** 2
    if ref.best_hypothesis < 0:
        assert res.best_hypothesis == -1
        assert not mask.any() and np.all(np.array(res.F[:]) == 0.0)
        return res, ref, None
    g, k = divmod(int(res.best_hypothesis), 4)
    s_chosen = trc["scores"][g, k]
    if res.best_hypothesis != ref.best_hypothesis:
        assert s_chosen - ref.model_score <= 1e-12 * fr.n * tau2, (
            res.best_hypothesis, ref.best_hypothesis, s_chosen, ref.model_score)
    assert res.model_score == pytest.approx(s_chosen, rel=1e-12, abs=1e-12)
    # mask of the chosen model: the oracle's e^2, in numpy
    T1, T2 = orc.hartley(fr.points1), orc.hartley(fr.points2)
    Fn = trc["models"][g, k].reshape(3, 3)
    e2 = numpy_sampson2(Fn, T1, T2, fr.points1, fr.points2)
    near = np.abs(e2 - tau2) <= 1e-9 * tau2
    diff = mask.astype(bool) != (e2 < tau2)
    assert not (diff & ~near).any(), np.flatnonzero(diff & ~near)[:10]
    assert res.num_inliers == mask.sum() == res.best_inliers
    if res.best_hypothesis == ref.best_hypothesis and not (
            mask.astype(bool) != rmask.astype(bool)).any():
        assert res.num_inliers == ref.num_inliers
    # refinement from the chosen model on the device's inliers
    x1n = T1[0] * (fr.points1 - T1[1:])
    x2n = T2[0] * (fr.points2 - T2[1:])
    inl = mask.astype(bool)
    Fr, it = orc.fundamental_irls(x1n[inl], x2n[inl], T1[0] ** 2, T2[0] ** 2,
                                  bo.max_iterations, bo.step_tol, Fn)
    G = orc.fundamental_denormalise(Fr, T1, T2).ravel()
    assert abs(int(res.refinements) - it) <= 1, (res.refinements, it)
    F = np.array(res.F[:])
    assert np.abs(F - G).max() < 1e-8, (F, G)
    if debug:
        S = debug
        assert np.array_equal(dbg["samples"], trc["samples"][:S])
        nd, nr = dbg["num_solutions"], trc["num_solutions"][:S]
        same = nd == nr
        assert same.mean() >= 0.99, same.mean()
        models_d = dbg["models"].reshape(S, 3, 9)
        models_r = trc["models"][:S].reshape(S, 3, 9)
        err = []
        for gg in np.flatnonzero(same):
            for kk in range(nr[gg]):
                err.append(np.abs(models_d[gg, kk] - models_r[gg, kk]).max())
        err = np.array(err)
        assert np.median(err) < 1e-12 and np.quantile(err, 0.99) < 1e-8, \
            np.quantile(err, [0.5, 0.99, 1])
        s32 = dbg["scores32"].astype(np.float64)
        s64 = trc["scores"][:S]
        both = np.isfinite(s32) & np.isfinite(s64) & same[:, None]
        assert np.abs(s32[both] - s64[both]).max() <= 1e-5 * fr.n * tau2
        assert np.all(np.isnan(s32[~np.isfinite(s64) & same[:, None]]))
    return res, ref, mask


@pytest.mark.parametrize("n,ratio,sigma,H", [(7, 1.0, 0.0, 64), (8, 1.0, 0.3, 64),
                                             (100, 0.7, 0.5, 256), (1000, 0.5, 0.8, 512),
                                             (5000, 0.3, 1.0, 1024)])
def test_fundamental_debug_parity(ctx, orc, n, ratio, sigma, H):
    rng = np.random.default_rng(900 + n)
    fr = synth.make_pair_frame(rng, n, ratio, sigma, hypotheses=H, threshold=1.5, seed=n)
    check_frame(ctx, orc, fr, debug=H)


def test_fundamental_stream(ctx, orc):
    med = []
    for fr in synth.pair_stream(seed=8, frames=64, hypotheses=1024):
        res, ref, mask = check_frame(ctx, orc, fr)
        F = np.array(res.F[:]).reshape(3, 3)
        e = np.array([synth.sampson_error(F, a, b) for a, b in
                      zip(fr.points1[fr.is_inlier], fr.points2[fr.is_inlier])])
        med.append(float(np.median(e)))
    # 1024 samples rarely hold an all-inlier 7-subset at 30% inliers, so a
    # few frames miss; the typical frame's true inliers sit at the noise
    assert np.median(med) < 1.0


def test_fundamental_large(ctx, orc):
    rng = np.random.default_rng(905)
    fr = synth.make_pair_frame(rng, 100000, 0.4, 0.7, hypotheses=1024, threshold=1.5, seed=5)
    check_frame(ctx, orc, fr)


def test_fundamental_noise_free(ctx):
    rng = np.random.default_rng(906)
    for seed in range(4):
        fr = synth.make_pair_frame(rng, 1500, 0.5, 0.0, hypotheses=512, threshold=1.0, seed=seed,
                                   outlier_margin_px=5.0)
        res, mask = ctx.estimate_fundamental(fr.points1, fr.points2, fr.ransac_options(),
                                             fr.bundle_options())
        assert np.abs(np.array(res.F[:]).reshape(3, 3) - fr.F).max() < 1e-9
        assert np.array_equal(mask.astype(bool), fr.is_inlier)
        assert res.num_inliers == fr.is_inlier.sum()


def test_fundamental_degenerate_and_errors(ctx, orc):
    p = np.tile([[100.0, 200.0]], (20, 1))
    ro = _abi.default_ransac_options(max_iterations=64, seed=1, max_epipolar_error=1.0)
    bo = _abi.default_bundle_options(max_iterations=5)
    res, mask = ctx.estimate_fundamental(p, p, ro, bo)
    st, ref, _, _ = orc.estimate_fundamental(p, p, ro, bo)
    assert res.best_hypothesis == ref.best_hypothesis == -1
    assert res.num_candidates == ref.num_candidates == 0
    assert not mask.any() and res.num_inliers == 0
    rng = np.random.default_rng(907)
    fr = synth.make_pair_frame(rng, 50, 0.8, 0.5)
    with pytest.raises(tr.TRError) as e:
        ctx.estimate_fundamental(fr.points1[:6], fr.points2[:6], ro, bo)
    assert e.value.status == _abi.TR_ERR_TOO_FEW_POINTS
    for bad in (dict(max_epipolar_error=0.0), dict(max_epipolar_error=float("nan")),
                dict(max_iterations=0)):
        with pytest.raises(tr.TRError) as e:
            ctx.estimate_fundamental(fr.points1, fr.points2,
                                     _abi.default_ransac_options(**{**dict(max_iterations=64), **bad}),
                                     bo)
        assert e.value.status == _abi.TR_ERR_INVALID_ARGUMENT
    p1 = fr.points1.copy()
    p1[3, 0] = np.nan
    with pytest.raises(tr.TRError) as e:
        ctx.estimate_fundamental(p1, fr.points2, ro, bo)
    assert e.value.status == _abi.TR_ERR_NOT_FINITE
    # the context still works after the errors
    check_frame(ctx, orc, fr, ro=_abi.default_ransac_options(max_iterations=64, seed=3,
                                                              max_epipolar_error=1.0))


def test_cpp_wrapper_is_the_c_abi_call(ctx):
    """turboransac::estimate_fundamental (include/turboransac.hpp) in a
    separate process gives the C-ABI call's bits."""
    exe = os.path.join(ROOT, "turboransac_amd", "lib", "tr_example")
    rng = np.random.default_rng(910)
    for n, ratio in ((1500, 0.6), (6000, 0.4)):
        fr = synth.make_pair_frame(rng, n, ratio, 0.7, hypotheses=1024, threshold=1.5, seed=n)
        bo = fr.bundle_options()
        with tempfile.TemporaryDirectory() as d:
            fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
            with open(fin, "wb") as f:
                f.write(struct.pack("<qd3Qd", n, fr.threshold, fr.hypotheses, fr.seed,
                                    int(bo.max_iterations), bo.step_tol))
                f.write(np.ascontiguousarray(fr.points1, np.float64).tobytes())
                f.write(np.ascontiguousarray(fr.points2, np.float64).tobytes())
            print(f"C++ fundamental wrapper on N={n}", flush=True)
            subprocess.run([exe, "--fundamental", fin, fout], check=True, timeout=120)
            blob = open(fout, "rb").read()
        F = np.frombuffer(blob[:72], np.float64)
        refinements, nin = struct.unpack("<2Q", blob[72:88])
        (score,) = struct.unpack("<d", blob[88:96])
        (best,) = struct.unpack("<q", blob[96:104])
        mask = np.frombuffer(blob[104:], np.uint8)
        res, m = ctx.estimate_fundamental(fr.points1, fr.points2, fr.ransac_options(), bo)
        assert list(F) == list(res.F)
        assert (refinements, nin, best) == (res.refinements, res.num_inliers, res.best_hypothesis)
        assert score == res.model_score
        assert np.array_equal(mask, m)


def test_cpp_batch_wrapper_is_the_c_abi_call(ctx):
    """turboransac::estimate_fundamental_batch (include/turboransac.hpp) in a
    separate process gives the C-ABI batched call's bits, frame by frame."""
    exe = os.path.join(ROOT, "turboransac_amd", "lib", "tr_example")
    rng = np.random.default_rng(911)
    frames = [synth.make_pair_frame(rng, n, ratio, 0.6, hypotheses=h, threshold=thr, seed=n + 7)
              for n, ratio, h, thr in ((300, 0.8, 256, 1.5), (1500, 0.6, 1024, 1.5),
                                       (4000, 0.4, 512, 2.0), (8, 1.0, 64, 1.5),
                                       (9000, 0.5, 1024, 1.0))]
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<q", len(frames)))
            for fr in frames:
                bo = fr.bundle_options()
                f.write(struct.pack("<qd3Qd", fr.n, fr.threshold, fr.hypotheses, fr.seed,
                                    int(bo.max_iterations), bo.step_tol))
                f.write(np.ascontiguousarray(fr.points1, np.float64).tobytes())
                f.write(np.ascontiguousarray(fr.points2, np.float64).tobytes())
        print(f"C++ fundamental batch wrapper on {len(frames)} frames", flush=True)
        subprocess.run([exe, "--fundamental-batch", fin, fout], check=True, timeout=120)
        blob = open(fout, "rb").read()
    res, masks = ctx.estimate_fundamental_batch(
        [(fr.points1, fr.points2, fr.ransac_options(), fr.bundle_options()) for fr in frames])
    o = 0
    for i, fr in enumerate(frames):
        F = np.frombuffer(blob[o:o + 72], np.float64)
        refinements, nin = struct.unpack("<2Q", blob[o + 72:o + 88])
        (score,) = struct.unpack("<d", blob[o + 88:o + 96])
        (best,) = struct.unpack("<q", blob[o + 96:o + 104])
        mask = np.frombuffer(blob[o + 104:o + 104 + fr.n], np.uint8)
        o += 104 + fr.n
        assert list(F) == list(res[i].F), i
        assert (refinements, nin, best) == (res[i].refinements, res[i].num_inliers,
                                            res[i].best_hypothesis), i
        assert score == res[i].model_score, i
        assert np.array_equal(mask, masks[i]), i
    assert o == len(blob)


def test_fundamental_batch_equals_one_frame_calls(ctx):
    rng = np.random.default_rng(908)
    frames = list(synth.pair_stream(seed=8, frames=64, hypotheses=1024))
    frames.append(synth.make_pair_frame(rng, 7, 1.0, 0.3, hypotheses=64, threshold=1.5, seed=71))
    frames.append(synth.make_pair_frame(rng, 8, 1.0, 0.3, hypotheses=64, threshold=1.5, seed=72))
    frames.append(synth.make_pair_frame(rng, 3000, 0.6, 0.5, hypotheses=200, threshold=3.0,
                                        seed=73))
    pairs = [(fr.points1, fr.points2, fr.ransac_options(), fr.bundle_options()) for fr in frames]
    deg = np.tile([[100.0, 200.0]], (20, 1))
    pairs.append((deg, deg, _abi.default_ransac_options(max_iterations=64, seed=1,
                                                         max_epipolar_error=1.0),
                  _abi.default_bundle_options(max_iterations=5)))
    print(f"fundamental batch of {len(pairs)} frames", flush=True)
    res, masks = ctx.estimate_fundamental_batch(pairs)
    for i, (a, b, ro, bo) in enumerate(pairs):
        one, m1 = ctx.estimate_fundamental(a, b, ro, bo)
        for name, _ in _abi.tr_fundamental_result._fields_:
            va, vb = getattr(res[i], name), getattr(one, name)
            if hasattr(va, "__len__"):
                assert list(va) == list(vb), (i, name)
            else:
                assert va == vb, (i, name, va, vb)
        assert np.array_equal(masks[i], m1), i
    assert res[-1].best_hypothesis == -1 and not masks[-1].any()
    # a frame the one-frame call refuses: the same status for the batch
    short = pairs[:3] + [(pairs[3][0][:6], pairs[3][1][:6], pairs[3][2], pairs[3][3])]
    with pytest.raises(tr.TRError) as e:
        ctx.estimate_fundamental_batch(short)
    assert e.value.status == _abi.TR_ERR_TOO_FEW_POINTS
    p1 = pairs[1][0].copy()
    p1[2, 1] = np.inf
    with pytest.raises(tr.TRError) as e:
        ctx.estimate_fundamental_batch([pairs[0], (p1, pairs[1][1], pairs[1][2], pairs[1][3])])
    assert e.value.status == _abi.TR_ERR_NOT_FINITE
    # and the context still works
    res2, _ = ctx.estimate_fundamental_batch(pairs[:2], masks=False)
    assert list(res2[1].F) == list(res[1].F)
    # inputs go up in pieces of 4 MiB as they are staged: a NaN found after
    # several pieces have been queued (the stream's 64 frames, about 5.3 MB,
    # then the NaN frame) ends the call with the same status, and a
    # following call over the same staging buffer gives the same bits
    big = pairs[:64] * 2
    p2 = big[100][1].copy()
    p2[-1, 0] = np.nan
    bad = big[:100] + [(big[100][0], p2, big[100][2], big[100][3])] + big[101:]
    print(f"fundamental batch of {len(bad)} frames, NaN in frame 100", flush=True)
    with pytest.raises(tr.TRError) as e:
        ctx.estimate_fundamental_batch(bad)
    assert e.value.status == _abi.TR_ERR_NOT_FINITE and "frame 100" in str(e.value)
    res3, masks3 = ctx.estimate_fundamental_batch(big)
    for i in range(len(big)):
        assert list(res3[i].F) == list(res[i % 64].F), i
        assert np.array_equal(masks3[i], masks[i % 64]), i
