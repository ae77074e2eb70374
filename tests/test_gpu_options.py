"""GPU parity of PROSAC, refinement weights and the initial model against
the oracle, and across the per-frame, batch and server paths."""
import numpy as np
import pytest

import turboransac_amd as tr
from turboransac_amd import synth

from test_gpu_parity import check_frame, pose_diff

pytestmark = pytest.mark.gpu


def test_prosac_stages_match_oracle(ctx, orc):
    rng = np.random.default_rng(51)
    fr = synth.make_frame(rng, 3000, 0.35, 1.0, hypotheses=2048, seed=12)
    ro = fr.ransac_options(progressive_sampling=1, max_prosac_iterations=1500)
    cam, bo = fr.camera(), fr.bundle_options()
    print("prosac frame N=3000 H=2048 max_prosac=1500", flush=True)
    res, mask, dbg = ctx.debug_estimate(fr.points2D, fr.points3D, cam, ro, bo, max_samples=2048)
    assert np.array_equal(dbg["samples"], orc.sample_prosac(fr.seed, fr.n, 1500, 1500, 0, 2048))
    st, ref, ref_mask, trc = orc.estimate(fr.points2D, fr.points3D, cam, ro, bo,
                                          trace_samples=2048)
    check_frame(fr, res, mask, dbg, ref, ref_mask, trc, orc)


def test_prosac_adaptive_batch_equals_per_frame(ctx, orc):
    rng = np.random.default_rng(52)
    frames = [synth.make_frame(rng, n, w, 1.0, hypotheses=256, seed=30 + k)
              for k, (n, w) in enumerate([(2000, 0.3), (900, 0.7), (4000, 0.2)])]
    ro = dict(progressive_sampling=1, max_prosac_iterations=3000, adaptive_stopping=1,
              min_iterations=0, max_total_iterations=4096)
    fs = tr.FrameSet(frames)
    for i, fr in enumerate(frames):
        fs.arr[i].ransac = fr.ransac_options(**ro)
    b = tr.Batch(ctx, fs)
    b.launch()
    b.sync()
    res, _ = b.results()
    b.close()
    for i, fr in enumerate(frames):
        r = fr.ransac_options(**ro)
        one, _ = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(), r,
                                            fr.bundle_options())
        st, ref, _, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(), r, fr.bundle_options())
        assert res[i].iterations == one.iterations == ref.iterations, i
        assert res[i].best_hypothesis == one.best_hypothesis == ref.best_hypothesis, i
        assert list(res[i].q) == list(one.q), i
        rot, cm = pose_diff(one, ref)
        assert rot <= 1e-6 and cm <= 1e-6


def test_weights_match_oracle_on_every_path(ctx, orc):
    rng = np.random.default_rng(53)
    fr = synth.make_frame(rng, 6000, 0.5, 1.5, hypotheses=512, seed=21)
    fr.weights = rng.uniform(0.1, 4.0, fr.n)
    one, m1 = ctx.estimate_frame(fr)
    st, ref, ref_mask, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(),
                                        fr.ransac_options(), fr.bundle_options(),
                                        weights=fr.weights)
    rot, cm = pose_diff(one, ref)
    print(f"weighted: vs oracle {rot:.2e} deg {cm:.2e} cm, LM {one.refinements}/{ref.refinements}",
          flush=True)
    assert rot <= 1e-6 and cm <= 1e-6 and one.refinements == ref.refinements
    assert np.array_equal(m1, ref_mask) or (m1 != ref_mask).sum() <= 2
    unweighted, _ = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(),
                                               fr.ransac_options(), fr.bundle_options())
    assert list(unweighted.q) != list(one.q)  # the weights act
    # unit weights are bit-identical to none
    ones = synth.Frame(**{**fr.__dict__, "weights": np.ones(fr.n)})
    u, _ = ctx.estimate_frame(ones)
    assert list(u.q) == list(unweighted.q) and u.lm_cost == unweighted.lm_cost
    # batch and server: the same bits
    b = tr.Batch(ctx, [fr])
    b.launch()
    b.sync()
    rb, _ = b.results()
    b.close()
    assert list(rb[0].q) == list(one.q)
    srv = tr.Server(0, max_batch_frames=8, max_inflight=1)
    fs = tr.FrameSet([fr])
    rs, _, _ = srv.wait(srv.submit(fs, 0))
    srv.close()
    assert list(rs.q) == list(one.q)


def test_partly_unit_weights_match_oracle_on_every_path(ctx, orc):
    """Half of the weights exactly 1 and half not: within a lane the Cauchy
    cost takes both the product form (unit weights) and per-term log1p."""
    rng = np.random.default_rng(55)
    frames = []
    for k, (n, w) in enumerate([(6000, 0.5), (3000, 0.6)]):
        fr = synth.make_frame(rng, n, w, 1.5, hypotheses=512, seed=60 + k)
        wts = rng.uniform(0.1, 4.0, n)
        wts[::2] = 1.0
        fr.weights = wts
        frames.append(fr)
    ones = []
    for fr in frames:
        print(f"partly unit weights: N={fr.n}", flush=True)
        one, m1 = ctx.estimate_frame(fr)
        st, ref, ref_mask, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(),
                                            fr.ransac_options(), fr.bundle_options(),
                                            weights=fr.weights)
        rot, cm = pose_diff(one, ref)
        assert rot <= 1e-6 and cm <= 1e-6 and one.refinements == ref.refinements
        assert abs(one.lm_cost - ref.lm_cost) <= 1e-9 * abs(ref.lm_cost)
        ones.append(one)
    # five frames: the batch refines frames above 2048 points one LM
    # iteration per launch; the same bits as one frame per call
    extra = [synth.make_frame(rng, 2500, 0.5, 1.0, hypotheses=512, seed=70 + k) for k in range(3)]
    b = tr.Batch(ctx, frames + extra)
    b.launch()
    b.sync()
    rb, _ = b.results()
    b.close()
    for i in range(2):
        assert list(rb[i].q) == list(ones[i].q) and list(rb[i].t) == list(ones[i].t)
        assert rb[i].lm_cost == ones[i].lm_cost


def test_initial_model_on_every_path(ctx, orc):
    rng = np.random.default_rng(54)
    fr = synth.make_frame(rng, 1500, 0.25, 1.0, hypotheses=16, seed=4)
    q_gt = orc.rotmat_to_quat(fr.R)
    fr.initial = (q_gt, fr.t)
    ro = fr.ransac_options(score_initial_model=1)
    st, ref, ref_mask, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(), ro,
                                        fr.bundle_options(), initial=fr.initial)
    assert ref.best_hypothesis == -2
    a, ma = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(), ro,
                                       fr.bundle_options(), initial=fr.initial)
    fr_opt = synth.Frame(**{**fr.__dict__})
    fs = tr.FrameSet([fr_opt])
    fs.arr[0].ransac = ro
    f1, mf = ctx.estimate_frame(fs)
    b = tr.Batch(ctx, fs)
    b.launch()
    b.sync()
    rb, mb = b.results(masks=True)
    b.close()
    for r in (a, f1, rb[0]):
        assert r.best_hypothesis == -2
        assert abs(r.model_score - ref.model_score) <= 1e-12 * ref.model_score
        assert r.best_inliers == ref.best_inliers
        rot, cm = pose_diff(r, ref)
        assert rot <= 1e-6 and cm <= 1e-6
    srv = tr.Server(0, max_batch_frames=8, max_inflight=1)
    rs, _, _ = srv.wait(srv.submit(fs, 0))
    srv.close()
    assert list(a.q) == list(f1.q) == list(rb[0].q) == list(rs.q)
    assert np.array_equal(ma, ref_mask) and np.array_equal(mb[0], ma)
    # a far initial pose changes nothing
    far = (np.array([0.0, 1.0, 0.0, 0.0]), np.array([5.0, 5.0, -3.0]))
    c, _ = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(), ro,
                                      fr.bundle_options(), initial=far)
    d, _ = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(), fr.ransac_options(),
                                      fr.bundle_options())
    assert c.best_hypothesis == d.best_hypothesis >= 0 and list(c.q) == list(d.q)
