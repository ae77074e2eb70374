"""Continuous-batching server: every result is the same alone and in the mix,
the HBM budget bounds what is admitted, and the served results equal the
per-frame path's."""
import numpy as np
import pytest

import turboransac_amd as tr
from turboransac_amd import _abi, synth

pytestmark = pytest.mark.gpu


def same(a, b):
    return (a.best_hypothesis == b.best_hypothesis and list(a.q) == list(b.q)
            and list(a.t) == list(b.t) and a.num_inliers == b.num_inliers)


def test_server_alone_equals_in_the_mix(ctx):
    frames = synth.serving_stream(seed=13, frames=48, hypotheses=512)
    fs = tr.FrameSet(frames)
    srv = tr.Server(0, hbm_budget_bytes=2 << 30, max_batch_frames=16, max_inflight=2,
                    return_masks=True)
    print("server: alone", flush=True)
    alone = []
    for i in range(0, 48, 6):
        t = srv.submit(fs, i)
        alone.append((i,) + srv.wait(t, frames[i].n)[:2])
    print("server: mix", flush=True)
    arr = synth.poisson_arrivals(4000.0, len(frames), seed=1)
    lat, wall = srv.replay(fs, arr)
    st = srv.stats()
    assert st["launches"] >= 3
    # replay returns latencies only; get the mixed results by resubmitting
    # all frames at once (one more mix, different batch boundaries)
    tickets = [srv.submit(fs, i) for i in range(len(frames))]
    mixed = [srv.wait(t, frames[i].n)[:2] for i, t in enumerate(tickets)]
    for i, r, m in alone:
        assert same(r, mixed[i][0]), i
        assert np.array_equal(m, mixed[i][1])
        one, m1 = ctx.estimate_absolute_pose(frames[i].points2D, frames[i].points3D,
                                             frames[i].camera(), frames[i].ransac_options(),
                                             frames[i].bundle_options())
        assert same(r, one), i
        assert np.array_equal(m, m1)
    assert np.all(lat > 0) and wall > 0
    srv.close()


def test_server_budget_admission():
    frames = synth.serving_stream(seed=17, frames=24, hypotheses=512)
    fs = tr.FrameSet(frames)
    # budget for a few frames at a time
    budget = 64 << 20
    srv = tr.Server(0, hbm_budget_bytes=budget, max_batch_frames=64, max_inflight=2)
    tickets = []
    for i in range(len(frames)):
        try:
            tickets.append(srv.submit(fs, i))
        except tr.TRError as e:
            assert e.status == tr._abi.TR_ERR_CAPACITY
    for t in tickets:
        res, _, lat = srv.wait(t)
        assert res.status == 0 and lat > 0
    st = srv.stats()
    assert st["peak_bytes"] <= budget
    assert st["launches"] >= 2
    srv.close()
    # a frame bigger than the whole budget is refused, not queued
    small = tr.Server(0, hbm_budget_bytes=8 << 20, max_batch_frames=8, max_inflight=1)
    big = synth.make_frame(np.random.default_rng(3), 60000, 0.5, 1.0, hypotheses=16384)
    with pytest.raises(tr.TRError) as e:
        small.submit(tr.FrameSet([big]), 0)
    assert e.value.status == tr._abi.TR_ERR_CAPACITY
    small.close()


def test_server_threads_timeouts_and_refusals(ctx):
    """Submit is thread-safe: threads submitting and waiting at once get the
    per-frame path's results.  A wait with a timeout returns NOT_READY while
    the frame runs; an unknown ticket and a frame whose stop rule may ask for
    more than 32 rounds are refused with INVALID_ARGUMENT."""
    import threading

    frames = synth.serving_stream(seed=29, frames=32, hypotheses=512)
    fs = tr.FrameSet(frames)
    srv = tr.Server(0, hbm_budget_bytes=1 << 30, max_batch_frames=8, max_inflight=2,
                    return_masks=True)
    out = [None] * len(frames)
    errors = []

    def worker(k):
        try:
            for i in range(k, len(frames), 4):
                t = srv.submit(fs, i)
                out[i] = srv.wait(t, frames[i].n)[:2]
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    print("server: 4 threads x 8 frames", flush=True)
    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not errors, errors
    for i, fr in enumerate(frames):
        one, m1 = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(),
                                             fr.ransac_options(), fr.bundle_options())
        assert same(out[i][0], one), i
        assert np.array_equal(out[i][1], m1), i

    # A frame of 50 000 points with 16384 samples is still running after 0 ms.
    big = synth.make_frame(np.random.default_rng(31), 50000, 0.3, 1.0, hypotheses=16384)
    bs = tr.FrameSet([big])
    t = srv.submit(bs, 0)
    with pytest.raises(tr.TRError) as e:
        srv.wait(t, big.n, timeout_ms=0)
    assert e.value.status == tr._abi.TR_ERR_NOT_READY
    res, mask, lat = srv.wait(t, big.n)
    assert res.status == 0 and mask.sum() == res.num_inliers and lat > 0

    with pytest.raises(tr.TRError) as e:
        srv.wait(123456789)
    assert e.value.status == tr._abi.TR_ERR_INVALID_ARGUMENT

    long_run = tr.FrameSet([frames[0]])
    ro = long_run.arr[0].ransac
    ro.adaptive_stopping = 1
    ro.max_iterations = 64
    ro.max_total_iterations = 64 * 40
    with pytest.raises(tr.TRError) as e:
        srv.submit(long_run, 0)
    assert e.value.status == tr._abi.TR_ERR_INVALID_ARGUMENT
    ro.max_total_iterations = 64 * 32  # 32 rounds fit
    t = srv.submit(long_run, 0)
    res, _, _ = srv.wait(t, frames[0].n)
    assert res.status == 0 and res.iterations <= 64 * 32
    srv.close()


def test_device_resident_frames_equal_host_frames(ctx):
    """Frames submitted with their correspondences in device memory
    (tr_server_submit_device, read in place) give the same bits as the same
    frames submitted from host memory, in the same launches and in a
    device-resident replay; weighted frames included."""
    frames = synth.serving_stream(seed=43, frames=36, hypotheses=512)
    rng = np.random.default_rng(44)
    for i in (3, 17):
        frames[i].weights = rng.uniform(0.5, 2.0, frames[i].n)
    fs = tr.FrameSet(frames)
    dfs = tr.DeviceFrameSet(frames)
    srv = tr.Server(0, hbm_budget_bytes=1 << 30, max_batch_frames=16, max_inflight=2,
                    return_masks=True)
    print("server: host and device submissions interleaved", flush=True)
    th, td = [], []
    for i in range(len(frames)):
        th.append(srv.submit(fs, i))
        td.append(srv.submit_device(dfs, i))
    host = [srv.wait(t, frames[i].n)[:2] for i, t in enumerate(th)]
    dev = [srv.wait(t, frames[i].n)[:2] for i, t in enumerate(td)]
    for i in range(len(frames)):
        assert host[i][0].status == 0 and dev[i][0].status == 0, i
        assert same(host[i][0], dev[i][0]), i
        assert np.array_equal(host[i][1], dev[i][1]), i
    for i in (0, 5, 11):
        fr = frames[i]
        one, m1 = ctx.estimate_absolute_pose(fr.points2D, fr.points3D, fr.camera(),
                                             fr.ransac_options(), fr.bundle_options())
        assert same(dev[i][0], one), i
        assert np.array_equal(dev[i][1], m1), i
    print("server: device-resident replay", flush=True)
    arr = synth.poisson_arrivals(6000.0, len(frames), seed=4)
    lat, wall, res, masks = srv.replay(dfs, arr, results=True)
    assert np.all(lat > 0) and wall > 0
    for i in range(len(frames)):
        assert same(res[i], host[i][0]), i
        assert np.array_equal(masks[i], host[i][1]), i
    srv.close()
    dfs.close()


def test_host_submit_refuses_non_finite_inputs(ctx):
    """tr_server_submit checks host arrays for finite values while it copies
    them: a NaN or an infinity in points2D, points3D or the weights is
    refused with TR_ERR_NOT_FINITE at submission, wherever it sits, and the
    server then serves a good frame as before."""
    import dataclasses

    rng = np.random.default_rng(31)
    base = synth.make_frame(rng, 700, 0.6, 0.5, hypotheses=256, seed=31)
    base = dataclasses.replace(base, weights=np.linspace(0.5, 1.5, base.n))
    bad = []
    for field, idx in (("points2D", (0, 0)), ("points2D", (699, 1)), ("points3D", (350, 2)),
                       ("weights", (0,)), ("weights", (699,))):
        for v in (np.nan, np.inf, -np.inf):
            a = getattr(base, field).copy()
            a[idx] = v
            bad.append(dataclasses.replace(base, **{field: a}))
    fs = tr.FrameSet(bad + [base])
    srv = tr.Server(0, hbm_budget_bytes=1 << 30, max_batch_frames=8, max_inflight=2,
                    return_masks=True)
    try:
        print(f"server: {len(bad)} frames with a non-finite value", flush=True)
        for i in range(len(bad)):
            with pytest.raises(tr.TRError) as e:
                srv.submit(fs, i)
            assert e.value.status == _abi.TR_ERR_NOT_FINITE, i
        t = srv.submit(fs, len(bad))
        res, mask, _ = srv.wait(t, base.n)
    finally:
        srv.close()
    one, m1 = ctx.estimate_frame(base)
    assert same(res, one)
    assert np.array_equal(mask, m1)
