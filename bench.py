"""Benchmark: poses/s of the batched engine vs the one-frame-per-call path.

Workload (one step): the mixed batch of 256 frames -- N log-uniform in
[500, 20000], inlier ratio U[0.1, 0.8], sigma U[0.5, 2] px, fx, fy, cx, cy
each scaled by U[0.9, 1.1], 2048 minimal samples per frame, 8 px MSAC
threshold, LM <= 25 iterations -- with the correspondences already in HBM.
A step is one launch of the captured graph over the 256 frames.  The K
timed steps go round --inflight batches (default 2) of those frames, each
on a stream of its own, so one batch's refinement (latency-bound) runs
under the next one's scoring; every step still completes inside the timed
region.  The one-batch-at-a-time rate and its launch-to-result latency are
reported beside it.

Printed (one JSON line): batched poses/s (value), the per-frame path's
poses/s over the same resident frames and its per-call p50/p99 latency,
served runs (Poisson arrivals at 500, 2000 and 8000 frames/s, 10 s each)
with p50/p99, pose error of both paths against the ground truth and the
FP64 oracle, the MSAC kernel's roofline (algorithmic FP32 FLOP / kernel
time from HIP events, HBM traffic from the calibrated rocprofv3 passes) and
the CPU oracle timed on a bounded sample of the same frames.

    python bench.py [--steps K] [--warmup W] [--frames F] [--serve-rates 500,2000,8000]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_RESIDUAL = 28  # 9 FMA projection, 1/z, 2 FMA residual, r^2, min + add
# MI355X FP32 peak: 256 CUs x 4 SIMDs x 32 lanes x 2 FLOP x 2.4 GHz.  On
# gfx950 a wave64 FP32 FMA whose operands are all VGPRs issues in about 2.3
# clocks (measured), so the figure applies to plain (not packed) FP32.
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
# FP64 FLOP of one inlier in one LM pass (Cauchy path, FMA = 2): camera
# transform 18, reciprocal depth and the two residuals 12, the Cauchy weight 5,
# projection derivatives 24, the cross-product Jacobian 18, the 21 J^T J and
# 6 J^T r double-FMAs 108, the weighted rows 12.
LM_FLOP_PER_INLIER_PASS = 197
PEAK_HBM_GBS = 8000.0
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 2  # wave64 FP32 VALU instructions/s over the chip
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_msac_traffic_mixed_batch_256.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pose_err(res, fr):
    from turboransac_amd import synth

    R = synth.quat_to_rotmat(res.q)
    return synth.pose_error(R, np.array(res.t[:]), fr.R, fr.t)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=256)
    ap.add_argument("--inflight", type=int, default=2,
                    help="batches in flight during the timed steps, each on its own stream")
    ap.add_argument("--samples", type=int, default=2048)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--serve-rates", default="500,2000,8000",
                    help="Poisson arrival rates (frames/s) of the served runs, comma separated")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="length of each served run")
    ap.add_argument("--serve-pool", type=int, default=2000,
                    help="distinct frames the served runs cycle through")
    ap.add_argument("--serve-slots", type=int, default=2,
                    help="launches the server keeps in flight (tr_server_options.max_inflight)")
    ap.add_argument("--serve-sat-frames", type=int, default=20000,
                    help="frames of the saturation run (all queued at once)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the frame-parallel CPU baseline (the box's CPU share)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-large", dest="large", action="store_false",
                    help="skip the N=200000 single-frame section")
    ap.add_argument("--no-fundamental", dest="fundamental", action="store_false",
                    help="skip the fundamental-matrix frames")
    ap.add_argument("--no-sweep", dest="sweep", action="store_false",
                    help="skip the one-frame-per-call sweep over N of the reference paper")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine scales out by frame, one server "
              "per GPU with nothing exchanged between them; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import turboransac_amd as tr
    from turboransac_amd import synth

    t0 = time.time()
    frames = synth.mixed_batch(seed=2, frames=args.frames, hypotheses=args.samples)
    log(f"workload generated in {time.time() - t0:.1f}s: {len(frames)} frames, "
        f"sum N = {sum(f.n for f in frames)}")

    ctx = tr.Context(0)
    batch = tr.Batch(ctx, frames)

    # ---- batched: warmup, then exactly K timed graph launches, going round
    # `inflight` batches of the same frames (each on its own stream) ----
    batches = [batch] + [tr.Batch(ctx, frames) for _ in range(max(1, args.inflight) - 1)]
    for i in range(args.warmup):
        batches[i % len(batches)].launch()
    for b in batches:
        b.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        batches[i % len(batches)].launch()
    for b in batches:
        b.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = len(frames) * args.steps / elapsed
    log(f"batched, {len(batches)} in flight: {ms_per_step:.3f} ms/step, {value:.0f} poses/s")
    res_b, _ = batch.results()
    for b in batches[1:]:
        r, _ = b.results()
        if any(x.q[:] != y.q[:] or x.t[:] != y.t[:] for x, y in zip(r, res_b)):
            raise RuntimeError("batches in flight disagree on the same frames")
        b.close()
    # One batch at a time: launch-to-result latency and its rate.
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        batch.launch()
        batch.sync()
        lat.append(time.perf_counter() - s)
    serial_value = len(frames) * args.steps / (time.perf_counter() - t0)
    log(f"batched, one at a time: {1e3 * np.median(lat):.3f} ms p50, {serial_value:.0f} poses/s")
    work = batch.work()

    # ---- roofline of the MSAC kernel: HIP events around it, same stream ----
    score_ms = []
    stages = {}
    for _ in range(args.profile_steps):
        batch.launch_profiled()
        batch.sync()
        st = batch.stage_times()
        score_ms.append(st["msac_score"])
        for k, v in st.items():
            stages[k] = stages.get(k, 0.0) + v / args.profile_steps
    t_score = float(np.mean(score_ms)) * 1e-3
    residuals = work["residuals"]
    achieved = FLOP_PER_RESIDUAL * residuals / t_score / 1e12
    log(f"msac_score: {t_score * 1e3:.3f} ms avg, {residuals:.3e} residuals, "
        f"{achieved:.2f} TFLOP/s; stages {json.dumps({k: round(v, 4) for k, v in stages.items()})}")
    # Algorithmic HBM bytes of one MSAC launch: each frame's FP32 points once
    # (20 B), each pose once (48 B) and one partial score per pose per chunk.
    algo_bytes = sum(20 * f.n for f in frames) + 52 * work["candidates"]
    shape = next(l.split("#")[0].split() for l in ctx.get_tuning().splitlines()
                 if l.startswith("batch"))
    engine, J, T, W, chunk = shape[2], int(shape[3]), int(shape[4]), int(shape[5]), int(shape[6])
    kernel = {"mfma": "k_msac_mfma", "gated": "k_msac_gated",
              "gated2": "k_msac_gated2"}.get(engine, "k_msac") + f"<{J}, {T}, {W}>"
    # HBM bytes per launch of the same kernel on the same workload, from the
    # two calibrated rocprofv3 --pmc passes of turboransac_amd.roofline.
    # The third pass counts the kernel's executed wave64 instructions; with
    # this run's HIP-event time they give the VALU issue rate against the
    # 1024 SIMDs x 2.4 GHz / 2 clocks per FP32 wave64 instruction.
    traffic = None
    valu = None
    try:
        with open(TRAFFIC_FILE) as f:
            tf = json.load(f)
        if tf.get("kernel") and kernel in tf["kernel"] and args.frames == 256 and args.samples == 2048:
            traffic = tf["traffic_bytes"]
            v = tf.get("valu")
            if v and abs(v["residuals"] - residuals) <= 1e-9 * residuals:
                n_valu = v["per_launch"]["SQ_INSTS_VALU"]
                valu = {
                    "valu_insts_per_launch": n_valu,
                    "valu_insts_per_64_residuals": v["per_64_residuals"]["SQ_INSTS_VALU"],
                    "salu_insts_per_64_residuals": v["per_64_residuals"]["SQ_INSTS_SALU"],
                    "issue_rate": n_valu / t_score / VALU_ISSUE_PEAK,
                    "fp32_ops_per_residual": v["fp32_ops_per_residual"],
                    "fp32_executed_tflops": v["fp32_ops_per_launch"] / t_score / 1e12,
                }
                iss = tf.get("issue")
                if iss:
                    # The issue pass: VALU instructions per SIMD quad-cycle
                    # (at most 2), the share issued in pairs, and the
                    # shader clock the counters saw during the kernel.
                    valu.update({
                        "valu_per_simd_quad": iss["valu_per_simd_quad"],
                        "paired_share": iss["paired"],
                        "wave_residency": iss["residency"],
                        "clock_ghz_counted": iss["clock_ghz"],
                        "max_clock_ghz": iss["max_clock_ghz"],
                    })
    except (OSError, ValueError, KeyError):
        pass

    # ---- LM (FP64) of the same profiled launches ----
    res_lm, _ = batch.results()
    lm_passes = sum((int(r.refinements) + 1) * int(r.num_inliers) for r in res_lm)
    t_lm = stages["lm"] * 1e-3
    lm_achieved = LM_FLOP_PER_INLIER_PASS * lm_passes / t_lm / 1e12
    lm = {
        "stage": "inlier compaction and every LM pass of the batch launch (HIP events)",
        "stage_ms": stages["lm"],
        "inlier_passes": lm_passes,
        "flop_per_inlier_pass": LM_FLOP_PER_INLIER_PASS,
        "achieved": lm_achieved,
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": lm_achieved / PEAK_FP64_TFLOPS,
        "bound": "latency: each pass waits for the previous step (folded chunk sums, "
                 "6x6 Cholesky, pose update)",
    }
    try:
        with open(TRAFFIC_FILE) as f:
            l64 = json.load(f).get("lm_fp64")
        if l64 and args.frames == 256 and args.samples == 2048 and \
                int(l64["inlier_passes"]) == lm_passes:
            launches = l64["dispatches_per_batch_launch"]
            lm.update({
                "launches": launches,
                "us_per_launch": stages["lm"] * 1e3 / launches,
                "executed_fp64_ops": l64["fp64_ops_per_batch_launch"],
                "executed_fp64_ops_per_inlier_pass": l64["fp64_ops_per_inlier_pass"],
                "executed_tflops": l64["fp64_ops_per_batch_launch"] / t_lm / 1e12,
                "counters_source": os.path.relpath(TRAFFIC_FILE, ROOT),
            })
    except (OSError, ValueError, KeyError):
        pass
    log(f"lm: {stages['lm']:.3f} ms, {lm_passes:.4g} inlier passes, {lm_achieved:.3f} TFLOP/s FP64")

    # ---- plain: one frame per call over the same resident frames ----
    for i in range(min(3, len(frames))):
        d2, d3, n = batch.frame_inputs(i)
        fr = frames[i]
        ctx.estimate_absolute_pose_device(d2, d3, n, fr.camera(), fr.ransac_options(),
                                          fr.bundle_options())
    plain_lat = []
    res_p = []
    t0 = time.perf_counter()
    for i, fr in enumerate(frames):
        d2, d3, n = batch.frame_inputs(i)
        s = time.perf_counter()
        res_p.append(ctx.estimate_absolute_pose_device(d2, d3, n, fr.camera(),
                                                       fr.ransac_options(),
                                                       fr.bundle_options()))
        plain_lat.append(time.perf_counter() - s)
    plain_elapsed = time.perf_counter() - t0
    plain_value = len(frames) / plain_elapsed
    log(f"plain: {plain_value:.0f} poses/s, p50 {1e3 * np.median(plain_lat):.3f} ms")
    same = sum(int(a.best_hypothesis == b.best_hypothesis and list(a.q) == list(b.q))
               for a, b in zip(res_b, res_p))

    err_b = np.array([pose_err(r, f) for r, f in zip(res_b, frames)])
    err_p = np.array([pose_err(r, f) for r, f in zip(res_p, frames)])

    # ---- served: Poisson arrivals, continuous batching, one run per rate ----
    # The server's slots each want a hardware queue of their own; HIP
    # multiplexes streams beyond its queue count (GPU_MAX_HW_QUEUES, 4 by
    # default) onto shared queues, and a slot that lands on one was measured
    # 1.4-1.8x slower.  The batch (and its stream) is closed first, as in a
    # process that only serves.
    batch.close()
    serve = []
    serve_sat = None
    serve_sat_dev = None
    rates = [float(r) for r in args.serve_rates.split(",") if r.strip()]
    if rates and args.serve_seconds > 0:
        pool_frames = synth.serving_stream(seed=3, frames=args.serve_pool,
                                           hypotheses=args.samples)
        pool = tr.FrameSet(pool_frames)
        srv = tr.Server(0, hbm_budget_bytes=8 << 30, max_batch_frames=256,
                        max_inflight=args.serve_slots)
        warm = pool.repeat(64)
        srv.replay(warm, np.zeros(64))

        def served_errors(results, count):
            # pose error against the truth of every 8th served frame
            idx = range(0, count, 8)
            e = np.array([pose_err(results[i], pool_frames[i % len(pool_frames)]) for i in idx])
            return {"frames": len(idx), "median_rot_deg": float(np.median(e[:, 0])),
                    "median_center_cm": float(np.median(e[:, 1]))}

        for rate in rates:
            count = max(1, int(rate * args.serve_seconds))
            arr = synth.poisson_arrivals(rate, count)
            before = srv.stats()
            lat_s, wall, sres, _ = srv.replay(pool.repeat(count), arr, results=True)
            after = srv.stats()
            launches = after["launches"] - before["launches"]
            run = {
                "arrival_rate_hz": rate,
                "frames": count,
                "seconds": float(arr[-1]),
                "poses_per_s": count / wall,
                "p50_ms": float(np.percentile(lat_s, 50) / 1e3),
                "p99_ms": float(np.percentile(lat_s, 99) / 1e3),
                "launches": launches,
                "frames_per_launch": count / max(1, launches),
                "pose_error_vs_truth": served_errors(sres, count),
            }
            serve.append(run)
            log(f"served at {rate:.0f}/s for {run['seconds']:.1f} s: {run['poses_per_s']:.0f} "
                f"poses/s, p50 {run['p50_ms']:.3f} ms, p99 {run['p99_ms']:.3f} ms, "
                f"{run['frames_per_launch']:.1f} frames/launch")
        # Saturation: every frame arrives at once (closed loop with the whole
        # trace queued), so the server runs full batches back to back; the
        # latency then counts the queueing.
        count = max(1, int(args.serve_sat_frames))
        before = srv.stats()
        lat_s, wall, sres, _ = srv.replay(pool.repeat(count), np.zeros(count), results=True)
        serve_sat_res = sres
        after = srv.stats()
        launches = after["launches"] - before["launches"]
        serve_sat = {
            "frames": count,
            "poses_per_s": count / wall,
            "p50_ms": float(np.percentile(lat_s, 50) / 1e3),
            "p99_ms": float(np.percentile(lat_s, 99) / 1e3),
            "launches": launches,
            "frames_per_launch": count / max(1, launches),
            "pose_error_vs_truth": served_errors(sres, count),
        }
        log(f"served at saturation: {serve_sat['poses_per_s']:.0f} poses/s over {count} frames, "
            f"{serve_sat['frames_per_launch']:.1f} frames/launch")
        # The same saturation run with the pool's correspondences already in
        # HBM (tr_server_submit_device: the server reads them in place, no
        # host copy), as when an earlier GPU stage produced them.
        dpool = tr.DeviceFrameSet(pool_frames)
        srv.replay(dpool.repeat(64), np.zeros(64))
        before = srv.stats()
        lat_s, wall, sres, _ = srv.replay(dpool.repeat(count), np.zeros(count), results=True)
        after = srv.stats()
        launches = after["launches"] - before["launches"]
        same_dev = sum(int(list(a.q) == list(b.q) and a.num_inliers == b.num_inliers)
                       for a, b in zip(sres[:len(pool_frames)], serve_sat_res[:len(pool_frames)]))
        serve_sat_dev = {
            "frames": count,
            "inputs": "device-resident (tr_server_submit_device), uploaded once before the run",
            "poses_per_s": count / wall,
            "p50_ms": float(np.percentile(lat_s, 50) / 1e3),
            "p99_ms": float(np.percentile(lat_s, 99) / 1e3),
            "launches": launches,
            "frames_per_launch": count / max(1, launches),
            "pose_error_vs_truth": served_errors(sres, count),
            "same_result_as_host_inputs": same_dev,
            "compared": min(count, len(pool_frames)),
        }
        log(f"served at saturation, device-resident inputs: {serve_sat_dev['poses_per_s']:.0f} "
            f"poses/s, {serve_sat_dev['frames_per_launch']:.1f} frames/launch, "
            f"{same_dev} of {serve_sat_dev['compared']} equal to the host-input results")
        srv.close()
        dpool.close()
        # The plain path on the same pool frames, for the served pose error
        # comparison ("at equal pose error").
        pb = tr.Batch(ctx, pool_frames[:256])
        perr = []
        for i, fr in enumerate(pool_frames[:256:8]):
            d2, d3, n = pb.frame_inputs(8 * i)
            perr.append(pose_err(ctx.estimate_absolute_pose_device(
                d2, d3, n, fr.camera(), fr.ransac_options(), fr.bundle_options()), fr))
        pb.close()
        perr = np.array(perr)
        serve_sat["plain_path_same_pool"] = {
            "frames": len(perr), "median_rot_deg": float(np.median(perr[:, 0])),
            "median_center_cm": float(np.median(perr[:, 1]))}

    # ---- large frame: N=200000, 5% inliers, 16384 samples (scoring roofline) ----
    large = None
    if args.large:
        lf = synth.large_frame()
        lb = tr.Batch(ctx, [lf])
        lb.launch()
        lb.sync()
        t_ms = []
        lstages = {}
        for _ in range(args.profile_steps):
            lb.launch_profiled()
            lb.sync()
            st = lb.stage_times()
            t_ms.append(sum(st.values()))
            for k, v in st.items():
                lstages[k] = lstages.get(k, 0.0) + v / args.profile_steps
        lres, _ = lb.results()
        lw = lb.work()
        d2, d3, n = lb.frame_inputs(0)
        # The first call grows the context's device arena to this frame's
        # size; the median of the calls after it is reported.
        lp = ctx.estimate_absolute_pose_device(d2, d3, n, lf.camera(), lf.ransac_options(),
                                               lf.bundle_options())
        pl = []
        for _ in range(5):
            s = time.perf_counter()
            lp = ctx.estimate_absolute_pose_device(d2, d3, n, lf.camera(), lf.ransac_options(),
                                                   lf.bundle_options())
            pl.append((time.perf_counter() - s) * 1e3)
        plain_large_ms = float(np.median(pl))
        ls = lstages["msac_score"] * 1e-3
        lach = FLOP_PER_RESIDUAL * lw["residuals"] / ls / 1e12
        lerr = pose_err(lres[0], lf)
        large = {
            "workload": "one frame N=200000, 5% inliers, sigma 1 px, 16384 samples, tau 8 px",
            "pipeline_ms": float(np.mean(t_ms)),
            "plain_call_ms": plain_large_ms,
            "msac_ms": lstages["msac_score"],
            "residuals": lw["residuals"],
            "candidates": lw["candidates"],
            "msac_tflops": lach,
            "msac_frac_of_157": lach / PEAK_FP32_TFLOPS,
            "stage_ms": {k: round(v, 5) for k, v in lstages.items()},
            "inliers": int(lres[0].num_inliers),
            "rot_err_deg": lerr[0],
            "center_err_cm": lerr[1],
            "plain_equals_batched": bool(list(lp.q) == list(lres[0].q)),
        }
        lb.close()
        log(f"large frame: pipeline {large['pipeline_ms']:.3f} ms, msac {large['msac_ms']:.3f} ms "
            f"({lach:.1f} TFLOP/s), inliers {large['inliers']}, err {lerr}")

    # ---- the reference paper's runtime-vs-N sweep (one frame per call) ----
    # Published (RTX 4090, kernels only, digitised from the paper's Fig. 4):
    # per-frame ms at each N.  Here: the drop-in call on device-resident
    # inputs, median of repeated calls, results included.
    sweep = None
    if args.sweep:
        published = {1000: 2.78, 5000: 5.04, 10000: 8.53, 50000: 32.0, 100000: 59.8,
                     200000: 116.6}
        pts = []
        for n in synth.PAPER_SWEEP_N:
            fr = synth.paper_sweep_frame(n)
            holder = tr.Batch(ctx, [fr])
            d2, d3, nn = holder.frame_inputs(0)
            ro, bo, cam = synth.paper_sweep_ransac_options(fr), fr.bundle_options(), fr.camera()
            for _ in range(2):
                r = ctx.estimate_absolute_pose_device(d2, d3, nn, cam, ro, bo)
            ts = []
            for _ in range(10):
                s = time.perf_counter()
                r = ctx.estimate_absolute_pose_device(d2, d3, nn, cam, ro, bo)
                ts.append((time.perf_counter() - s) * 1e3)
            holder.close()
            e = pose_err(r, fr)
            pts.append({"n": n, "ms": float(np.median(ts)), "published_rtx4090_ms": published[n],
                        "samples": int(r.iterations), "inliers": int(r.num_inliers),
                        "rot_err_deg": e[0], "center_err_cm": e[1]})
            log(f"paper sweep N={n}: {pts[-1]['ms']:.3f} ms per call "
                f"(published RTX 4090: {published[n]} ms), {r.iterations} samples")
        sweep = {"workload": "one frame per call, device-resident inputs; N in "
                             f"{list(synth.PAPER_SWEEP_N)}, 50% inliers, sigma 0.5 px, tau 1.2 px, "
                             "3000 samples per round, the reference's stop rule "
                             "(confidence 0.9999, x3); LM <= 25",
                 "published": "RTX 4090, kernels only, digitised from the reference paper's Fig. 4 "
                              "(its synthetic data are not published)",
                 "points": pts}

    # ---- fundamental matrix: 2D-2D frames through the drop-in call ----
    # Host inputs in, F and the mask out (H2D of 32 B per point and D2H of
    # the result and mask inside each call); 1024 7-point samples per frame.
    fund = None
    pfs = synth.pair_stream(seed=8, frames=64, hypotheses=1024) if args.fundamental else []
    fund_res = []
    if args.fundamental:
        for fr in pfs[:3]:  # the first calls grow the context's arena
            ctx.estimate_fundamental(fr.points1, fr.points2, fr.ransac_options(),
                                     fr.bundle_options())
        flat, fresid, ferr = [], 0, []
        t0 = time.perf_counter()
        for fr in pfs:
            s = time.perf_counter()
            r, _ = ctx.estimate_fundamental(fr.points1, fr.points2, fr.ransac_options(),
                                            fr.bundle_options())
            flat.append((time.perf_counter() - s) * 1e3)
            fund_res.append(r)
            fresid += int(r.num_candidates) * fr.n
        fwall = time.perf_counter() - t0
        for fr, r in zip(pfs, fund_res):
            F = np.array(r.F[:]).reshape(3, 3)
            ferr.append(float(np.median(_sampson_px(F, fr.points1[fr.is_inlier],
                                                    fr.points2[fr.is_inlier]))))
        fund = {
            "workload": "64 two-view frames (synth.pair_stream seed 8): N log-uniform [200, "
                        "8000], inlier ratio U[0.3, 0.9], sigma U[0.3, 1] px, Sampson threshold "
                        "1.5 px, 1024 7-point samples, IRLS <= 20; host inputs, one call each",
            "frames_per_s": len(pfs) / fwall,
            "p50_ms": float(np.median(flat)),
            "p99_ms": float(np.quantile(flat, 0.99)),
            "sampson_residuals": fresid,
            "residuals_per_s": fresid / fwall,
            "median_true_inlier_sampson_px": float(np.median(ferr)),
        }
        log(f"fundamental: {fund['frames_per_s']:.0f} frames/s, p50 {fund['p50_ms']:.3f} ms, "
            f"median Sampson error of the true inliers {fund['median_true_inlier_sampson_px']:.3f} px")
        # The same 64 frames, and 4 copies of them, in one batched call each
        # (tr_estimate_fundamental_batch: host inputs staged and uploaded in
        # one copy, every kernel over all frames); every result checked
        # against the one-frame calls above.
        pairs = [(fr.points1, fr.points2, fr.ransac_options(), fr.bundle_options()) for fr in pfs]
        fb = {}
        for reps in (1, 4):
            batch_pairs = pairs * reps
            ctx.estimate_fundamental_batch(batch_pairs, masks=False)  # grows the arena
            times = []
            for _ in range(5):
                s = time.perf_counter()
                bres, _ = ctx.estimate_fundamental_batch(batch_pairs)
                times.append(time.perf_counter() - s)
            same_f = sum(int(list(a.F) == list(b.F) and a.num_inliers == b.num_inliers)
                         for a, b in zip(bres[:len(pfs)], fund_res))
            fb[str(len(batch_pairs))] = {
                "frames": len(batch_pairs),
                "ms_per_call": float(np.median(times) * 1e3),
                "frames_per_s": len(batch_pairs) / float(np.median(times)),
                "same_result_as_one_frame_calls": same_f,
            }
            log(f"fundamental batch of {len(batch_pairs)}: "
                f"{fb[str(len(batch_pairs))]['frames_per_s']:.0f} frames/s, "
                f"{fb[str(len(batch_pairs))]['ms_per_call']:.3f} ms per call, "
                f"{same_f}/{len(pfs)} equal to the one-frame calls")
        fund["batch"] = fb
        fund["batch_vs_one_frame_calls"] = fb["256"]["frames_per_s"] / fund["frames_per_s"]

    # ---- CPU baseline: the FP64 oracle on a bounded sample of the frames ----
    # One thread (frame by frame, also checked against the batched results),
    # then a frame-parallel pool of --cpu-threads threads.
    cpu = None
    oracle_diff = None
    fp32_err = None
    if not args.no_cpu_baseline:
        from oracle import load_oracle

        orc = load_oracle()
        done = 0
        diffs = []
        # The single-thread figure runs pinned to one core (this thread's
        # affinity), then the affinity is restored for the threaded pool.
        allowed = sorted(os.sched_getaffinity(0))
        pinned = allowed[0]
        os.sched_setaffinity(0, {pinned})
        t0 = time.perf_counter()
        for i, fr in enumerate(frames):
            st, ref, _, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(),
                                         fr.ransac_options(), fr.bundle_options(),
                                         want_mask=False)
            done += 1
            Rb = synth.quat_to_rotmat(res_b[i].q)
            Rr = synth.quat_to_rotmat(ref.q)
            diffs.append(synth.pose_error(Rb, np.array(res_b[i].t[:]), Rr, np.array(ref.t[:])))
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        t_one = time.perf_counter() - t0
        os.sched_setaffinity(0, set(allowed))
        one = done / t_one
        d = np.array(diffs)
        oracle_diff = {"frames": done, "max_rot_deg": float(d[:, 0].max()),
                       "max_center_cm": float(d[:, 1].max())}
        # FP32 MSAC scores of the batch kernel's shape against the oracle's
        # FP64 scores, every candidate of 16 frames (the per-frame debug
        # path of a context whose one-frame row is the batch row).
        fp32_err = None
        batch_row = next(l for l in ctx.get_tuning().splitlines() if l.startswith("batch"))
        cdbg = tr.Context(0)
        cdbg.set_tuning(batch_row.replace("batch", "single", 1).split("#")[0] + "\n" + batch_row)
        rel, rel_low = [], []
        for fr in frames[::16]:
            ro, bo = fr.ransac_options(), fr.bundle_options()
            _, _, dbg = cdbg.debug_estimate(fr.points2D, fr.points3D, fr.camera(), ro, bo,
                                            max_samples=fr.hypotheses)
            _, _, _, trc = orc.estimate(fr.points2D, fr.points3D, fr.camera(), ro, bo,
                                        trace_samples=fr.hypotheses, want_mask=False)
            s32 = dbg["scores32"].astype(np.float64)
            s64 = trc["scores"]
            ok = np.isfinite(s32) & np.isfinite(s64)
            tau2 = (fr.max_reproj_error / (0.5 * (fr.K[0] + fr.K[1]))) ** 2
            e = np.abs(s32[ok] - s64[ok]) / (fr.n * tau2)
            low = s64[ok] <= np.quantile(s64[ok], 0.01)
            rel.append(e)
            rel_low.append(e[low])
        cdbg.close()
        rel, rel_low = np.concatenate(rel), np.concatenate(rel_low)
        fp32_err = {"frames": len(frames[::16]), "candidates": int(rel.size),
                    "kernel_shape": batch_row.split("#")[0].split()[2:],
                    "unit": "|FP32 - FP64 score| / (N tau^2)",
                    "max": float(rel.max()), "p99_9": float(np.quantile(rel, 0.999)),
                    "median": float(np.median(rel)),
                    "max_lowest_1pct": float(rel_low.max())}
        log(f"FP32 MSAC error vs FP64 oracle ({rel.size} candidates): max {rel.max():.2e}, "
            f"lowest 1%: max {rel_low.max():.2e}, median {np.median(rel):.2e} (of N tau^2)")
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        fs_all = tr.FrameSet(frames)
        rounds, t0 = 0, time.perf_counter()
        while True:
            orc.estimate_many(list(fs_all.arr), threads=threads)
            rounds += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        t_mt = time.perf_counter() - t0
        cpu = {
            "value": rounds * len(frames) / t_mt,
            "unit": "poses/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{rounds} pass(es) over the {len(frames)} mixed-batch frames, FP64 scalar "
                      f"oracle, {threads} threads frame-parallel, {t_mt:.1f} s",
            "single_thread": {"value": one, "frames": done, "seconds": t_one,
                              "pinned_cpu": pinned},
            "cpu_model": _cpu_model(),
            "nproc": len(allowed),
            "os_cpu_count": os.cpu_count(),
        }
        log(f"cpu oracle: {cpu['value']:.1f} poses/s on {threads} threads, {one:.2f} poses/s on "
            f"one; max diff vs batched {oracle_diff}")
        if fund is not None:
            # The oracle on the same two-view frames, one thread, about
            # --cpu-seconds / 2 of them; its choices against the GPU's.
            done, same = 0, 0
            t0 = time.perf_counter()
            for fr, r in zip(pfs, fund_res):
                st, ref, _, _ = orc.estimate_fundamental(fr.points1, fr.points2,
                                                         fr.ransac_options(), fr.bundle_options())
                done += 1
                same += int(ref.best_hypothesis == r.best_hypothesis)
                if time.perf_counter() - t0 > args.cpu_seconds / 2:
                    break
            t_f = time.perf_counter() - t0
            fund["cpu_baseline"] = {"value": done / t_f, "unit": "frames/s", "cores": 1,
                                    "kind": "port", "sample": f"the first {done} frames, FP64 "
                                    f"scalar oracle, one thread, {t_f:.1f} s",
                                    "same_best_hypothesis": same}
            log(f"cpu oracle, fundamental: {done / t_f:.1f} frames/s on one thread; "
                f"same hypothesis as the GPU on {same}/{done}")

    out = {
        "metric": "poses/sec + p50/p99 latency, batched/served vs per-frame GPU, at equal pose error",
        "value": value,
        "unit": "poses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "mixed_batch: 256 frames, N log-uniform [500,20000], inlier ratio "
                        "U[0.1,0.8], sigma U[0.5,2] px, K +-10%, 2048 samples/frame, tau 8 px, "
                        "LM <= 25; inputs resident in HBM; one captured-graph launch per step, "
                        f"{len(batches)} batches in flight on their own streams",
            "frames": len(frames),
            "samples_per_frame": args.samples,
            "parallelism": "single GPU, frames batched in one launch",
            "inflight": len(batches),
        },
        "serial_poses_per_s": serial_value,
        "batch_latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                             "p99": float(np.percentile(lat, 99) * 1e3)},
        "plain_poses_per_s": plain_value,
        "plain_latency_ms": {"p50": float(np.percentile(plain_lat, 50) * 1e3),
                             "p99": float(np.percentile(plain_lat, 99) * 1e3)},
        "speedup_vs_plain": value / plain_value,
        "batched_equals_plain_frames": same,
        "pose_error_vs_truth": {
            "batched_median_rot_deg": float(np.median(err_b[:, 0])),
            "batched_median_center_cm": float(np.median(err_b[:, 1])),
            "plain_median_rot_deg": float(np.median(err_p[:, 0])),
            "plain_median_center_cm": float(np.median(err_p[:, 1])),
        },
        "vs_oracle": oracle_diff,
        "fp32_msac_error": fp32_err,
        "serve": serve,
        "serve_saturation": serve_sat,
        "served_saturation_vs_plain": (serve_sat["poses_per_s"] / plain_value
                                       if serve_sat else None),
        "serve_saturation_device": serve_sat_dev,
        "served_device_saturation_vs_plain": (serve_sat_dev["poses_per_s"] / plain_value
                                              if serve_sat_dev else None),
        "serve_workload": "frames cycle through a pool of 2000: N lognormal median 3000, sigma_ln "
                          "1.0, clipped [100,50000], inlier ratio U[0.05,0.8], sigma 1 px, "
                          "2048 samples; latency from the scheduled arrival to the result",
        "large_frame": large,
        "paper_fig4_sweep": sweep,
        "fundamental": fund,
        "stage_ms": {k: round(v, 5) for k, v in stages.items()},
        "roofline": {
            # the tuned batch kernel's arithmetic runs on the vector ALUs
            # (engine mfma alone puts the projection on the matrix cores)
            "bound": "mfma" if engine == "mfma" else "valu",
            "compute_unit": {"mfma": "FP32: projection on MFMA 16x16x4, residual on the VALU",
                             "gated": "FP32 VALU (points on lanes; the division only where "
                                      "a wavefront holds an inlier)",
                             "gated2": "FP32 VALU (points on lanes, two hypotheses per "
                                       "pass; the division only where a wavefront holds an "
                                       "inlier)"}.get(engine, "FP32 VALU") +
                            f"; peak {PEAK_FP32_TFLOPS} TF/s, FP32 on all VGPR operands",
            "kernel": kernel,
            "chunk_pts": chunk,
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            # One wave64 FP32 FMA per SIMD every 4 clocks (no dual issue):
            # 78.6 TF/s.  The kernel passes it by issuing two waves' VALU
            # operations in one quad-cycle (valu.valu_per_simd_quad > 1).
            "single_issue_peak": PEAK_FP32_TFLOPS / 2,
            "frac_of_single_issue": achieved / (PEAK_FP32_TFLOPS / 2),
            "traffic": traffic,
            "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if traffic else None,
            "valu": valu,
            "kernel_ms": t_score * 1e3,
            "residuals": residuals,
            "flop_per_residual": FLOP_PER_RESIDUAL,
            "algorithmic_hbm_bytes": algo_bytes,
            "algorithmic_hbm_gbs": algo_bytes / t_score / 1e9,
        },
        "lm_roofline": lm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def _sampson_px(F, p1, p2):
    """Sampson distance (pixels) of each pair under F."""
    h1 = np.column_stack([p1, np.ones(len(p1))])
    h2 = np.column_stack([p2, np.ones(len(p2))])
    a = h1 @ F.T
    b = h2 @ F
    num = np.einsum("ij,ij->i", h2, a)
    return np.abs(num) / np.sqrt(a[:, 0] ** 2 + a[:, 1] ** 2 + b[:, 0] ** 2 + b[:, 1] ** 2)


def _cpu_model():
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
