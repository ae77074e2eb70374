"""Autotuner of the MSAC kernel's shape.

The search space is every compiled (J hypotheses per lane, T points per LDS
tile, W wavefronts per workgroup) variant times the points per work item
(chunk).  For one frame per call the objective is the call's latency, per
N bucket; for batches it is the MSAC stage time of the 256-frame mixed
batch.  The winners are written as a table (the format of
tr_context_set_tuning) that every context loads from tuning/gfx950.txt.
A matrix-core (mfma) shape wins only when it is more than MFMA_MARGIN
faster than the best vector-ALU shape: the projection counts as a real
contraction only where the timing shows it.

    python -m turboransac_amd.autotune --out turboransac_amd/tuning/gfx950.txt
"""
from __future__ import annotations

import argparse
import datetime
import math
import sys
import time

import numpy as np

import turboransac_amd as tr
from turboransac_amd import synth

SINGLE_BUCKETS = [  # representative N, samples
    (1000, 1024),
    (4000, 2048),
    (16000, 2048),
    (64000, 4096),
    (200000, 16384),
]
CHUNKS = [512, 1024, 2048, 4096, 8192]
BATCH_CHUNKS = [2048, 3072, 4096, 6144, 8192]
MFMA_MARGIN = 0.01


def pick(rows):
    """rows: (ms, engine, J, T, W, chunk).  The fastest shape, except that an
    mfma shape must beat the fastest other shape by MFMA_MARGIN."""
    best = min(rows)
    if best[1] != "mfma":
        return best
    other = [r for r in rows if r[1] != "mfma"]
    if other and min(other)[0] <= best[0] * (1.0 + MFMA_MARGIN):
        return min(other)
    return best


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_single(ctx, batch, frames, reps):
    out = []
    for _ in range(reps):
        s = time.perf_counter()
        for i, fr in enumerate(frames):
            d2, d3, n = batch.frame_inputs(i)
            ctx.estimate_absolute_pose_device(d2, d3, n, fr.camera(), fr.ransac_options(),
                                              fr.bundle_options())
        out.append((time.perf_counter() - s) / len(frames))
    return float(np.median(out)) * 1e3


def tune_single(ctx, reps=5, frames_per_bucket=3):
    entries = []
    log_rows = []
    for bi, (n, H) in enumerate(SINGLE_BUCKETS):
        rng = np.random.default_rng(100 + bi)
        frames = [synth.make_frame(rng, n, 0.3, 1.0, hypotheses=H, seed=7 + k)
                  for k in range(frames_per_bucket)]
        holder = tr.Batch(ctx, frames)  # device-resident copies of the inputs
        cands = []
        for E, J, T, W in tr.score_variants():
            for chunk in CHUNKS:
                if chunk > max(512, 2 * n):
                    continue
                ctx.set_tuning(f"single -1 {E} {J} {T} {W} {chunk}\n"
                               f"batch -1 valu 2 1024 4 512\n")
                time_single(ctx, holder, frames[:1], 1)
                ms = time_single(ctx, holder, frames, reps)
                log_rows.append(("single", n, E, J, T, W, chunk, ms))
                cands.append((ms, E, J, T, W, chunk))
        holder.close()
        best = pick(cands)
        log(f"single N={n} H={H}: best {best}")
        if bi + 1 < len(SINGLE_BUCKETS):
            max_n = int(math.sqrt(n * SINGLE_BUCKETS[bi + 1][0]))
        else:
            max_n = -1
        entries.append(("single", max_n) + best[1:] + (best[0],))
    return entries, log_rows


def tune_batch(ctx, reps=5):
    frames = synth.mixed_batch(seed=2, frames=256, hypotheses=2048)
    fs = tr.FrameSet(frames)
    rows = []
    cands = []
    for E, J, T, W in tr.score_variants():
        for chunk in BATCH_CHUNKS:
            ctx.set_tuning(f"single -1 valu 1 1024 4 1024\nbatch -1 {E} {J} {T} {W} {chunk}\n")
            b = tr.Batch(ctx, fs)
            b.launch()
            b.sync()
            ms = []
            for _ in range(reps):
                b.launch_profiled()
                b.sync()
                ms.append(b.stage_times()["msac_score"])
            b.close()
            m = float(np.median(ms))
            rows.append(("batch", -1, E, J, T, W, chunk, m))
            cands.append((m, E, J, T, W, chunk))
    best = pick(cands)
    log(f"batch (256 mixed frames): best {best}")
    return [("batch", -1) + best[1:] + (best[0],)], rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--log", default=None, help="also write every measurement here")
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    ctx = tr.Context(0)
    single, rows1 = tune_single(ctx, args.reps)
    batch, rows2 = tune_batch(ctx, args.reps)
    stamp = datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%d %H:%M UTC")
    lines = [f"# MSAC kernel shapes, autotuned on gfx950 ({stamp}) by turboransac_amd.autotune",
             "# kind MAX_N ENGINE J T W CHUNK    # measured ms (single: per call; "
             "batch: MSAC stage)"]
    for kind, max_n, E, J, T, W, chunk, ms in single + batch:
        lines.append(f"{kind} {max_n} {E} {J} {T} {W} {chunk}    # {ms:.4f}")
    text = "\n".join(lines) + "\n"
    with open(args.out, "w") as f:
        f.write(text)
    print(text)
    if args.log:
        with open(args.log, "w") as f:
            f.write("kind,n,engine,J,T,W,chunk,ms\n")
            for r in rows1 + rows2:
                f.write(",".join(str(v) for v in r) + "\n")


if __name__ == "__main__":
    main()
