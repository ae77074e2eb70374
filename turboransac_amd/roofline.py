"""HBM traffic and executed VALU work of the MSAC kernel from rocprofv3 counters.

Runs the 256-frame mixed batch (the benchmark's workload) and a plain device
copy of known size under three `rocprofv3 --pmc` passes, each its own run:
FETCH_SIZE, then WRITE_SIZE (KiB per dispatch), then the SQ instruction
counters (VALU, SALU and the FP32 VALU classes).  Each traffic counter is
scaled by what it reads for the copy (bytes moved / counter), which
calibrates the counter on this device, and the scaled values give the MSAC
kernel's HBM bytes per launch.  The instruction pass gives the wave64
instructions one launch executes, per 64 residuals, and the FP32 operations
they perform (FMA counted twice); bench.py divides them by its own HIP-event
time of the kernel to report the VALU issue rate.  The summary lands in a
JSON file that bench.py reports as roofline.traffic and roofline.valu, and
in a markdown table next to it.

    python -m turboransac_amd.roofline --out profiles/k_msac_traffic_mixed_batch_256.json

The parent process only starts rocprofv3; the GPU work runs in the child
(`--child`), started by rocprofv3 itself.
"""
from __future__ import annotations

import argparse
import datetime
import glob
import json
import os
import sqlite3
import subprocess
import sys
import tempfile

COPY_BYTES = 256 << 20  # one direction of the calibration copy
SCORE_KERNELS = ("k_msac<", "k_msac_mfma<", "k_msac_gated<")
# One pass: 7 SQ counters, within the 8 one SQ pass can hold.
INST_COUNTERS = ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VALU_FMA_F32",
                 "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32",
                 "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_VALU_INT32")


def child(reps: int, work_out: str | None) -> None:
    import torch

    import turboransac_amd as tr
    from turboransac_amd import synth

    # The calibration copy: an elementwise kernel reading and writing
    # COPY_BYTES each (not a DMA copy, which the counters would not see).
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    ctx = tr.Context(0)
    frames = synth.mixed_batch(seed=2, frames=256, hypotheses=2048)
    b = tr.Batch(ctx, frames)
    for _ in range(reps):
        b.launch()
        b.sync()
    print("batch shape:", [l for l in ctx.get_tuning().splitlines() if l.startswith("batch")],
          flush=True)
    if work_out:
        b.results()  # the work counts are those of the last results() call
        with open(work_out, "w") as f:
            json.dump(b.work(), f)
    b.close()
    ctx.close()


def _counter_rows(db_path: str, counter: str):
    db = sqlite3.connect(db_path)
    rows = db.execute(
        "select dispatch_id, kernel_name, grid_size, value from counters_collection "
        "where counter_name = ?", (counter,)).fetchall()
    per = {}
    for did, name, grid, val in rows:
        key = (did, name, grid)
        per[key] = per.get(key, 0.0) + float(val)
    return [(name, grid, v) for (did, name, grid), v in sorted(per.items())]


def _run_pass(counters, outdir: str, reps: int, work_out: str | None = None) -> str:
    cmd = ["rocprofv3", "--pmc", *counters, "-d", outdir, "-o", "run", "--",
           sys.executable, "-m", "turboransac_amd.roofline", "--child", "--reps", str(reps)]
    if work_out:
        cmd += ["--work-out", work_out]
    env = dict(os.environ)
    env["PYTHONPATH"] = os.path.dirname(os.path.dirname(os.path.abspath(__file__))) + (
        ":" + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, env=env, timeout=240)
    dbs = glob.glob(os.path.join(outdir, "**", "*.db"), recursive=True)
    if not dbs:
        raise RuntimeError(f"no rocprofv3 database under {outdir}")
    return dbs[0]


def summarise(fetch_rows, write_rows):
    def pick(rows, pred):
        vals = [v for name, grid, v in rows if pred(name, grid)]
        return sum(vals) / len(vals) if vals else None, len(vals)

    is_copy = lambda name, grid: "copy" in name.lower() or "elementwise" in name.lower()
    is_score = lambda name, grid: any(k in name for k in SCORE_KERNELS)
    copy_fetch, _ = pick(fetch_rows, is_copy)
    copy_write, _ = pick(write_rows, is_copy)
    score_fetch, nf = pick(fetch_rows, is_score)
    score_write, nw = pick(write_rows, is_score)
    kib = COPY_BYTES / 1024.0
    f_fetch = kib / copy_fetch if copy_fetch else None
    f_write = kib / copy_write if copy_write else None
    traffic = None
    if None not in (f_fetch, f_write, score_fetch, score_write):
        traffic = 1024.0 * (score_fetch * f_fetch + score_write * f_write)
    name = next((n for n, g, v in fetch_rows if is_score(n, g)), None)
    return {
        "kernel": name,
        "copy_bytes_each_way": COPY_BYTES,
        "copy_fetch_kib": copy_fetch,
        "copy_write_kib": copy_write,
        "fetch_scale": f_fetch,
        "write_scale": f_write,
        "score_fetch_kib": score_fetch,
        "score_write_kib": score_write,
        "score_dispatches": [nf, nw],
        "traffic_bytes": traffic,
    }


def summarise_insts(rows_by_counter, work):
    """Mean per-launch SQ instruction counts of the MSAC kernel, per launch
    and per 64 residuals (one wave64 instruction's worth of residuals), and
    the FP32 operations they perform: 2 per FMA lane, 1 per add, multiply or
    transcendental lane."""
    is_score = lambda name: any(k in name for k in SCORE_KERNELS)
    per = {}
    for counter, rows in rows_by_counter.items():
        vals = [v for name, grid, v in rows if is_score(name)]
        per[counter] = sum(vals) / len(vals) if vals else None
    if any(v is None for v in per.values()):
        return None
    res = float(work["residuals"])
    fp32_ops = 64.0 * (2 * per["SQ_INSTS_VALU_FMA_F32"] + per["SQ_INSTS_VALU_ADD_F32"]
                       + per["SQ_INSTS_VALU_MUL_F32"] + per["SQ_INSTS_VALU_TRANS_F32"])
    out = {"residuals": res, "per_launch": per,
           "per_64_residuals": {k: 64.0 * v / res for k, v in per.items()},
           "fp32_ops_per_launch": fp32_ops,
           "fp32_ops_per_residual": fp32_ops / res}
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--child", action="store_true")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default="profiles/k_msac_traffic_mixed_batch_256.json")
    ap.add_argument("--workdir", default=None, help="rocprofv3 output directory")
    ap.add_argument("--work-out", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.child:
        child(args.reps, args.work_out)
        return
    work = args.workdir or tempfile.mkdtemp(prefix="tr_roofline_")
    fetch_db = _run_pass(["FETCH_SIZE"], os.path.join(work, "fetch"), args.reps)
    write_db = _run_pass(["WRITE_SIZE"], os.path.join(work, "write"), args.reps)
    work_json = os.path.join(work, "work.json")
    inst_db = _run_pass(INST_COUNTERS, os.path.join(work, "insts"), args.reps, work_json)
    s = summarise(_counter_rows(fetch_db, "FETCH_SIZE"), _counter_rows(write_db, "WRITE_SIZE"))
    with open(work_json) as f:
        s["valu"] = summarise_insts({c: _counter_rows(inst_db, c) for c in INST_COUNTERS},
                                    json.load(f))
    s["workload"] = ("mixed_batch: 256 frames, N log-uniform [500,20000], 2048 samples/frame "
                     "(seed 2), tuned batch shape")
    s["command"] = "python -m turboransac_amd.roofline --out " + args.out
    s["date"] = datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%d %H:%M UTC")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(s, f, indent=1)
    md = os.path.splitext(args.out)[0] + ".md"
    with open(md, "w") as f:
        f.write(f"# HBM traffic of `{s['kernel']}` on the 256-frame mixed batch\n\n")
        f.write(f"Command (from the repository root): `{s['command']}` ({s['date']}).\n")
        f.write("Two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), "
                "each scaled by what the counter reads for an elementwise copy of "
                f"{COPY_BYTES >> 20} MiB each way.\n\n")
        f.write("| quantity | value |\n|---|---|\n")
        for k in ("copy_fetch_kib", "copy_write_kib", "fetch_scale", "write_scale",
                  "score_fetch_kib", "score_write_kib", "score_dispatches", "traffic_bytes"):
            f.write(f"| {k} | {s[k]} |\n")
        v = s["valu"]
        if v:
            f.write("\nThird pass: SQ instruction counters (wave64 instructions per launch, "
                    f"{v['residuals']:.4g} residuals per launch).\n\n")
            f.write("| counter | per launch | per 64 residuals |\n|---|---|---|\n")
            for k in INST_COUNTERS:
                f.write(f"| {k} | {v['per_launch'][k]:.6g} | {v['per_64_residuals'][k]:.3f} |\n")
            f.write(f"\nFP32 operations executed: {v['fp32_ops_per_launch']:.6g} per launch, "
                    f"{v['fp32_ops_per_residual']:.2f} per residual.\n")
    print(json.dumps(s, indent=1))


if __name__ == "__main__":
    main()
