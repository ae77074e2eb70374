"""HBM traffic and executed work of the MSAC and LM kernels from rocprofv3
counters.

Runs the 256-frame mixed batch (the benchmark's workload) and a device copy
of known size under `rocprofv3 --pmc` passes, each its own run:

  1. FETCH_SIZE                         KiB per dispatch
  2. WRITE_SIZE                         KiB per dispatch
  3. TCC_EA0_RDREQ_sum, TCC_BUBBLE_sum,  the read requests behind FETCH_SIZE
     TCC_EA0_RDREQ_32B_sum
  4. SQ instruction counters (VALU, SALU and the FP32 classes): the MSAC
     kernel's executed instructions per 64 residuals and FP32 operations
  5. SQ FP64 counters (FMA, ADD, MUL, TRANS): the LM kernel's executed
     FP64 operations
  6. Issue counters (GRBM_GUI_ACTIVE, SQ busy, wave and VALU-issue
     cycles): the MSAC kernel's VALU instructions per SIMD quad-cycle (at
     most 2: a gfx950 SIMD issues two wave64 VALU instructions of different
     waves in one quad-cycle), the share issued in pairs, the waves'
     residency and the shader clock during the kernel

The copy is the HIP runtime's own device-to-device copy (a blit kernel the
counters see, `__amd_rocclr_copyBuffer`), COPY_BYTES each way, through the
runtime the engine library itself links (no second runtime in the process).
Each traffic counter is scaled by what it reads for the copy (bytes moved /
counter): on gfx950 every read request to the fabric carries 128 bytes
(TCC_EA0_RDREQ_sum = bytes / 128 for a copy) while FETCH_SIZE counts
requests that TCC_BUBBLE does not flag as 64 bytes, so FETCH_SIZE reads half
of a copy's bytes; WRITE_SIZE reads all of them.  Pass 3 records the request
counts that show it for every workload.  The scaled values give the MSAC
kernel's HBM bytes per launch.

    python -m turboransac_amd.roofline --out profiles/k_msac_traffic_mixed_batch_256.json

The parent process only starts rocprofv3; the GPU work runs in the child
(`--child`), started by rocprofv3 itself.
"""
from __future__ import annotations

import argparse
import ctypes as C
import datetime
import glob
import json
import os
import sqlite3
import subprocess
import sys
import tempfile

COPY_BYTES = 256 << 20  # one direction of the calibration copy
SCORE_KERNELS = ("k_msac<", "k_msac_mfma<", "k_msac_gated<", "k_msac_gated2<")
LM_KERNELS = ("k_lm_iter",)
# Each pass within what one pass can hold (8 SQ, 4 TCC counters).
READ_COUNTERS = ("TCC_EA0_RDREQ_sum", "TCC_BUBBLE_sum", "TCC_EA0_RDREQ_32B_sum")
INST_COUNTERS = ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VALU_FMA_F32",
                 "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32",
                 "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_VALU_INT32")
F64_COUNTERS = ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
                "SQ_INSTS_VALU_TRANS_F64")
# 7 SQ counters and 1 GRBM counter: within one pass.
ISSUE_COUNTERS = ("GRBM_GUI_ACTIVE", "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES", "SQ_INSTS_VALU",
                  "SQ_ACTIVE_INST_VALU2", "SQ_ACTIVE_INST_ANY", "SQ_WAIT_INST_ANY",
                  "SQ_WAIT_ANY")


def child(reps: int, work_out: str | None) -> None:
    import turboransac_amd as tr
    from turboransac_amd import synth

    tr.load_library()  # the HIP runtime of the engine library, loaded first
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), COPY_BYTES) or hip.hipMalloc(C.byref(dst), COPY_BYTES):
        raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, COPY_BYTES):
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(dst, src, COPY_BYTES, 3):  # device to device
            raise RuntimeError("hipMemcpy failed")
    if hip.hipDeviceSynchronize():
        raise RuntimeError("hipDeviceSynchronize failed")
    hip.hipFree(src)
    hip.hipFree(dst)
    ctx = tr.Context(0)
    frames = synth.mixed_batch(seed=2, frames=256, hypotheses=2048)
    b = tr.Batch(ctx, frames)
    for _ in range(reps):
        b.launch()
        b.sync()
    print("batch shape:", [l for l in ctx.get_tuning().splitlines() if l.startswith("batch")],
          flush=True)
    if work_out:
        res, _ = b.results()  # the work counts are those of the last results() call
        w = b.work()
        # LM work of the launch: passes per frame (the initial one and one
        # per iteration) times its inliers.
        w["lm_inlier_passes"] = sum((int(r.refinements) + 1) * int(r.num_inliers) for r in res)
        with open(work_out, "w") as f:
            json.dump(w, f)
    b.close()
    ctx.close()


def _counter_rows(db_path: str, counter: str):
    """(kernel name, grid, value) per dispatch."""
    db = sqlite3.connect(db_path)
    rows = db.execute(
        "select dispatch_id, kernel_name, grid_size, value from counters_collection "
        "where counter_name = ?", (counter,)).fetchall()
    per = {}
    for did, name, grid, val in rows:
        key = (did, name, grid)
        per[key] = per.get(key, 0.0) + float(val)
    return [(name, grid, v) for (did, name, grid), v in sorted(per.items())]


def _agent(db_path: str) -> dict:
    """The GPU agent's properties as rocprofv3 recorded them (XCDs, SIMDs,
    shader engines, the maximum shader clock in MHz)."""
    db = sqlite3.connect(db_path)
    for name, in db.execute("select name from sqlite_master where type = 'table' and "
                            "name like 'rocpd_info_agent%'"):
        for typ, ext in db.execute(f"select type, extdata from {name}"):
            if typ == "GPU":
                return json.loads(ext)
    return {}


def _score_dispatches(db_path: str) -> dict:
    """dispatch id -> (duration in ns, threads) of the MSAC kernel's dispatches."""
    db = sqlite3.connect(db_path)
    out = {}
    for did, name, start, end in db.execute("select dispatch_id, name, start, end from kernels"):
        if _is_score(name):
            out[did] = float(end - start)
    return out


def summarise_issue(db_path: str) -> dict | None:
    """The MSAC kernel's issue counters, per dispatch (mean over the
    launches): GRBM_GUI_ACTIVE is the sum over the agent's XCDs, the SQ
    cycle counters are in quad-cycles summed over shader engines.

    * clock: GUI-active cycles per XCD / the dispatch's duration;
    * valu_per_simd_quad: SQ_INSTS_VALU / (SIMDs x quad-cycles), at most 2;
    * paired: the share of VALU instructions issued two in one quad-cycle
      (2 SQ_ACTIVE_INST_VALU2 / SQ_INSTS_VALU);
    * residency: wave quad-cycles / (waves x quad-cycles of the kernel)."""
    ag = _agent(db_path)
    durs = _score_dispatches(db_path)
    db = sqlite3.connect(db_path)
    per = {}
    waves = {}
    for did, name, c, v, grid in db.execute(
            "select dispatch_id, kernel_name, counter_name, value, grid_size "
            "from counters_collection"):
        if not _is_score(name):
            continue
        per.setdefault(did, {})
        per[did][c] = per[did].get(c, 0.0) + float(v)
        waves[did] = grid / 64.0
    if not per or not ag:
        return None
    rows = []
    for did, d in per.items():
        if did not in durs or any(c not in d for c in ISSUE_COUNTERS):
            continue
        cyc = d["GRBM_GUI_ACTIVE"] / ag["num_xcc"]
        quads = cyc / 4.0
        v = d["SQ_INSTS_VALU"]
        rows.append({
            "duration_ms": durs[did] * 1e-6,
            "cycles_per_xcd": cyc,
            "clock_ghz": cyc / durs[did],
            "valu_per_simd_quad": v / (ag["simd_count"] * quads),
            "paired": 2.0 * d["SQ_ACTIVE_INST_VALU2"] / v,
            "residency": d["SQ_WAVE_CYCLES"] / (waves[did] * quads),
            "busy_share_of_wave_time": d["SQ_ACTIVE_INST_ANY"] / d["SQ_WAVE_CYCLES"],
            "wait_issue_share_of_wave_time": d["SQ_WAIT_INST_ANY"] / d["SQ_WAVE_CYCLES"],
            "wait_share_of_wave_time": d["SQ_WAIT_ANY"] / d["SQ_WAVE_CYCLES"],
            "counters": d,
        })
    if not rows:
        return None
    keys = [k for k in rows[0] if k != "counters"]
    mean = {k: sum(r[k] for r in rows) / len(rows) for k in keys}
    mean["dispatches"] = len(rows)
    mean["max_clock_ghz"] = ag.get("max_engine_clk_fcompute", 0) / 1000.0
    mean["simds"] = ag["simd_count"]
    mean["xcds"] = ag["num_xcc"]
    mean["counters_first_dispatch"] = rows[0]["counters"]
    return mean


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


def _is_copy(name):
    return "copybuffer" in name.lower()


def _is_score(name):
    return any(k in name for k in SCORE_KERNELS)


def _is_lm(name):
    return any(k in name for k in LM_KERNELS)


def _mean(rows, pred):
    vals = [v for name, grid, v in rows if pred(name)]
    return (sum(vals) / len(vals) if vals else None), len(vals)


def summarise(fetch_rows, write_rows, read_rows):
    copy_fetch, _ = _mean(fetch_rows, _is_copy)
    copy_write, _ = _mean(write_rows, _is_copy)
    score_fetch, nf = _mean(fetch_rows, _is_score)
    score_write, nw = _mean(write_rows, _is_score)
    kib = COPY_BYTES / 1024.0
    f_fetch = kib / copy_fetch if copy_fetch else None
    f_write = kib / copy_write if copy_write else None
    traffic = None
    if None not in (f_fetch, f_write, score_fetch, score_write):
        traffic = 1024.0 * (score_fetch * f_fetch + score_write * f_write)
    reads = {}
    for counter, rows in read_rows.items():
        reads[counter] = {"copy": _mean(rows, _is_copy)[0], "score": _mean(rows, _is_score)[0]}
    rq = reads.get("TCC_EA0_RDREQ_sum", {})
    bytes_per_read = (COPY_BYTES / rq["copy"]) if rq.get("copy") else None
    name = next((n for n, g, v in fetch_rows if _is_score(n)), None)
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
        "read_requests": reads,
        "copy_bytes_per_read_request": bytes_per_read,
        "score_read_bytes_from_requests": (rq["score"] * bytes_per_read
                                           if rq.get("score") and bytes_per_read else None),
    }


def summarise_insts(rows_by_counter, work):
    """Mean per-launch SQ instruction counts of the MSAC kernel, per launch
    and per 64 residuals (one wave64 instruction's worth of residuals), and
    the FP32 operations they perform: 2 per FMA lane, 1 per add, multiply or
    transcendental lane."""
    per = {}
    for counter, rows in rows_by_counter.items():
        per[counter] = _mean(rows, _is_score)[0]
    if any(v is None for v in per.values()):
        return None
    res = float(work["residuals"])
    fp32_ops = 64.0 * (2 * per["SQ_INSTS_VALU_FMA_F32"] + per["SQ_INSTS_VALU_ADD_F32"]
                       + per["SQ_INSTS_VALU_MUL_F32"] + per["SQ_INSTS_VALU_TRANS_F32"])
    return {"residuals": res, "per_launch": per,
            "per_64_residuals": {k: 64.0 * v / res for k, v in per.items()},
            "fp32_ops_per_launch": fp32_ops,
            "fp32_ops_per_residual": fp32_ops / res}


def summarise_f64(rows_by_counter, work, reps):
    """FP64 operations of the LM kernel's launches of one batch launch (all
    its dispatches summed, divided by the batch launches): 2 per FMA lane, 1
    per add, multiply or transcendental lane, and per inlier pass."""
    tot = {}
    disp = 0
    for counter, rows in rows_by_counter.items():
        vals = [v for name, grid, v in rows if _is_lm(name)]
        tot[counter] = sum(vals) / reps
        disp = max(disp, len(vals))
    if not disp:
        return None
    ops = 64.0 * (2 * tot["SQ_INSTS_VALU_FMA_F64"] + tot["SQ_INSTS_VALU_ADD_F64"]
                  + tot["SQ_INSTS_VALU_MUL_F64"] + tot["SQ_INSTS_VALU_TRANS_F64"])
    passes = float(work["lm_inlier_passes"])
    return {"dispatches_per_batch_launch": disp / reps, "per_batch_launch": tot,
            "fp64_ops_per_batch_launch": ops, "inlier_passes": passes,
            "fp64_ops_per_inlier_pass": ops / passes if passes else None}


def _f_frames_per_launch(name, gx, gy, gz, wx):
    """Frames in one launch of a fundamental-matrix kernel (None for the
    others): the grid dimension that indexes the frame table -- z for the
    scoring kernel, x (one workgroup per frame) for selection, choice and
    IRLS, y for the rest."""
    if "k_f_" not in name:
        return None
    if "k_f_score" in name:
        return int(gz)
    if any(k in name for k in ("k_f_select", "k_f_choose", "k_f_irls")):
        return int(gx) // max(1, int(wx))
    return int(gy)


def kernel_stats_summary(stats_csv: str, bench_log: str, out_md: str, command: str) -> None:
    """Markdown summary of a `rocprofv3 --kernel-trace --stats` run of
    bench.py: every kernel's calls and durations, and the MSAC kernel's
    FP32 rate from the profiler's mean duration beside the bench's own
    HIP-event figure (the JSON line in bench_log)."""
    import csv

    if stats_csv.endswith(".db"):
        # rocprofv3's rocpd database: the same statistics from its kernels
        # view, the fundamental-matrix kernels split by frames per launch
        db = sqlite3.connect(stats_csv)
        groups = {}
        for n, d, gx, gy, gz, wx in db.execute(
                "select name, duration, grid_x, grid_y, grid_z, workgroup_x from kernels"):
            nf = _f_frames_per_launch(n, gx, gy, gz, wx)
            groups.setdefault((n, nf), []).append(float(d))
        rows = [{"Name": n, "Frames": nf, "Calls": len(v), "AverageNs": sum(v) / len(v), "MinNs": min(v),
                 "MaxNs": max(v), "TotalDurationNs": sum(v)}
                for (n, nf), v in groups.items()]
    else:
        rows = list(csv.DictReader(open(stats_csv)))
    bench = None
    for line in open(bench_log):
        if line.startswith("{"):
            bench = json.loads(line)
    rl = bench["roofline"] if bench else None
    lm = bench.get("lm_roofline") if bench else None
    with open(out_md, "w") as f:
        f.write("# Kernel trace of bench.py\n\n")
        f.write(f"Command (from the repository root): `{command}`.\n\n")
        f.write("| kernel | calls | mean us | min us | max us | total ms |\n|---|---|---|---|---|---|\n")
        msac = lmk = None
        for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
            name = r["Name"]
            nf = r.get("Frames")
            tag = "" if nf is None else f" ({nf} frame{'s' if nf > 1 else ''} per launch)"
            f.write(f"| `{name[:90]}`{tag} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                    f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                    f"{float(r['TotalDurationNs']) / 1e6:.3f} |\n")
            if rl and rl["kernel"] in name:
                msac = r
            if "k_lm_iter" in name:
                lmk = r
        if rl and msac:
            flop = rl["flop_per_residual"] * rl["residuals"]
            avg = float(msac["AverageNs"]) * 1e-9
            f.write(f"\n`{rl['kernel']}`: {rl['residuals']:.6g} residuals x "
                    f"{rl['flop_per_residual']} FLOP per launch.  Mean duration under the "
                    f"profiler {avg * 1e3:.4f} ms: {flop / avg / 1e12:.2f} TFLOP/s, "
                    f"{flop / avg / 1e12 / rl['peak']:.3f} of {rl['peak']}.  The same run's HIP "
                    f"events: {rl['kernel_ms']:.4f} ms, {rl['achieved']:.2f} TFLOP/s, "
                    f"{rl['frac']:.3f} (the bench's `roofline.frac`).\n")
        fund = bench.get("fundamental") if bench else None
        fk = next((r for r in rows if "k_f_score" in r["Name"] and r.get("Frames", 1) == 1),
                  None)
        if fund and fk:
            # 34 FLOP per Sampson residual: a = A (p1, 1) 12, b_0, b_1 8, the
            # numerator 4, num^2 1, the denominator 7, the test 1, the MSAC
            # weight where a wave holds an inlier 1 (amortised)
            calls = int(fk["Calls"])
            per_call = fund["sampson_residuals"] / 64.0
            avg = float(fk["AverageNs"]) * 1e-9
            f.write(f"\n`k_f_score` (fundamental matrix, {calls} calls): mean "
                    f"{avg * 1e6:.2f} us for {per_call:.4g} Sampson residuals per call (the "
                    f"64-frame stream's mean) x 34 FLOP: {34 * per_call / avg / 1e12:.2f} TFLOP/s, "
                    f"{34 * per_call / avg / 1e12 / 157.3:.3f} of 157.3 (latency-bound: one frame "
                    f"per call).  The call: {fund['p50_ms']:.3f} ms p50, "
                    f"{fund['frames_per_s']:.0f} frames/s.\n")
            for r in rows:
                nf = r.get("Frames")
                if "k_f_score" not in r["Name"] or not nf or nf == 1 or nf % 64:
                    continue
                # a batch of nf frames is the 64-frame stream nf/64 times over
                per_launch = fund["sampson_residuals"] * (nf // 64)
                avg = float(r["AverageNs"]) * 1e-9
                fb = (fund.get("batch") or {}).get(str(nf))
                f.write(f"\n`k_f_score` over a batch of {nf} frames ({r['Calls']} calls): mean "
                        f"{avg * 1e6:.2f} us for {per_launch:.4g} Sampson residuals x 34 FLOP: "
                        f"{34 * per_launch / avg / 1e12:.2f} TFLOP/s, "
                        f"{34 * per_launch / avg / 1e12 / 157.3:.3f} of 157.3."
                        + (f"  The batched call: {fb['ms_per_call']:.3f} ms, "
                           f"{fb['frames_per_s']:.0f} frames/s." if fb else "") + "\n")
        if lm and lmk:
            f.write(f"\n`k_lm_iter`: {lmk['Calls']} calls, mean "
                    f"{float(lmk['AverageNs']) / 1e3:.2f} us; the bench's LM stage "
                    f"{lm['stage_ms']:.4f} ms, {lm['inlier_passes']:.6g} inlier passes x "
                    f"{lm['flop_per_inlier_pass']} FLOP: {lm['achieved']:.3f} TFLOP/s FP64, "
                    f"{lm['frac']:.4f} of {lm['peak']}.\n")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--child", action="store_true")
    ap.add_argument("--kernel-stats", default=None,
                    help="summarise this rocprofv3 kernel_stats.csv (or results .db) of a "
                         "bench.py run")
    ap.add_argument("--bench-log", default=None, help="that run's bench.py output")
    ap.add_argument("--command", default=None, help="the profiled command, for the summary")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default="profiles/k_msac_traffic_mixed_batch_256.json")
    ap.add_argument("--workdir", default=None, help="rocprofv3 output directory")
    ap.add_argument("--work-out", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.child:
        child(args.reps, args.work_out)
        return
    if args.kernel_stats:
        kernel_stats_summary(args.kernel_stats, args.bench_log, args.out, args.command or "")
        return
    work = args.workdir or tempfile.mkdtemp(prefix="tr_roofline_")
    fetch_db = _run_pass(["FETCH_SIZE"], os.path.join(work, "fetch"), args.reps)
    write_db = _run_pass(["WRITE_SIZE"], os.path.join(work, "write"), args.reps)
    read_db = _run_pass(READ_COUNTERS, os.path.join(work, "reads"), args.reps)
    work_json = os.path.join(work, "work.json")
    inst_db = _run_pass(INST_COUNTERS, os.path.join(work, "insts"), args.reps, work_json)
    f64_db = _run_pass(F64_COUNTERS, os.path.join(work, "f64"), args.reps)
    issue_db = _run_pass(ISSUE_COUNTERS, os.path.join(work, "issue"), args.reps)
    s = summarise(_counter_rows(fetch_db, "FETCH_SIZE"), _counter_rows(write_db, "WRITE_SIZE"),
                  {c: _counter_rows(read_db, c) for c in READ_COUNTERS})
    with open(work_json) as f:
        wk = json.load(f)
    s["valu"] = summarise_insts({c: _counter_rows(inst_db, c) for c in INST_COUNTERS}, wk)
    s["lm_fp64"] = summarise_f64({c: _counter_rows(f64_db, c) for c in F64_COUNTERS}, wk,
                                 args.reps)
    s["issue"] = summarise_issue(issue_db)
    s["workload"] = ("mixed_batch: 256 frames, N log-uniform [500,20000], 2048 samples/frame "
                     "(seed 2), tuned batch shape")
    s["command"] = "python -m turboransac_amd.roofline --out " + args.out
    s["date"] = datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%d %H:%M UTC")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(s, f, indent=1)
    md = os.path.splitext(args.out)[0] + ".md"
    with open(md, "w") as f:
        f.write(f"# HBM traffic and executed work of `{s['kernel']}` (and the LM) on the "
                "256-frame mixed batch\n\n")
        f.write(f"Command (from the repository root): `{s['command']}` ({s['date']}).\n")
        f.write("rocprofv3 --pmc passes, each its own run: FETCH_SIZE, WRITE_SIZE (KiB per "
                "dispatch), the read requests behind FETCH_SIZE, SQ instruction counters, "
                "SQ FP64 counters, issue counters.  The traffic counters are scaled by what they read for "
                f"the HIP runtime's device-to-device copy of {COPY_BYTES >> 20} MiB each "
                "way.\n\n")
        f.write("| quantity | value |\n|---|---|\n")
        for k in ("copy_fetch_kib", "copy_write_kib", "fetch_scale", "write_scale",
                  "score_fetch_kib", "score_write_kib", "score_dispatches", "traffic_bytes",
                  "copy_bytes_per_read_request", "score_read_bytes_from_requests"):
            f.write(f"| {k} | {s[k]} |\n")
        f.write("\nRead requests (mean per dispatch):\n\n| counter | copy | MSAC |\n|---|---|---|\n")
        for c, v in s["read_requests"].items():
            f.write(f"| {c} | {v['copy']} | {v['score']} |\n")
        v = s["valu"]
        if v:
            f.write("\nSQ instruction counters of the MSAC kernel (wave64 instructions per "
                    f"launch, {v['residuals']:.4g} residuals per launch).\n\n")
            f.write("| counter | per launch | per 64 residuals |\n|---|---|---|\n")
            for k in INST_COUNTERS:
                f.write(f"| {k} | {v['per_launch'][k]:.6g} | {v['per_64_residuals'][k]:.3f} |\n")
            f.write(f"\nFP32 operations executed: {v['fp32_ops_per_launch']:.6g} per launch, "
                    f"{v['fp32_ops_per_residual']:.2f} per residual.\n")
        iss = s["issue"]
        if iss:
            f.write("\nIssue counters of the MSAC kernel (mean over its dispatches in the issue "
                    "pass; GRBM_GUI_ACTIVE summed over the XCDs, SQ cycles in quad-cycles):\n\n")
            f.write("| quantity | value |\n|---|---|\n")
            for k in ("duration_ms", "cycles_per_xcd", "clock_ghz", "max_clock_ghz",
                      "valu_per_simd_quad", "paired", "residency", "busy_share_of_wave_time",
                      "wait_issue_share_of_wave_time", "wait_share_of_wave_time"):
                f.write(f"| {k} | {iss[k]:.4g} |\n")
            f.write("\n`valu_per_simd_quad` is at most 2 (two wave64 VALU instructions of "
                    "different waves in one quad-cycle); `paired` is the share of VALU "
                    "instructions issued two in a quad-cycle.\n")
        l = s["lm_fp64"]
        if l:
            f.write("\nSQ FP64 counters of the LM kernel (`k_lm_iter`), all its dispatches of "
                    f"one batch launch ({l['dispatches_per_batch_launch']:.0f}):\n\n")
            f.write("| counter | per batch launch |\n|---|---|\n")
            for k in F64_COUNTERS:
                f.write(f"| {k} | {l['per_batch_launch'][k]:.6g} |\n")
            f.write(f"\nFP64 operations executed: {l['fp64_ops_per_batch_launch']:.6g} per batch "
                    f"launch over {l['inlier_passes']:.6g} inlier passes, "
                    f"{l['fp64_ops_per_inlier_pass']:.1f} per inlier pass.\n")
    print(json.dumps(s, indent=1))


if __name__ == "__main__":
    main()
