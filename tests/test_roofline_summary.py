"""Arithmetic of the rocprofv3 counter summaries (no GPU, no profiler)."""
import pytest

from turboransac_amd import roofline as rf

KERNEL = "void tr::k_msac_gated2<12, 64, 4>(tr::BatchDev const*, int)"
COPY = "__amd_rocclr_copyBuffer"
LM = "tr::k_lm_iter(tr::BatchDev const*, int, int)"


def test_traffic_scales_each_counter_by_the_copy():
    kib = rf.COPY_BYTES / 1024
    # The fetch counter reads half the bytes the copy moved, the write counter all.
    fetch = [(COPY, 1, kib / 2), (KERNEL, 4096, 100.0), (KERNEL, 4096, 300.0), ("other", 1, 7.0)]
    write = [(COPY, 1, kib), (KERNEL, 4096, 10.0)]
    # 128-byte read requests: the copy's bytes / 128.
    reads = {"TCC_EA0_RDREQ_sum": [(COPY, 1, rf.COPY_BYTES / 128), (KERNEL, 4096, 1000.0)],
             "TCC_BUBBLE_sum": [(COPY, 1, 0.0), (KERNEL, 4096, 0.0)],
             "TCC_EA0_RDREQ_32B_sum": [(COPY, 1, 0.0), (KERNEL, 4096, 0.0)]}
    s = rf.summarise(fetch, write, reads)
    assert s["kernel"] == KERNEL
    assert s["fetch_scale"] == pytest.approx(2.0)
    assert s["write_scale"] == pytest.approx(1.0)
    assert s["score_dispatches"] == [2, 1]
    assert s["traffic_bytes"] == pytest.approx(1024 * (200.0 * 2 + 10.0))
    assert s["copy_bytes_per_read_request"] == pytest.approx(128.0)
    assert s["score_read_bytes_from_requests"] == pytest.approx(128000.0)


def test_instruction_summary_per_launch_and_per_residual():
    counts = {"SQ_INSTS_VALU": 1800.0, "SQ_INSTS_SALU": 300.0, "SQ_INSTS_VALU_FMA_F32": 900.0,
              "SQ_INSTS_VALU_ADD_F32": 100.0, "SQ_INSTS_VALU_MUL_F32": 200.0,
              "SQ_INSTS_VALU_TRANS_F32": 10.0, "SQ_INSTS_VALU_INT32": 50.0}
    rows = {c: [(KERNEL, 4096, v), (KERNEL, 4096, 3 * v), (COPY, 1, 1e9)] for c, v in counts.items()}
    v = rf.summarise_insts(rows, {"residuals": 6400})
    # Mean of the two kernel dispatches: twice each count; 6400 residuals = 100 x 64.
    assert v["per_launch"]["SQ_INSTS_VALU"] == pytest.approx(3600.0)
    assert v["per_64_residuals"]["SQ_INSTS_VALU"] == pytest.approx(36.0)
    ops = 64 * (2 * 1800 + 200 + 400 + 20)
    assert v["fp32_ops_per_launch"] == pytest.approx(ops)
    assert v["fp32_ops_per_residual"] == pytest.approx(ops / 6400)


def test_instruction_summary_needs_every_counter():
    rows = {c: [(KERNEL, 1, 1.0)] for c in rf.INST_COUNTERS}
    rows["SQ_INSTS_SALU"] = [(COPY, 1, 1.0)]
    assert rf.summarise_insts(rows, {"residuals": 64}) is None


def test_lm_fp64_summary_sums_dispatches_per_batch_launch():
    # 2 batch launches x 3 LM dispatches each; per dispatch FMA 10, ADD 4, MUL 2, TRANS 1
    per = {"SQ_INSTS_VALU_FMA_F64": 10.0, "SQ_INSTS_VALU_ADD_F64": 4.0,
           "SQ_INSTS_VALU_MUL_F64": 2.0, "SQ_INSTS_VALU_TRANS_F64": 1.0}
    rows = {c: [(LM, 256, v)] * 6 + [(KERNEL, 4096, 1e9)] for c, v in per.items()}
    l = rf.summarise_f64(rows, {"lm_inlier_passes": 64}, reps=2)
    assert l["dispatches_per_batch_launch"] == 3
    ops = 64 * 3 * (2 * 10 + 4 + 2 + 1)
    assert l["fp64_ops_per_batch_launch"] == pytest.approx(ops)
    assert l["fp64_ops_per_inlier_pass"] == pytest.approx(ops / 64)


def _issue_db(path, dispatches):
    """A database with the tables summarise_issue reads, as rocprofv3 writes
    them: the agent's properties, the dispatches and their counters."""
    import json
    import sqlite3

    db = sqlite3.connect(path)
    db.execute("create table rocpd_info_agent_x (type text, extdata text)")
    db.execute("insert into rocpd_info_agent_x values ('CPU', '{}')")
    db.execute("insert into rocpd_info_agent_x values ('GPU', ?)",
               (json.dumps({"num_xcc": 8, "simd_count": 1024,
                            "max_engine_clk_fcompute": 2400}),))
    db.execute("create table kernels (dispatch_id int, name text, start int, end int)")
    db.execute("create table counters_collection (dispatch_id int, kernel_name text, "
               "counter_name text, value real, grid_size int)")
    for did, name, ns, counters in dispatches:
        db.execute("insert into kernels values (?, ?, 1000, ?)", (did, name, 1000 + ns))
        for c, v in counters.items():
            db.execute("insert into counters_collection values (?, ?, ?, ?, 262144)",
                       (did, name, c, v))
    db.commit()
    db.close()


def test_issue_summary(tmp_path):
    # 1 ms at 2.0 GHz: 2e6 cycles per XCD, 5e5 quad-cycles per SIMD.  4096
    # waves alive 90% of it; 7.5e8 VALU instructions, 2.5e8 quad-cycles with
    # a pair: 7.5e8 / (1024 x 5e5) = 1.46 per SIMD quad-cycle, 2/3 paired.
    quads = 5e5
    c = {"GRBM_GUI_ACTIVE": 8 * 2e6, "SQ_BUSY_CYCLES": 32 * 2e6,
         "SQ_WAVE_CYCLES": 0.9 * 4096 * quads, "SQ_INSTS_VALU": 7.5e8,
         "SQ_ACTIVE_INST_VALU2": 2.5e8, "SQ_ACTIVE_INST_ANY": 0.5 * 0.9 * 4096 * quads,
         "SQ_WAIT_INST_ANY": 0.25 * 0.9 * 4096 * quads,
         "SQ_WAIT_ANY": 0.25 * 0.9 * 4096 * quads}
    path = str(tmp_path / "issue.db")
    _issue_db(path, [(1, KERNEL, 1_000_000, c), (2, KERNEL, 1_000_000, c),
                     (3, COPY, 5, {k: 1.0 for k in c})])
    s = rf.summarise_issue(path)
    assert s["dispatches"] == 2
    assert s["clock_ghz"] == pytest.approx(2.0)
    assert s["max_clock_ghz"] == pytest.approx(2.4)
    assert s["valu_per_simd_quad"] == pytest.approx(7.5e8 / (1024 * quads))
    assert s["paired"] == pytest.approx(2 / 3)
    assert s["residency"] == pytest.approx(0.9)
    assert s["busy_share_of_wave_time"] == pytest.approx(0.5)


def test_issue_summary_needs_every_counter(tmp_path):
    path = str(tmp_path / "issue.db")
    _issue_db(path, [(1, KERNEL, 1000, {"GRBM_GUI_ACTIVE": 1.0, "SQ_INSTS_VALU": 1.0})])
    assert rf.summarise_issue(path) is None


def test_fundamental_kernels_split_by_frames_per_launch():
    # rocprofv3 reports grid sizes in work items; the frame table is indexed
    # by z (scoring), by workgroup along x (selection, choice, IRLS), by y
    # (the rest); other kernels are not split
    f = rf._f_frames_per_launch
    assert f("tr::(anonymous namespace)::k_f_score(FFrame const*)", 512, 3, 256, 512) == 256
    assert f("tr::(anonymous namespace)::k_f_irls(FFrame const*)", 512 * 64, 1, 1, 512) == 64
    assert f("tr::(anonymous namespace)::k_f_select(FFrame const*)", 256, 1, 1, 256) == 1
    assert f("tr::(anonymous namespace)::k_f_mask(FFrame const*)", 4096, 64, 1, 256) == 64
    assert f("tr::k_msac_gated2<12, 64, 4>(tr::BatchDev const*, int)", 4096, 1, 1, 256) is None
