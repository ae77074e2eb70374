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
