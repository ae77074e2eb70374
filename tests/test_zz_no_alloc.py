"""The allocation-free entry point (tr_estimate_absolute_pose_device_noalloc)
against the entry point it varies, tr_estimate_absolute_pose_device, through
the C ABI: the same result bits and inlier mask on two calls in a row, the
second with a smaller frame (the scratch array keeps the first call's
contents), and TR_ERR_CAPACITY beyond its limits before any device work.
"""
import numpy as np
import pytest

import turboransac_amd as tr
from turboransac_amd import _abi, synth

from _hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def _device(a):
    return DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=np.float64))


def _run_both(ctx, fr, ro, bo):
    bufs = [_device(fr.points2D), _device(fr.points3D), DeviceArray(fr.n).zero(),
            DeviceArray(fr.n).zero()]
    d2, d3, m_a, m_b = bufs
    try:
        a = ctx.estimate_absolute_pose_device(d2.ptr, d3.ptr, fr.n, fr.camera(), ro, bo,
                                              d_mask=m_a.ptr)
        b = ctx.estimate_absolute_pose_device_noalloc(d2.ptr, d3.ptr, fr.n, fr.camera(), ro, bo,
                                                      d_mask=m_b.ptr)
        return a, b, m_a.to_numpy(np.uint8, fr.n), m_b.to_numpy(np.uint8, fr.n)
    finally:
        for d in bufs:
            d.close()


def _same(a, b):
    for name, _ in _abi.tr_pose_result._fields_:
        va, vb = getattr(a, name), getattr(b, name)
        if hasattr(va, "__len__"):
            assert list(va) == list(vb), name
        else:
            assert va == vb, name


def test_no_alloc_matches_the_device_entry_point():
    ctx = tr.Context(0)
    try:
        rng = np.random.default_rng(88)
        # a large frame (N 120 000, 8192 samples, several rounds of the stop
        # rule), then a smaller one (N 3000, 1024 samples, one round)
        big = synth.make_frame(rng, 120000, 0.1, 1.0, hypotheses=8192, seed=5)
        small = synth.make_frame(rng, 3000, 0.5, 1.0, hypotheses=1024, seed=6)
        cases = [
            (big, big.ransac_options(adaptive_stopping=1, min_iterations=0,
                                     max_total_iterations=8192 * 3)),
            (small, small.ransac_options()),
        ]
        for fr, ro in cases:
            print(f"no-alloc call: N={fr.n} samples/round={ro.max_iterations}", flush=True)
            a, b, ma, mb = _run_both(ctx, fr, ro, fr.bundle_options())
            assert a.status == 0 and a.num_inliers > 0
            _same(a, b)
            assert np.array_equal(ma, mb)
    finally:
        ctx.close()


def test_no_alloc_rejects_what_exceeds_its_limits():
    ctx = tr.Context(0)
    try:
        rng = np.random.default_rng(89)
        fr = synth.make_frame(rng, 200001, 0.5, 1.0, hypotheses=64, seed=7)
        d2, d3 = _device(fr.points2D), _device(fr.points3D)
        try:
            with pytest.raises(tr.TRError) as e:
                ctx.estimate_absolute_pose_device_noalloc(d2.ptr, d3.ptr, fr.n, fr.camera(),
                                                          fr.ransac_options(),
                                                          fr.bundle_options())
            assert e.value.status == _abi.TR_ERR_CAPACITY
            with pytest.raises(tr.TRError) as e:
                ctx.estimate_absolute_pose_device_noalloc(
                    d2.ptr, d3.ptr, 1000, fr.camera(), fr.ransac_options(max_iterations=16385),
                    fr.bundle_options())
            assert e.value.status == _abi.TR_ERR_CAPACITY
        finally:
            d2.close()
            d3.close()
        # the entry point still works afterwards
        a, b, ma, mb = _run_both(ctx, synth.make_frame(rng, 2000, 0.6, 1.0, hypotheses=512,
                                                       seed=8),
                                 _abi.default_ransac_options(max_iterations=512, min_iterations=0,
                                                             adaptive_stopping=0,
                                                             max_reproj_error=8.0, seed=8),
                                 _abi.default_bundle_options(max_iterations=25))
        _same(a, b)
        assert np.array_equal(ma, mb)
    finally:
        ctx.close()
