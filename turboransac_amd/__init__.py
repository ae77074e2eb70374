"""MI355X (gfx950) absolute-pose RANSAC engine: Python binding of the C ABI.

The engine is the HIP library lib/libturboransac_amd.so; this module binds
it with ctypes for tests and benchmarks.  There is no fallback: without the
library every call raises.

    ctx = Context()
    res, mask = ctx.estimate_absolute_pose(points2D, points3D, camera,
                                           ransac_options, bundle_options)

mirrors TbrCudaPnP::estimate_absolute_pose_cuda (entry.h:13-19): pixels and
world points in, world-to-camera q = [w, x, y, z], t and the inlier mask of
the RANSAC pose out.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _abi
from ._abi import (  # noqa: F401
    TR_LOSS_CAUCHY,
    TR_LOSS_HUBER,
    TR_LOSS_TRIVIAL,
    TR_LOSS_TRUNCATED,
    TR_LOSS_TRUNCATED_LE_ZACH,
    default_bundle_options,
    default_ransac_options,
    load_library,
    make_camera,
    tr_bundle_options,
    tr_bundle_stats,
    tr_camera,
    tr_frame,
    tr_pose_result,
    tr_ransac_options,
    tr_server_options,
)


class TRError(RuntimeError):
    def __init__(self, status: int, message: str = ""):
        self.status = status
        name = _abi.STATUS_NAMES.get(status, str(status))
        super().__init__(f"{name}: {message}" if message else name)


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _f64(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"expected an N x {cols} array")
    return a


def score_variants() -> list[tuple[str, int, int, int]]:
    """(engine, J, T, W) shapes of the MSAC kernel compiled into the library;
    engine is "valu", "mfma", "gated" or "gated2"."""
    lib = load_library()
    n = lib.tr_score_variants(None, 0)
    buf = (C.c_int32 * (4 * n))()
    lib.tr_score_variants(buf, n)
    names = ("valu", "mfma", "gated", "gated2")
    return [(names[buf[4 * i]], buf[4 * i + 1], buf[4 * i + 2], buf[4 * i + 3])
            for i in range(n)]


def result_dict(r: tr_pose_result) -> dict:
    return {
        "q": np.array(r.q[:]),
        "t": np.array(r.t[:]),
        "iterations": int(r.iterations),
        "refinements": int(r.refinements),
        "num_inliers": int(r.num_inliers),
        "inlier_ratio": float(r.inlier_ratio),
        "model_score": float(r.model_score),
        "num_candidates": int(r.num_candidates),
        "best_hypothesis": int(r.best_hypothesis),
        "best_inliers": int(r.best_inliers),
        "lm_initial_cost": float(r.lm_initial_cost),
        "lm_cost": float(r.lm_cost),
        "lm_invalid_steps": int(r.lm_invalid_steps),
        "lm_lambda": float(r.lm_lambda),
        "lm_step_norm": float(r.lm_step_norm),
        "lm_grad_norm": float(r.lm_grad_norm),
        "status": int(r.status),
    }


class Context:
    """One HIP device, one stream, reusable device buffers (tr_context)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.tr_context_create(int(device), C.byref(h))
        if st != 0:
            raise TRError(st, f"tr_context_create(device={device})")
        self.h = h
        self._batches = weakref.WeakSet()

    def close(self):
        """Destroys the batches made on this context, then the context."""
        if getattr(self, "h", None):
            for b in list(self._batches):
                b.close()
            self.lib.tr_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        return self.lib.tr_context_last_error(self.h).decode()

    def _check(self, st, what):
        if st != 0:
            raise TRError(st, f"{what}: {self.last_error()}")

    def estimate_absolute_pose(self, points2D, points3D, camera, ransac=None, bundle=None,
                               return_mask=True, initial=None):
        """initial: (q, t) scored first when ransac.score_initial_model (it
        travels in the result, as the reference's in/out pose)."""
        p2 = _f64(points2D, 2)
        p3 = _f64(points3D, 3)
        if len(p2) != len(p3):
            raise TRError(_abi.TR_ERR_INVALID_ARGUMENT, "points2D and points3D sizes differ")
        ransac = ransac or default_ransac_options()
        bundle = bundle or default_bundle_options()
        res = tr_pose_result()
        if initial is not None:
            res.q[:] = [float(v) for v in initial[0]]
            res.t[:] = [float(v) for v in initial[1]]
        mask = np.zeros(len(p2), dtype=np.uint8) if return_mask else None
        st = self.lib.tr_estimate_absolute_pose(
            self.h, _ptr(p2), _ptr(p3), len(p2), C.byref(camera), C.byref(ransac),
            C.byref(bundle), C.byref(res), _ptr(mask))
        self._check(st, "tr_estimate_absolute_pose")
        return res, mask

    def estimate_absolute_pose_device(self, d_points2D: int, d_points3D: int, n: int, camera,
                                      ransac=None, bundle=None, d_mask: int = 0):
        """Inputs are device addresses (e.g. torch tensor data_ptr())."""
        ransac = ransac or default_ransac_options()
        bundle = bundle or default_bundle_options()
        res = tr_pose_result()
        st = self.lib.tr_estimate_absolute_pose_device(
            self.h, C.c_void_p(d_points2D), C.c_void_p(d_points3D), int(n), C.byref(camera),
            C.byref(ransac), C.byref(bundle), C.byref(res),
            C.c_void_p(d_mask) if d_mask else None)
        self._check(st, "tr_estimate_absolute_pose_device")
        return res

    def bundle_adjust(self, x, X, camera, bundle=None, q=(1.0, 0.0, 0.0, 0.0),
                      t=(0.0, 0.0, 0.0), weights=None):
        """bundle_adjust_cuda (ba.cuh:17-24): refines (q, t) on x (N x 2, in
        the camera's pixel units) and X (N x 3).  Returns q, t and the
        tr_bundle_stats."""
        x = _f64(np.asarray(x).reshape(-1, 2), 2)
        X = _f64(np.asarray(X).reshape(-1, 3), 3)
        if len(x) != len(X):
            raise TRError(_abi.TR_ERR_INVALID_ARGUMENT, "x and X sizes differ")
        bundle = bundle or default_bundle_options()
        qq = np.ascontiguousarray(q, dtype=np.float64).copy()
        tt = np.ascontiguousarray(t, dtype=np.float64).copy()
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        st = tr_bundle_stats()
        rc = self.lib.tr_bundle_adjust(self.h, _ptr(x), _ptr(X), len(x), C.byref(camera),
                                       C.byref(bundle), _ptr(w), _ptr(qq), _ptr(tt),
                                       C.byref(st))
        self._check(rc, "tr_bundle_adjust")
        return qq, tt, st

    def get_inliers(self, q, t, x, X, sq_threshold):
        """get_inliers_mx (entry.h:22-25) on normalised image points x."""
        x = _f64(np.asarray(x).reshape(-1, 2), 2)
        X = _f64(np.asarray(X).reshape(-1, 3), 3)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        tt = np.ascontiguousarray(t, dtype=np.float64)
        mask = np.zeros(len(x), dtype=np.uint8)
        cnt = C.c_int64()
        rc = self.lib.tr_get_inliers(self.h, _ptr(qq), _ptr(tt), _ptr(x), _ptr(X), len(x),
                                     float(sq_threshold), _ptr(mask), C.byref(cnt))
        self._check(rc, "tr_get_inliers")
        return mask, cnt.value

    def estimate_fundamental_batch(self, pairs, masks=True):
        """Many two-view frames in one call (tr_estimate_fundamental_batch):
        pairs is a sequence of (points1, points2, ransac, bundle) with
        ransac/bundle None for the defaults; returns the results and, with
        masks=True, each frame's inlier mask."""
        n = len(pairs)
        arr = (_abi.tr_pair_frame * n)()
        keep = []
        counts = np.empty(n, dtype=np.int64)
        dro = dbo = None
        for i, (a, b, ro, bo) in enumerate(pairs):
            p1, p2 = _f64(a, 2), _f64(b, 2)
            if len(p1) != len(p2):
                raise ValueError(f"frame {i}: points1 and points2 differ in length")
            keep += [p1, p2]
            f = arr[i]
            f.points1 = p1.ctypes.data
            f.points2 = p2.ctypes.data
            f.num_points = counts[i] = len(p1)
            if ro is None:
                ro = dro = dro or default_ransac_options()
            if bo is None:
                bo = dbo = dbo or default_bundle_options()
            f.ransac = ro
            f.bundle = bo
        res = (_abi.tr_fundamental_result * n)()
        mlist = marr = None
        if masks:
            # one buffer, a view per frame
            starts = np.zeros(n + 1, dtype=np.int64)
            np.cumsum(counts, out=starts[1:])
            flat = np.zeros(int(starts[-1]), dtype=np.uint8)
            mlist = [flat[starts[i]:starts[i + 1]] for i in range(n)]
            ptrs = (starts[:-1] + flat.ctypes.data).astype(np.uint64)
            marr = (C.c_void_p * n).from_buffer(ptrs)
        st = self.lib.tr_estimate_fundamental_batch(self.h, arr, n, res, marr)
        self._check(st, "tr_estimate_fundamental_batch")
        return list(res), mlist

    def estimate_fundamental(self, points1, points2, ransac=None, bundle=None, debug_samples=0):
        """Fundamental matrix of 2D-2D pixel correspondences
        (tr_estimate_fundamental): returns (result, mask) or, with
        debug_samples > 0, (result, mask, debug arrays of the first
        debug_samples samples)."""
        p1 = _f64(points1, 2)
        p2 = _f64(points2, 2)
        if len(p1) != len(p2):
            raise ValueError("points1 and points2 differ in length")
        ransac = ransac or default_ransac_options()
        bundle = bundle or default_bundle_options()
        res = _abi.tr_fundamental_result()
        mask = np.zeros(len(p1), dtype=np.uint8)
        if not debug_samples:
            st = self.lib.tr_estimate_fundamental(self.h, _ptr(p1), _ptr(p2), len(p1),
                                                  C.byref(ransac), C.byref(bundle), C.byref(res),
                                                  _ptr(mask))
            self._check(st, "tr_estimate_fundamental")
            return res, mask
        S = int(debug_samples)
        out = dict(samples=np.full((S, 7), -1, dtype=np.int32),
                   num_solutions=np.full(S, -1, dtype=np.int32),
                   models=np.full((S, 3, 3, 3), np.nan),
                   scores32=np.full((S, 3), np.nan, dtype=np.float32))
        dbg = _abi.tr_fundamental_debug(
            max_samples=S, samples=out["samples"].ctypes.data,
            num_solutions=out["num_solutions"].ctypes.data, models=out["models"].ctypes.data,
            scores32=out["scores32"].ctypes.data)
        st = self.lib.tr_debug_estimate_fundamental(self.h, _ptr(p1), _ptr(p2), len(p1),
                                                    C.byref(ransac), C.byref(bundle),
                                                    C.byref(res), _ptr(mask), C.byref(dbg))
        self._check(st, "tr_debug_estimate_fundamental")
        return res, mask, out

    def estimate_absolute_pose_device_noalloc(self, d_points2D: int, d_points3D: int, n: int,
                                              camera, ransac=None, bundle=None, d_mask: int = 0):
        """tr_estimate_absolute_pose_device with no allocation in the call
        (the GPU's 64 MiB scratch array; N <= 200000, <= 16384 samples per
        round)."""
        ransac = ransac or default_ransac_options()
        bundle = bundle or default_bundle_options()
        res = tr_pose_result()
        st = self.lib.tr_estimate_absolute_pose_device_noalloc(
            self.h, C.c_void_p(d_points2D), C.c_void_p(d_points3D), int(n), C.byref(camera),
            C.byref(ransac), C.byref(bundle), C.byref(res),
            C.c_void_p(d_mask) if d_mask else None)
        self._check(st, "tr_estimate_absolute_pose_device_noalloc")
        return res

    def estimate_frame(self, frame, return_mask=True):
        """One frame with every option (tr_estimate_frame): `frame` is a
        synth.Frame-like object or a one-frame FrameSet."""
        fs = frame if isinstance(frame, FrameSet) else FrameSet([frame])
        res = tr_pose_result()
        mask = np.zeros(fs.n[0], dtype=np.uint8) if return_mask else None
        st = self.lib.tr_estimate_frame(self.h, C.byref(fs.arr[0]), C.byref(res), _ptr(mask))
        self._check(st, "tr_estimate_frame")
        return res, mask

    def debug_estimate(self, points2D, points3D, camera, ransac, bundle, max_samples,
                       weights=None, initial=None):
        p2 = _f64(points2D, 2)
        p3 = _f64(points3D, 3)
        fr = tr_frame()
        fr.points2D = p2.ctypes.data
        fr.points3D = p3.ctypes.data
        fr.num_points = len(p2)
        fr.camera = camera
        fr.ransac = ransac
        fr.bundle = bundle
        wv = None
        if weights is not None:
            wv = _f64(weights, 1).reshape(-1)
            fr.weights = wv.ctypes.data
        fr.initial_q[:] = [1.0, 0.0, 0.0, 0.0] if initial is None else [float(v) for v in initial[0]]
        fr.initial_t[:] = [0.0, 0.0, 0.0] if initial is None else [float(v) for v in initial[1]]
        S = int(max_samples)
        out = dict(
            samples=np.full((S, 3), -1, dtype=np.int32),
            num_solutions=np.full(S, -1, dtype=np.int32),
            poses=np.full((S, 4, 3, 4), np.nan),
            scores32=np.full((S, 4), np.nan, dtype=np.float32),
            topk_ids=np.full(16, -1, dtype=np.int64),
            topk_scores64=np.full(16, np.nan),
        )
        dbg = _abi.tr_debug_frame(
            max_samples=S, samples=out["samples"].ctypes.data,
            num_solutions=out["num_solutions"].ctypes.data, poses=out["poses"].ctypes.data,
            scores32=out["scores32"].ctypes.data, topk_ids=out["topk_ids"].ctypes.data,
            topk_scores64=out["topk_scores64"].ctypes.data)
        res = tr_pose_result()
        mask = np.zeros(len(p2), dtype=np.uint8)
        st = self.lib.tr_debug_estimate(self.h, C.byref(fr), C.byref(res), _ptr(mask),
                                        C.byref(dbg))
        self._check(st, "tr_debug_estimate")
        k = int(dbg.topk_count)
        out["topk_ids"] = out["topk_ids"][:k]
        out["topk_scores64"] = out["topk_scores64"][:k]
        return res, mask, out

    def set_tuning(self, table_text: str):
        """Replaces the MSAC shape table ("single|batch MAX_N ENGINE J T W CHUNK")."""
        self._check(self.lib.tr_context_set_tuning(self.h, table_text.encode()),
                    "tr_context_set_tuning")

    def get_tuning(self) -> str:
        n = self.lib.tr_context_get_tuning(self.h, None, 0)
        buf = C.create_string_buffer(n + 1)
        self.lib.tr_context_get_tuning(self.h, buf, n + 1)
        return buf.value.decode()

    def sample_triples(self, seed: int, n: int, first: int, count: int) -> np.ndarray:
        out = np.empty((int(count), 3), dtype=np.int32)
        st = self.lib.tr_sample_triples(self.h, int(seed), int(n), int(first), int(count),
                                        _ptr(out))
        self._check(st, "tr_sample_triples")
        return out


class FrameSet:
    """Keeps tr_frame structs and the arrays they point to alive."""

    def __init__(self, frames):
        self.keep = []
        arr = (tr_frame * len(frames))()
        for i, f in enumerate(frames):
            p2 = np.ascontiguousarray(f.points2D, dtype=np.float64)
            p3 = np.ascontiguousarray(f.points3D, dtype=np.float64)
            self.keep += [p2, p3]
            arr[i].points2D = p2.ctypes.data
            arr[i].points3D = p3.ctypes.data
            arr[i].num_points = len(p2)
            arr[i].camera = f.camera()
            arr[i].ransac = f.ransac_options()
            arr[i].bundle = f.bundle_options()
            w = getattr(f, "weights", None)
            if w is not None:
                w = np.ascontiguousarray(w, dtype=np.float64)
                self.keep.append(w)
                arr[i].weights = w.ctypes.data
            init = getattr(f, "initial", None)
            arr[i].initial_q[:] = [1.0, 0.0, 0.0, 0.0] if init is None else [float(v) for v in init[0]]
            arr[i].initial_t[:] = [0.0, 0.0, 0.0] if init is None else [float(v) for v in init[1]]
        self.arr = arr
        self.n = [len(f.points2D) for f in frames]

    def __len__(self):
        return len(self.n)

    def repeat(self, count: int) -> "FrameSet":
        """`count` frames cycling through this set's frames (same arrays)."""
        out = FrameSet.__new__(FrameSet)
        out.keep = [self]
        arr = (tr_frame * count)()
        for i in range(count):
            arr[i] = self.arr[i % len(self.n)]
        out.arr = arr
        out.n = [self.n[i % len(self.n)] for i in range(count)]
        return out


class DeviceFrameSet:
    """Frames whose correspondences (and weights) live in device memory:
    tr_frame structs with device addresses, for tr_server_submit_device.
    All frames share one allocation on the current device, made and filled
    through the HIP runtime the engine library links; free it with close()
    once no server request reads it any more."""

    def __init__(self, frames):
        load_library()  # the engine's HIP runtime, loaded before ctypes opens it by name
        hip = C.CDLL("libamdhip64.so.7")
        hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        hip.hipFree.argtypes = [C.c_void_p]
        self._hip = hip
        host = FrameSet(frames)  # options, initial poses and host arrays
        sizes = []
        for f in frames:
            w = getattr(f, "weights", None)
            sizes.append(8 * len(f.points2D) * (6 if w is not None else 5))
        offs = np.concatenate([[0], np.cumsum([(b + 255) // 256 * 256 for b in sizes])])
        p = C.c_void_p()
        if hip.hipMalloc(C.byref(p), C.c_size_t(max(256, int(offs[-1])))) != 0:
            raise RuntimeError(f"hipMalloc({int(offs[-1])}) failed")
        self.base = p.value
        staging = np.zeros(max(1, int(offs[-1]) // 8), dtype=np.float64)
        arr = (tr_frame * len(frames))()
        for i, f in enumerate(frames):
            n = len(f.points2D)
            o = int(offs[i]) // 8
            staging[o:o + 2 * n] = np.ascontiguousarray(f.points2D, dtype=np.float64).ravel()
            staging[o + 2 * n:o + 5 * n] = np.ascontiguousarray(f.points3D, dtype=np.float64).ravel()
            arr[i] = host.arr[i]
            arr[i].points2D = self.base + 8 * o
            arr[i].points3D = self.base + 8 * (o + 2 * n)
            w = getattr(f, "weights", None)
            if w is not None:
                staging[o + 5 * n:o + 6 * n] = np.ascontiguousarray(w, dtype=np.float64)
                arr[i].weights = self.base + 8 * (o + 5 * n)
        if hip.hipMemcpy(C.c_void_p(self.base), staging.ctypes.data_as(C.c_void_p),
                         int(offs[-1]), 1) != 0:
            raise RuntimeError("hipMemcpy to device failed")
        self.arr = arr
        self.n = list(host.n)
        self.keep = []

    def __len__(self):
        return len(self.n)

    def repeat(self, count: int) -> "DeviceFrameSet":
        """`count` frames cycling through this set's frames (same device
        arrays; this set must stay open while the repeat is used)."""
        out = DeviceFrameSet.__new__(DeviceFrameSet)
        out.keep = [self]
        out.base = None
        arr = (tr_frame * count)()
        for i in range(count):
            arr[i] = self.arr[i % len(self.n)]
        out.arr = arr
        out.n = [self.n[i % len(self.n)] for i in range(count)]
        return out

    def close(self):
        if getattr(self, "base", None):
            self._hip.hipFree(C.c_void_p(self.base))
            self.base = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Frames uploaded once, pipeline captured in a HIP graph (tr_batch)."""

    def __init__(self, ctx: Context, frames):
        self.ctx = ctx
        self.lib = ctx.lib
        self.fs = frames if isinstance(frames, FrameSet) else FrameSet(frames)
        h = C.c_void_p()
        st = self.lib.tr_batch_create(ctx.h, self.fs.arr, len(self.fs), C.byref(h))
        ctx._check(st, "tr_batch_create")
        self.h = h
        ctx._batches.add(self)

    def launch(self):
        self.ctx._check(self.lib.tr_batch_launch(self.h), "tr_batch_launch")

    def launch_profiled(self):
        self.ctx._check(self.lib.tr_batch_launch_profiled(self.h), "tr_batch_launch_profiled")

    def sync(self):
        self.ctx._check(self.lib.tr_batch_sync(self.h), "tr_batch_sync")

    def stage_times(self) -> dict:
        ms = (C.c_double * 16)()
        names = (C.c_char_p * 16)()
        k = self.lib.tr_batch_stage_times(self.h, ms, names, 16)
        return {names[i].decode(): float(ms[i]) for i in range(k)}

    def work(self) -> dict:
        r, c, s = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.lib.tr_batch_work(self.h, C.byref(r), C.byref(c), C.byref(s))
        return {"residuals": r.value, "candidates": c.value, "samples": s.value}

    def results(self, masks=False):
        n = len(self.fs)
        res = (tr_pose_result * n)()
        mlist = None
        marr = None
        if masks:
            mlist = [np.zeros(k, dtype=np.uint8) for k in self.fs.n]
            marr = (C.c_void_p * n)(*[m.ctypes.data for m in mlist])
        st = self.lib.tr_batch_results(self.h, res, marr)
        self.ctx._check(st, "tr_batch_results")
        return list(res), mlist

    def score_shape(self) -> tuple:
        """(engine, J, T, W, chunk) of the MSAC kernel this batch runs."""
        v = (C.c_int32 * 5)()
        self.ctx._check(self.lib.tr_batch_score_shape(self.h, v), "tr_batch_score_shape")
        return (("valu", "mfma", "gated", "gated2")[v[0]], v[1], v[2], v[3], v[4])

    def frame_inputs(self, f: int):
        """Device addresses (points2D, points3D) and N of uploaded frame f."""
        a, b, n = C.c_void_p(), C.c_void_p(), C.c_int64()
        st = self.lib.tr_batch_frame_inputs(self.h, int(f), C.byref(a), C.byref(b), C.byref(n))
        self.ctx._check(st, "tr_batch_frame_inputs")
        return a.value, b.value, n.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.tr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server (tr_server)."""

    def __init__(self, device=0, hbm_budget_bytes=8 << 30, max_batch_frames=256,
                 max_inflight=2, return_masks=False):
        self.lib = load_library()
        o = tr_server_options(hbm_budget_bytes=int(hbm_budget_bytes),
                              max_batch_frames=int(max_batch_frames),
                              max_inflight=int(max_inflight), return_masks=int(return_masks))
        h = C.c_void_p()
        st = self.lib.tr_server_create(int(device), C.byref(o), C.byref(h))
        if st != 0:
            raise TRError(st, "tr_server_create")
        self.h = h
        self.return_masks = return_masks

    def _check(self, st, what):
        if st != 0:
            raise TRError(st, f"{what}: {self.lib.tr_server_last_error(self.h).decode()}")

    def submit(self, frame_set: FrameSet, i: int) -> int:
        t = C.c_uint64()
        st = self.lib.tr_server_submit(self.h, C.byref(frame_set.arr[i]), C.byref(t))
        self._check(st, "tr_server_submit")
        return t.value

    def submit_device(self, frame_set: "DeviceFrameSet", i: int) -> int:
        """Submits a device-resident frame: the server reads its arrays in
        place, so frame_set must stay open until the ticket is waited for."""
        t = C.c_uint64()
        st = self.lib.tr_server_submit_device(self.h, C.byref(frame_set.arr[i]), C.byref(t))
        self._check(st, "tr_server_submit_device")
        return t.value

    def wait(self, ticket: int, n_points: int = 0, timeout_ms: int = -1):
        res = tr_pose_result()
        lat = C.c_double()
        mask = np.zeros(n_points, dtype=np.uint8) if (self.return_masks and n_points) else None
        st = self.lib.tr_server_wait(self.h, int(ticket), int(timeout_ms), C.byref(res),
                                     _ptr(mask), C.byref(lat))
        self._check(st, "tr_server_wait")
        return res, mask, lat.value

    def replay(self, frame_set, arrival_s, results=False):
        """Open-loop replay: latencies (us) and wall time (s); with
        results=True also every frame's result and (if the server keeps
        masks) mask.  A DeviceFrameSet is submitted device-resident
        (tr_server_replay_device), a FrameSet copied at submit."""
        n = len(frame_set)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        tickets = np.zeros(n, dtype=np.uint64)
        lat = np.zeros(n)
        wall = C.c_double()
        res = (tr_pose_result * n)() if results else None
        mlist = marr = None
        if results and self.return_masks:
            mlist = [np.zeros(k, dtype=np.uint8) for k in frame_set.n]
            marr = (C.c_void_p * n)(*[m.ctypes.data for m in mlist])
        fn = (self.lib.tr_server_replay_device if isinstance(frame_set, DeviceFrameSet)
              else self.lib.tr_server_replay)
        st = fn(self.h, frame_set.arr, arr.ctypes.data_as(
            C.POINTER(C.c_double)), n, tickets.ctypes.data_as(C.POINTER(C.c_uint64)),
            lat.ctypes.data_as(C.POINTER(C.c_double)), C.byref(wall), res, marr)
        self._check(st, "tr_server_replay")
        if results:
            return lat, wall.value, list(res), mlist
        return lat, wall.value

    def stats(self) -> dict:
        d, l, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.lib.tr_server_stats(self.h, C.byref(d), C.byref(l), C.byref(p))
        return {"frames_done": d.value, "launches": l.value, "peak_bytes": p.value}

    def close(self):
        if getattr(self, "h", None):
            self.lib.tr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
