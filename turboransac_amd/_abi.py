"""ctypes mirror of include/turboransac.h (structs, enums, prototypes).

The struct layouts here must match the C header field for field; the
CPU test suite checks sizes against the library's own sizeof report.
"""
from __future__ import annotations

import ctypes as C
import os

TR_OK = 0
TR_ERR_INVALID_ARGUMENT = 1
TR_ERR_TOO_FEW_POINTS = 2
TR_ERR_NOT_FINITE = 3
TR_ERR_DEVICE = 4
TR_ERR_OUT_OF_MEMORY = 5
TR_ERR_CAPACITY = 6
TR_ERR_NOT_READY = 7
TR_ERR_NO_DEVICE = 8
TR_ERR_INTERNAL = 9

STATUS_NAMES = {
    0: "TR_OK",
    1: "TR_ERR_INVALID_ARGUMENT",
    2: "TR_ERR_TOO_FEW_POINTS",
    3: "TR_ERR_NOT_FINITE",
    4: "TR_ERR_DEVICE",
    5: "TR_ERR_OUT_OF_MEMORY",
    6: "TR_ERR_CAPACITY",
    7: "TR_ERR_NOT_READY",
    8: "TR_ERR_NO_DEVICE",
    9: "TR_ERR_INTERNAL",
}

TR_LOSS_TRIVIAL = 0
TR_LOSS_TRUNCATED = 1
TR_LOSS_HUBER = 2
TR_LOSS_CAUCHY = 3
TR_LOSS_TRUNCATED_LE_ZACH = 4


class tr_camera(C.Structure):
    _fields_ = [
        ("model_id", C.c_int32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("reserved", C.c_int32),
        ("params", C.c_double * 4),
    ]


class tr_ransac_options(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint64),
        ("min_iterations", C.c_uint64),
        ("dyn_num_trials_mult", C.c_double),
        ("success_prob", C.c_double),
        ("max_reproj_error", C.c_double),
        ("seed", C.c_uint64),
        ("max_total_iterations", C.c_uint64),
        ("adaptive_stopping", C.c_int32),
        ("progressive_sampling", C.c_int32),
        ("max_prosac_iterations", C.c_uint64),
        ("score_initial_model", C.c_int32),
        ("reserved", C.c_int32),
        ("max_epipolar_error", C.c_double),
    ]


class tr_bundle_options(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint64),
        ("loss_type", C.c_int32),
        ("reserved", C.c_int32),
        ("loss_scale", C.c_double),
        ("gradient_tol", C.c_double),
        ("step_tol", C.c_double),
        ("initial_lambda", C.c_double),
        ("min_lambda", C.c_double),
        ("max_lambda", C.c_double),
    ]


class tr_pose_result(C.Structure):
    _fields_ = [
        ("q", C.c_double * 4),
        ("t", C.c_double * 3),
        ("iterations", C.c_uint64),
        ("refinements", C.c_uint64),
        ("num_inliers", C.c_uint64),
        ("inlier_ratio", C.c_double),
        ("model_score", C.c_double),
        ("num_candidates", C.c_uint64),
        ("best_hypothesis", C.c_int64),
        ("best_inliers", C.c_uint64),
        ("lm_initial_cost", C.c_double),
        ("lm_cost", C.c_double),
        ("lm_invalid_steps", C.c_uint64),
        ("lm_lambda", C.c_double),
        ("lm_step_norm", C.c_double),
        ("lm_grad_norm", C.c_double),
        ("status", C.c_int32),
        ("reserved", C.c_int32),
    ]


class tr_bundle_stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint64),
        ("initial_cost", C.c_double),
        ("cost", C.c_double),
        ("lambda_", C.c_double),
        ("invalid_steps", C.c_uint64),
        ("step_norm", C.c_double),
        ("grad_norm", C.c_double),
    ]


class tr_fundamental_result(C.Structure):
    _fields_ = [
        ("F", C.c_double * 9),
        ("iterations", C.c_uint64),
        ("refinements", C.c_uint64),
        ("num_inliers", C.c_uint64),
        ("inlier_ratio", C.c_double),
        ("model_score", C.c_double),
        ("num_candidates", C.c_uint64),
        ("best_hypothesis", C.c_int64),
        ("best_inliers", C.c_uint64),
        ("status", C.c_int32),
        ("reserved", C.c_int32),
    ]


class tr_fundamental_debug(C.Structure):
    _fields_ = [
        ("max_samples", C.c_int64),
        ("samples", C.c_void_p),
        ("num_solutions", C.c_void_p),
        ("models", C.c_void_p),
        ("scores32", C.c_void_p),
    ]


class tr_pair_frame(C.Structure):
    _fields_ = [
        ("points1", C.c_void_p),
        ("points2", C.c_void_p),
        ("num_points", C.c_int64),
        ("ransac", tr_ransac_options),
        ("bundle", tr_bundle_options),
    ]


class tr_frame(C.Structure):
    _fields_ = [
        ("points2D", C.c_void_p),
        ("points3D", C.c_void_p),
        ("num_points", C.c_int64),
        ("camera", tr_camera),
        ("ransac", tr_ransac_options),
        ("bundle", tr_bundle_options),
        ("weights", C.c_void_p),
        ("initial_q", C.c_double * 4),
        ("initial_t", C.c_double * 3),
    ]


class tr_server_options(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("max_batch_frames", C.c_int32),
        ("max_inflight", C.c_int32),
        ("return_masks", C.c_int32),
        ("reserved", C.c_int32),
    ]


class tr_debug_frame(C.Structure):
    _fields_ = [
        ("max_samples", C.c_int64),
        ("samples", C.c_void_p),
        ("num_solutions", C.c_void_p),
        ("poses", C.c_void_p),
        ("scores32", C.c_void_p),
        ("topk_ids", C.c_void_p),
        ("topk_scores64", C.c_void_p),
        ("topk_count", C.c_int32),
        ("reserved", C.c_int32),
    ]


def default_ransac_options(**kw) -> tr_ransac_options:
    """Defaults of RansacOptions_cuda (utils.h:14-31) plus the hard cap of
    turbo_ransac.cu:145; keyword arguments override fields."""
    o = tr_ransac_options(
        max_iterations=100000,
        min_iterations=1000,
        dyn_num_trials_mult=3.0,
        success_prob=0.9999,
        max_reproj_error=12.0,
        seed=0,
        max_total_iterations=100000,
        adaptive_stopping=1,
        progressive_sampling=0,
        max_prosac_iterations=100000,
        score_initial_model=0,
        reserved=0,
        max_epipolar_error=1.0,
    )
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def default_bundle_options(**kw) -> tr_bundle_options:
    """Defaults of BundleOptions_cuda (utils.h:33-50)."""
    o = tr_bundle_options(
        max_iterations=100,
        loss_type=TR_LOSS_CAUCHY,
        reserved=0,
        loss_scale=1.0,
        gradient_tol=1e-10,
        step_tol=1e-8,
        initial_lambda=1e-3,
        min_lambda=1e-10,
        max_lambda=1e10,
    )
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def make_camera(fx, fy, cx, cy, width=None, height=None) -> tr_camera:
    cam = tr_camera()
    cam.model_id = 1
    cam.width = int(width if width is not None else round(2 * cx))
    cam.height = int(height if height is not None else round(2 * cy))
    cam.params[0], cam.params[1], cam.params[2], cam.params[3] = fx, fy, cx, cy
    return cam


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libturboransac_amd.so")

# Every symbol include/turboransac.h declares.
EXPORTED_SYMBOLS = [
    "tr_ransac_options_default",
    "tr_bundle_options_default",
    "tr_status_string",
    "tr_abi_version",
    "tr_context_create",
    "tr_context_destroy",
    "tr_context_last_error",
    "tr_estimate_absolute_pose",
    "tr_estimate_absolute_pose_device",
    "tr_estimate_absolute_pose_device_noalloc",
    "tr_estimate_frame",
    "tr_bundle_adjust",
    "tr_get_inliers",
    "tr_estimate_fundamental",
    "tr_estimate_fundamental_batch",
    "tr_debug_estimate_fundamental",
    "tr_batch_create",
    "tr_batch_launch",
    "tr_batch_launch_profiled",
    "tr_batch_sync",
    "tr_batch_stage_times",
    "tr_batch_work",
    "tr_batch_results",
    "tr_batch_score_shape",
    "tr_batch_frame_inputs",
    "tr_batch_destroy",
    "tr_estimate_absolute_pose_batch",
    "tr_server_options_default",
    "tr_server_create",
    "tr_server_submit",
    "tr_server_submit_device",
    "tr_server_wait",
    "tr_server_replay",
    "tr_server_replay_device",
    "tr_server_stats",
    "tr_server_destroy",
    "tr_server_last_error",
    "tr_sample_triples",
    "tr_debug_estimate",
    "tr_score_variants",
    "tr_context_set_tuning",
    "tr_context_get_tuning",
]


def _proto(lib):
    P = C.POINTER
    vp = C.c_void_p
    i32, i64, u64 = C.c_int32, C.c_int64, C.c_uint64
    sigs = {
        "tr_ransac_options_default": (None, [P(tr_ransac_options)]),
        "tr_bundle_options_default": (None, [P(tr_bundle_options)]),
        "tr_status_string": (C.c_char_p, [i32]),
        "tr_abi_version": (i32, []),
        "tr_context_create": (i32, [i32, P(vp)]),
        "tr_context_destroy": (None, [vp]),
        "tr_context_last_error": (C.c_char_p, [vp]),
        "tr_estimate_absolute_pose": (
            i32,
            [vp, vp, vp, i64, P(tr_camera), P(tr_ransac_options), P(tr_bundle_options),
             P(tr_pose_result), vp],
        ),
        "tr_estimate_absolute_pose_device": (
            i32,
            [vp, vp, vp, i64, P(tr_camera), P(tr_ransac_options), P(tr_bundle_options),
             P(tr_pose_result), vp],
        ),
        "tr_estimate_absolute_pose_device_noalloc": (
            i32,
            [vp, vp, vp, i64, P(tr_camera), P(tr_ransac_options), P(tr_bundle_options),
             P(tr_pose_result), vp],
        ),
        "tr_estimate_frame": (i32, [vp, P(tr_frame), P(tr_pose_result), vp]),
        "tr_bundle_adjust": (
            i32, [vp, vp, vp, i64, P(tr_camera), P(tr_bundle_options), vp, vp, vp,
                  P(tr_bundle_stats)]),
        "tr_get_inliers": (i32, [vp, vp, vp, vp, vp, i64, C.c_double, vp, P(i64)]),
        "tr_estimate_fundamental": (
            i32, [vp, vp, vp, i64, P(tr_ransac_options), P(tr_bundle_options),
                  P(tr_fundamental_result), vp]),
        "tr_estimate_fundamental_batch": (
            i32, [vp, P(tr_pair_frame), i32, P(tr_fundamental_result), P(vp)]),
        "tr_debug_estimate_fundamental": (
            i32, [vp, vp, vp, i64, P(tr_ransac_options), P(tr_bundle_options),
                  P(tr_fundamental_result), vp, P(tr_fundamental_debug)]),
        "tr_batch_create": (i32, [vp, P(tr_frame), i32, P(vp)]),
        "tr_batch_launch": (i32, [vp]),
        "tr_batch_launch_profiled": (i32, [vp]),
        "tr_batch_sync": (i32, [vp]),
        "tr_batch_stage_times": (i32, [vp, P(C.c_double), P(C.c_char_p), i32]),
        "tr_batch_work": (i32, [vp, P(u64), P(u64), P(u64)]),
        "tr_batch_results": (i32, [vp, P(tr_pose_result), P(vp)]),
        "tr_batch_score_shape": (i32, [vp, P(i32)]),
        "tr_batch_frame_inputs": (i32, [vp, i32, P(vp), P(vp), P(i64)]),
        "tr_batch_destroy": (None, [vp]),
        "tr_estimate_absolute_pose_batch": (i32, [vp, P(tr_frame), i32, P(tr_pose_result), P(vp)]),
        "tr_server_options_default": (None, [P(tr_server_options)]),
        "tr_server_create": (i32, [i32, P(tr_server_options), P(vp)]),
        "tr_server_submit": (i32, [vp, P(tr_frame), P(u64)]),
        "tr_server_wait": (i32, [vp, u64, i64, P(tr_pose_result), vp, P(C.c_double)]),
        "tr_server_submit_device": (i32, [vp, P(tr_frame), P(u64)]),
        "tr_server_replay": (
            i32, [vp, P(tr_frame), P(C.c_double), i32, P(u64), P(C.c_double), P(C.c_double),
                  P(tr_pose_result), P(vp)]),
        "tr_server_replay_device": (
            i32, [vp, P(tr_frame), P(C.c_double), i32, P(u64), P(C.c_double), P(C.c_double),
                  P(tr_pose_result), P(vp)]),
        "tr_server_stats": (i32, [vp, P(u64), P(u64), P(u64)]),
        "tr_server_destroy": (None, [vp]),
        "tr_server_last_error": (C.c_char_p, [vp]),
        "tr_sample_triples": (i32, [vp, u64, i64, i64, i64, vp]),
        "tr_score_variants": (i32, [vp, i32]),
        "tr_context_set_tuning": (i32, [vp, C.c_char_p]),
        "tr_context_get_tuning": (i32, [vp, vp, i32]),
        "tr_debug_estimate": (i32, [vp, P(tr_frame), P(tr_pose_result), vp, P(tr_debug_frame)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None


ABI_VERSION = 5  # TR_ABI_VERSION of include/turboransac.h


def load_library(path: str | None = None):
    """Loads the HIP engine library.  Raises if it is missing: there is no
    fallback implementation."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"turboransac_amd engine library not built: {p} is missing "
            "(run __graft_entry__.build() or python -m turboransac_amd.build)"
        )
    lib = _proto(C.CDLL(p))
    got = lib.tr_abi_version()
    if got != ABI_VERSION:
        raise RuntimeError(f"{p} has C ABI {got}, this binding needs {ABI_VERSION}: rebuild it")
    if path is None:
        _LIB = lib
    return lib
