"""Synthetic absolute-pose workloads (the reference ships no generator).

A frame: camera rotation uniform on SO(3) (Shoemake's uniform unit
quaternion), world points in the box [0,4] x [0,3] x [0,3] m, camera centre
2.25 m behind the box centre along its optical axis.  Inliers are box points
with depth in [0.5, 4] m that project inside the 640 x 480 image, observed
with Gaussian pixel noise; outliers pair a uniform pixel with an
independent uniform box point.  Correspondences are shuffled.

The named workloads mirror the benchmark configurations:
  plain_oracle_frame   N=2000, 60% inliers, sigma 1 px, 1024 samples
  plain_gpu_frame      N=5000, 30% inliers, sigma 1.5 px, 4096 samples
  mixed_batch          256 frames, N log-uniform [500, 20000], inlier ratio
                       U[0.1, 0.8], sigma U[0.5, 2] px, K +-10%, 2048 samples
  serving_stream       N lognormal (median 3000, sigma_ln 1.0) clipped to
                       [100, 50000], inlier ratio U[0.05, 0.8], 2048 samples
  large_frame          N=200000, 5% inliers, sigma 1 px, 16384 samples
All use an 8 px MSAC threshold and at most 25 LM iterations, and draw
exactly the stated number of minimal samples (one round, no adaptive stop).
"""
from __future__ import annotations

import dataclasses
import math

from typing import Optional

import numpy as np

from . import _abi

BOX_LO = np.array([0.0, 0.0, 0.0])
BOX_HI = np.array([4.0, 3.0, 3.0])
WIDTH, HEIGHT = 640, 480
BASE_K = (525.0, 525.0, 320.0, 240.0)


@dataclasses.dataclass
class Frame:
    points2D: np.ndarray  # N x 2 pixels, float64
    points3D: np.ndarray  # N x 3 world, float64
    K: tuple  # fx, fy, cx, cy
    R: np.ndarray  # ground-truth world-to-camera rotation
    t: np.ndarray
    is_inlier: np.ndarray  # bool N
    hypotheses: int = 1024
    max_reproj_error: float = 8.0
    lm_max_iterations: int = 25
    seed: int = 0
    weights: Optional[np.ndarray] = None  # N refinement weights, or None
    initial: Optional[tuple] = None       # (q, t) initial pose, or None

    @property
    def n(self) -> int:
        return len(self.points2D)

    def camera(self) -> _abi.tr_camera:
        return _abi.make_camera(*self.K, width=WIDTH, height=HEIGHT)

    def ransac_options(self, **kw) -> _abi.tr_ransac_options:
        base = dict(max_iterations=self.hypotheses, min_iterations=0,
                    max_reproj_error=self.max_reproj_error, seed=self.seed,
                    adaptive_stopping=0, max_total_iterations=max(100000, self.hypotheses))
        base.update(kw)
        return _abi.default_ransac_options(**base)

    def bundle_options(self, **kw) -> _abi.tr_bundle_options:
        base = dict(max_iterations=self.lm_max_iterations)
        base.update(kw)
        return _abi.default_bundle_options(**base)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform rotation from a uniform unit quaternion (Shoemake 1992)."""
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    w, x = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
    y, z = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def make_frame(rng: np.random.Generator, n: int, inlier_ratio: float, sigma: float,
               K=BASE_K, hypotheses: int = 1024, max_reproj_error: float = 8.0,
               lm_max_iterations: int = 25, seed: int = 0) -> Frame:
    fx, fy, cx, cy = K
    R = random_rotation(rng)
    centre = 0.5 * (BOX_LO + BOX_HI)
    C = centre - 2.25 * R[2]  # the optical axis in world coordinates is R^T e3
    t = -R @ C
    n_in = int(round(n * inlier_ratio))
    n_out = n - n_in

    pts_in = np.empty((0, 3))
    uv_in = np.empty((0, 2))
    while len(pts_in) < n_in:
        need = n_in - len(pts_in)
        X = BOX_LO + (BOX_HI - BOX_LO) * rng.random((max(64, 4 * need), 3))
        Xc = X @ R.T + t
        z = Xc[:, 2]
        ok = (z >= 0.5) & (z <= 4.0)
        u = fx * Xc[:, 0] / np.where(ok, z, 1.0) + cx
        v = fy * Xc[:, 1] / np.where(ok, z, 1.0) + cy
        ok &= (u >= 0) & (u < WIDTH) & (v >= 0) & (v < HEIGHT)
        pts_in = np.vstack([pts_in, X[ok][:need]])
        uv_in = np.vstack([uv_in, np.stack([u[ok], v[ok]], axis=1)[:need]])
    uv_in = uv_in + rng.normal(0.0, sigma, size=uv_in.shape)

    uv_out = rng.random((n_out, 2)) * np.array([WIDTH, HEIGHT])
    pts_out = BOX_LO + (BOX_HI - BOX_LO) * rng.random((n_out, 3))

    p2 = np.vstack([uv_in, uv_out])
    p3 = np.vstack([pts_in, pts_out])
    flags = np.concatenate([np.ones(n_in, bool), np.zeros(n_out, bool)])
    perm = rng.permutation(n)
    return Frame(np.ascontiguousarray(p2[perm]), np.ascontiguousarray(p3[perm]),
                 tuple(float(k) for k in K), R, t, flags[perm], hypotheses,
                 max_reproj_error, lm_max_iterations, seed)


def _rng(seed: int, index: int = 0) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([int(seed), int(index)]))


def _frame_seed(seed: int, index: int) -> int:
    return int(np.random.SeedSequence([int(seed), int(index), 7]).generate_state(1, np.uint64)[0])


def plain_oracle_frame(seed: int = 0) -> Frame:
    """N=2000, 60% inliers, sigma 1 px, tau 8 px, 1024 samples, LM <= 25."""
    return make_frame(_rng(seed), 2000, 0.6, 1.0, hypotheses=1024,
                      seed=_frame_seed(seed, 0))


def plain_gpu_frame(seed: int = 1) -> Frame:
    """N=5000, 30% inliers, sigma 1.5 px, 4096 samples."""
    return make_frame(_rng(seed), 5000, 0.3, 1.5, hypotheses=4096,
                      seed=_frame_seed(seed, 0))


def mixed_batch(seed: int = 2, frames: int = 256, hypotheses: int = 2048) -> list[Frame]:
    """N log-uniform in [500, 20000], inlier ratio U[0.1, 0.8], sigma
    U[0.5, 2] px, each of fx, fy, cx, cy scaled by U[0.9, 1.1]."""
    out = []
    for i in range(frames):
        rng = _rng(seed, i)
        n = int(round(math.exp(rng.uniform(math.log(500), math.log(20000)))))
        w = rng.uniform(0.1, 0.8)
        sigma = rng.uniform(0.5, 2.0)
        K = tuple(k * rng.uniform(0.9, 1.1) for k in BASE_K)
        out.append(make_frame(rng, n, w, sigma, K=K, hypotheses=hypotheses,
                              seed=_frame_seed(seed, i)))
    return out


def serving_stream(seed: int = 3, frames: int = 512, hypotheses: int = 2048,
                   sigma_ln: float = 1.0) -> list[Frame]:
    """N lognormal (median 3000, sigma_ln) clipped to [100, 50000], inlier
    ratio U[0.05, 0.8], sigma 1 px."""
    out = []
    for i in range(frames):
        rng = _rng(seed, i)
        n = int(np.clip(round(3000 * math.exp(sigma_ln * rng.standard_normal())), 100, 50000))
        w = rng.uniform(0.05, 0.8)
        out.append(make_frame(rng, n, w, 1.0, hypotheses=hypotheses,
                              seed=_frame_seed(seed, i)))
    return out


def large_frame(seed: int = 4, n: int = 200000, hypotheses: int = 16384) -> Frame:
    """N=200000, 5% inliers, sigma 1 px, 16384 samples."""
    return make_frame(_rng(seed), n, 0.05, 1.0, hypotheses=hypotheses,
                      seed=_frame_seed(seed, 0))


# The reference paper's runtime-vs-N sweep (Fig. 4): synthetic frames at
# these N, 1.2 px threshold, at most 3000 samples per round, confidence
# 0.9999.  Its synthetic data are not published; the frames here have 50%
# inliers with 0.5 px noise.
PAPER_SWEEP_N = (1000, 5000, 10000, 50000, 100000, 200000)


def paper_sweep_frame(n: int, seed: int = 6) -> Frame:
    """One frame of the runtime-vs-N sweep: 50% inliers, sigma 0.5 px, tau
    1.2 px, 3000 samples per round (LM <= 25)."""
    return make_frame(_rng(seed, n), n, 0.5, 0.5, hypotheses=3000, max_reproj_error=1.2,
                      seed=_frame_seed(seed, n))


def paper_sweep_ransac_options(fr: Frame):
    """The reference's stop rule with its hard-coded constants
    (turbo_ransac.cu:223-257): rounds of 3000 samples until
    ceil(log(1 - 0.9999) / log(1 - w^3) * 3) are drawn."""
    return fr.ransac_options(adaptive_stopping=1, success_prob=0.9999, dyn_num_trials_mult=3.0,
                             max_total_iterations=300000)


def poisson_arrivals(rate_hz: float, count: int, seed: int = 5) -> np.ndarray:
    """Arrival times (s) of a Poisson process with the given rate."""
    rng = _rng(seed, int(rate_hz))
    return np.cumsum(rng.exponential(1.0 / rate_hz, size=count))


def quat_to_rotmat(q) -> np.ndarray:
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def pose_error(R_pred, t_pred, R_gt, t_gt) -> tuple[float, float]:
    """Rotation angle of R_gt R_pred^T (deg) and camera-centre distance (cm),
    the metric of the reference's evaluator (main.cu:141-158)."""
    E = np.asarray(R_gt) @ np.asarray(R_pred).T
    c = np.clip((np.trace(E) - 1.0) / 2.0, -1.0, 1.0)
    s = np.linalg.norm([E[2, 1] - E[1, 2], E[0, 2] - E[2, 0], E[1, 0] - E[0, 1]]) / 2.0
    ang = math.degrees(math.atan2(s, c))
    Cp = -np.asarray(R_pred).T @ np.asarray(t_pred)
    Cg = -np.asarray(R_gt).T @ np.asarray(t_gt)
    return ang, float(np.linalg.norm(Cp - Cg) * 100.0)


def seven_scenes_record(fr: Frame) -> dict:
    """One frame in the 7-Scenes correspondence format (README.md:54-62):
    world-to-camera 4x4, 3x3 intrinsics and [u, v, X, Y, Z] rows."""
    fx, fy, cx, cy = fr.K
    T = np.eye(4)
    T[:3, :3] = fr.R
    T[:3, 3] = fr.t
    return {
        "gt_w2c": T.tolist(),
        "intrinsic K": [[fx, 0.0, cx], [0.0, fy, cy], [0.0, 0.0, 1.0]],
        "point pair": np.column_stack([fr.points2D, fr.points3D]).tolist(),
    }


def write_seven_scenes(path: str, frames: dict) -> None:
    """Writes {name: Frame} as a scene file of the 7-Scenes format."""
    import json

    with open(path, "w") as f:
        json.dump({k: seven_scenes_record(v) for k, v in frames.items()}, f)


# ---------------------------------------------------------------------------
# Two-view frames for the fundamental matrix
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class PairFrame:
    """2D-2D correspondences of two pinhole views: x2^T F x1 = 0 for the
    inliers (before noise).  F is the true matrix in pixels, unit Frobenius
    norm, the largest entry of its last row positive."""
    points1: np.ndarray  # N x 2 pixels, view 1
    points2: np.ndarray  # N x 2 pixels, view 2
    F: np.ndarray        # 3 x 3
    K1: tuple
    K2: tuple
    R: np.ndarray        # view 2 from view 1: X2 = R X1 + t
    t: np.ndarray
    is_inlier: np.ndarray
    hypotheses: int = 1024
    threshold: float = 1.0  # Sampson error, pixels
    seed: int = 0
    refine_iterations: int = 20

    @property
    def n(self) -> int:
        return len(self.points1)

    def ransac_options(self, **kw) -> _abi.tr_ransac_options:
        base = dict(max_iterations=self.hypotheses, min_iterations=0, seed=self.seed,
                    adaptive_stopping=0, max_epipolar_error=self.threshold,
                    max_total_iterations=max(100000, self.hypotheses))
        base.update(kw)
        return _abi.default_ransac_options(**base)

    def bundle_options(self, **kw) -> _abi.tr_bundle_options:
        # A step tolerance three decades above the FP64 noise of the IRLS
        # step (~1e-12 on a unit-norm F): the stop decision does not hang on
        # the summation order.
        base = dict(max_iterations=self.refine_iterations, step_tol=1e-9)
        base.update(kw)
        return _abi.default_bundle_options(**base)


def _kmat(K):
    fx, fy, cx, cy = K
    return np.array([[fx, 0.0, cx], [0.0, fy, cy], [0.0, 0.0, 1.0]])


def fundamental_from_pose(K1, K2, R, t) -> np.ndarray:
    """F = K2^-T [t]x R K1^-1, unit norm, the largest entry of its last row
    positive (the engine's sign convention)."""
    tx = np.array([[0.0, -t[2], t[1]], [t[2], 0.0, -t[0]], [-t[1], t[0], 0.0]])
    F = np.linalg.inv(_kmat(K2)).T @ tx @ R @ np.linalg.inv(_kmat(K1))
    F = F / np.linalg.norm(F)
    if F[2, np.argmax(np.abs(F[2]))] < 0:
        F = -F
    return F


def make_pair_frame(rng: np.random.Generator, n: int, inlier_ratio: float, sigma: float,
                    K1=BASE_K, K2=BASE_K, hypotheses: int = 1024, threshold: float = 1.0,
                    seed: int = 0, max_angle_deg: float = 25.0, outlier_margin_px: float = 0.0):
    """View 2 rotated up to max_angle_deg from view 1 and displaced 0.3-1 m;
    inliers are points 2-8 m in front of view 1 that both views see inside
    the 640 x 480 image, observed with Gaussian pixel noise sigma; outliers
    pair a uniform pixel of each image (at least outlier_margin_px from the
    epipolar line of its partner, when given).  Shuffled."""
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    ang = np.radians(rng.uniform(0.0, max_angle_deg))
    c, s = math.cos(ang), math.sin(ang)
    A = np.array([[0.0, -ax[2], ax[1]], [ax[2], 0.0, -ax[0]], [-ax[1], ax[0], 0.0]])
    R = np.eye(3) + s * A + (1.0 - c) * (A @ A)
    t = rng.normal(size=3)
    t *= rng.uniform(0.3, 1.0) / np.linalg.norm(t)
    k1, k2 = _kmat(K1), _kmat(K2)
    n_in = int(round(n * inlier_ratio))
    p1 = np.empty((0, 2))
    p2 = np.empty((0, 2))
    while len(p1) < n_in:
        m = max(64, 4 * (n_in - len(p1)))
        uv = rng.random((m, 2)) * [WIDTH, HEIGHT]
        d = rng.uniform(2.0, 8.0, m)
        X1 = (np.linalg.inv(k1) @ np.column_stack([uv, np.ones(m)]).T).T * d[:, None]
        X2 = X1 @ R.T + t
        ok = X2[:, 2] > 0.5
        q = (k2 @ X2.T).T
        q = q[:, :2] / q[:, 2:]
        ok &= (q[:, 0] >= 0) & (q[:, 0] < WIDTH) & (q[:, 1] >= 0) & (q[:, 1] < HEIGHT)
        need = n_in - len(p1)
        p1 = np.vstack([p1, uv[ok][:need]])
        p2 = np.vstack([p2, q[ok][:need]])
    p1 = p1 + rng.normal(0.0, sigma, p1.shape)
    p2 = p2 + rng.normal(0.0, sigma, p2.shape)
    F = fundamental_from_pose(K1, K2, R, t)
    n_out = n - n_in
    o1 = rng.random((n_out, 2)) * [WIDTH, HEIGHT]
    o2 = rng.random((n_out, 2)) * [WIDTH, HEIGHT]
    if outlier_margin_px > 0:
        # redraw both pixels: near the epipole of image 1 no partner is far
        # from its epipolar line
        for i in range(n_out):
            while sampson_error(F, o1[i], o2[i]) < outlier_margin_px:
                o1[i] = rng.random(2) * [WIDTH, HEIGHT]
                o2[i] = rng.random(2) * [WIDTH, HEIGHT]
    P1 = np.vstack([p1, o1])
    P2 = np.vstack([p2, o2])
    flags = np.concatenate([np.ones(n_in, bool), np.zeros(n_out, bool)])
    perm = rng.permutation(n)
    return PairFrame(np.ascontiguousarray(P1[perm]), np.ascontiguousarray(P2[perm]), F,
                     tuple(float(k) for k in K1), tuple(float(k) for k in K2), R, t,
                     flags[perm], hypotheses, threshold, seed)


def sampson_error(F, x1, x2) -> float:
    """Sampson distance (pixels) of one pair under F (Hartley and Zisserman
    eq. 11.9)."""
    a = F @ np.array([x1[0], x1[1], 1.0])
    b = F.T @ np.array([x2[0], x2[1], 1.0])
    num = x2[0] * a[0] + x2[1] * a[1] + a[2]
    return float(abs(num) / math.sqrt(a[0] ** 2 + a[1] ** 2 + b[0] ** 2 + b[1] ** 2))


def pair_stream(seed: int = 8, frames: int = 64, hypotheses: int = 1024) -> list:
    """Two-view frames: N log-uniform in [200, 8000], inlier ratio
    U[0.3, 0.9], sigma U[0.3, 1] px, threshold 1.5 px."""
    out = []
    for i in range(frames):
        rng = _rng(seed, i)
        n = int(round(math.exp(rng.uniform(math.log(200), math.log(8000)))))
        out.append(make_pair_frame(rng, n, rng.uniform(0.3, 0.9), rng.uniform(0.3, 1.0),
                                   hypotheses=hypotheses, threshold=1.5,
                                   seed=_frame_seed(seed, i)))
    return out
