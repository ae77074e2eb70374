"""Oracle: the options the reference declares without implementing them --
PROSAC sampling, per-point refinement weights and the initial model --
pinned by independent restatements and exact identities."""
import math

import numpy as np
import pytest
from scipy.optimize import least_squares
from scipy.spatial.transform import Rotation

from turboransac_amd import _abi, synth


# --------------------------------------------------------------------------
# PROSAC (Chum and Matas, CVPR 2005)
# --------------------------------------------------------------------------
def prosac_plan(N, T_N, t_max):
    """The paper's algorithm run sample by sample: for t = 1..t_max, the
    subset size n and whether the sample is drawn from the top n (True) or
    is u_n plus m-1 of the top n-1 (False)."""
    m = 3
    Tn = float(T_N)
    for i in range(m):
        Tn *= (m - i) / (N - i)
    n, Tnp = m, 1
    out = []
    for t in range(1, t_max + 1):
        if t == Tnp and n < N:
            Tn1 = Tn * (n + 1) / (n + 1 - m)
            Tnp = Tnp + math.ceil(Tn1 - Tn)
            Tn = Tn1
            n += 1
        out.append((n, Tnp < t))
    return out


def _words(orc, seed, g):
    blk = 0
    while True:
        for w in orc.philox([g & 0xFFFFFFFF, g >> 32, blk, 0], [seed & 0xFFFFFFFF, seed >> 32]):
            yield w
        blk += 1


def _bounded(it, b):
    x = next(it)
    m = x * b
    thr = (1 << 32) % b
    while (m & 0xFFFFFFFF) < thr:
        x = next(it)
        m = x * b
    return m >> 32


@pytest.mark.parametrize("N,T_N,t_max", [(3, 100, 20), (5, 50, 60), (40, 2000, 400),
                                         (1000, 100000, 3000), (200000, 100000, 500)])
def test_prosac_follows_the_paper(orc, N, T_N, t_max):
    seed = 0x1234567890ABCDEF
    got = orc.sample_prosac(seed, N, T_N, t_max, 0, t_max)
    for t, (n, full) in enumerate(prosac_plan(N, T_N, t_max), start=1):
        s = got[t - 1]
        assert len(set(s.tolist())) == 3 and s.min() >= 0
        it = _words(orc, seed, t - 1)
        if full:  # Floyd 3-subset of the top n
            i0 = _bounded(it, n - 2)
            i1 = _bounded(it, n - 1)
            i1 = n - 2 if i1 == i0 else i1
            i2 = _bounded(it, n)
            i2 = n - 1 if i2 in (i0, i1) else i2
            assert s.tolist() == [i0, i1, i2], (t, n)
        else:  # u_n and a Floyd 2-subset of the top n - 1
            k = n - 1
            i0 = _bounded(it, k - 1)
            i1 = _bounded(it, k)
            i1 = k - 1 if i1 == i0 else i1
            assert s.tolist() == [i0, i1, n - 1], (t, n)


def test_prosac_is_uniform_after_max_prosac(orc):
    seed, N, T_N = 5, 300, 200
    got = orc.sample_prosac(seed, N, T_N, T_N, T_N, 100)  # samples t = T_N + 1 ...
    ref = orc.sample_triples(seed, N, T_N, 100)
    assert np.array_equal(got, ref)


def test_prosac_prefers_the_top_ranked(orc):
    got = orc.sample_prosac(9, 5000, 100000, 2000, 0, 2000)
    assert got.max() < 2003  # the subset grows by at most one per sample
    assert np.mean(got) < 0.25 * 5000


# --------------------------------------------------------------------------
# Refinement weights (RealWeightVector_cuda, jaco_utils.cuh:29-41)
# --------------------------------------------------------------------------
def _lm_problem(rng, n=200, noise=0.002):
    R = Rotation.random(random_state=rng.integers(1 << 31)).as_matrix()
    t = np.array([0.1, -0.2, 4.0])
    X = rng.uniform(-1, 1, (n, 3)) @ R
    Z = X @ R.T + t
    x = Z[:, :2] / Z[:, 2:] + rng.normal(0, noise, (n, 2))
    K = np.array([1.0, 1.0, 0.0, 0.0])
    return x, X, R, t, K


def test_unit_weights_equal_no_weights(orc):
    x, X, R, t, K = _lm_problem(np.random.default_rng(2))
    opt = _abi.default_bundle_options(loss_scale=0.01)
    q0 = orc.rotmat_to_quat(R @ Rotation.from_rotvec([0.02, -0.01, 0.01]).as_matrix())
    a = orc.lm_refine(x, X, K, opt, q0, t + 0.03)
    b = orc.lm_refine_weighted(x, X, np.ones(len(x)), K, opt, q0, t + 0.03)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2].iterations == b[2].iterations


@pytest.mark.parametrize("loss", [_abi.TR_LOSS_CAUCHY, _abi.TR_LOSS_HUBER])
def test_integer_weights_equal_repeated_points(orc, loss):
    """A weight k is the point taken k times (any loss)."""
    rng = np.random.default_rng(3)
    x, X, R, t, K = _lm_problem(rng, n=120)
    w = rng.integers(1, 4, len(x)).astype(np.float64)
    opt = _abi.default_bundle_options(loss_type=loss, loss_scale=0.004, max_iterations=60)
    q0 = orc.rotmat_to_quat(R @ Rotation.from_rotvec([0.03, 0.01, -0.02]).as_matrix())
    qa, ta, sa = orc.lm_refine_weighted(x, X, w, K, opt, q0, t - 0.05)
    rep = np.repeat(np.arange(len(x)), w.astype(int))
    qb, tb, sb = orc.lm_refine(x[rep], X[rep], K, opt, q0, t - 0.05)
    assert abs(sa.cost - sb.cost) <= 1e-12 * max(1.0, sb.cost)
    assert np.allclose(qa * np.sign(qa[0]), qb * np.sign(qb[0]), atol=1e-10)
    assert np.allclose(ta, tb, atol=1e-10)
    cw = orc.lm_cost_weighted(x, X, w, K, opt, q0, t - 0.05)
    cr = orc.lm_cost(x[rep], X[rep], K, opt, q0, t - 0.05)
    assert abs(cw - cr) <= 1e-12 * cr


def test_weighted_linear_loss_matches_scipy(orc):
    """Trivial loss: sum w_i |r_i|^2 is least squares of sqrt(w_i) r_i."""
    rng = np.random.default_rng(4)
    x, X, R, t, K = _lm_problem(rng, n=150, noise=0.004)
    w = rng.uniform(0.2, 3.0, len(x))
    opt = _abi.default_bundle_options(loss_type=_abi.TR_LOSS_TRIVIAL, max_iterations=200,
                                      gradient_tol=1e-15, step_tol=1e-15)
    q0 = orc.rotmat_to_quat(R @ Rotation.from_rotvec([0.01, 0.02, -0.01]).as_matrix())
    q, tt, st = orc.lm_refine_weighted(x, X, w, K, opt, q0, t + 0.02)

    def res(p):
        Rm = Rotation.from_rotvec(p[:3]).as_matrix()
        Z = X @ Rm.T + p[3:]
        return (np.sqrt(w)[:, None] * (Z[:, :2] / Z[:, 2:] - x)).ravel()

    r0 = Rotation.from_quat([q0[1], q0[2], q0[3], q0[0]]).as_rotvec()
    sc = least_squares(res, np.concatenate([r0, t + 0.02]), xtol=1e-15, ftol=1e-15, gtol=1e-15)
    Rs = Rotation.from_rotvec(sc.x[:3]).as_matrix()
    assert np.allclose(orc.quat_to_rotmat(q), Rs, atol=1e-9)
    assert np.allclose(tt, sc.x[3:], atol=1e-9)


# --------------------------------------------------------------------------
# Initial model (score_initial_model)
# --------------------------------------------------------------------------
def test_initial_model_is_scored_first_and_kept_when_best(orc):
    rng = np.random.default_rng(5)
    fr = synth.make_frame(rng, 1500, 0.25, 1.0, hypotheses=16, seed=4)
    ro = fr.ransac_options(score_initial_model=1)
    q_gt = orc.rotmat_to_quat(fr.R)
    st, res, mask, trc = orc.estimate(fr.points2D, fr.points3D, fr.camera(), ro,
                                      fr.bundle_options(), initial=(q_gt, fr.t))
    assert st == 0 and res.best_hypothesis == -2
    f = 0.5 * (fr.K[0] + fr.K[1])
    xn = np.column_stack([(fr.points2D[:, 0] - fr.K[2]) / fr.K[0],
                          (fr.points2D[:, 1] - fr.K[3]) / fr.K[1]])
    P = np.column_stack([orc.quat_to_rotmat(q_gt), fr.t])
    score, nin = orc.msac_score(xn, fr.points3D, P, (fr.max_reproj_error / f) ** 2)
    assert res.model_score == score and res.best_inliers == nin
    # without the option the same frame picks a sampled hypothesis
    st, plain, _, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(), fr.ransac_options(),
                                   fr.bundle_options())
    assert plain.best_hypothesis >= 0 and plain.model_score > res.model_score


def test_a_worse_initial_model_changes_nothing(orc):
    rng = np.random.default_rng(6)
    fr = synth.make_frame(rng, 1200, 0.6, 1.0, hypotheses=256, seed=8)
    far = (np.array([0.0, 1.0, 0.0, 0.0]), np.array([5.0, 5.0, -3.0]))
    st, a, ma, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(),
                                fr.ransac_options(score_initial_model=1), fr.bundle_options(),
                                initial=far)
    st, b, mb, _ = orc.estimate(fr.points2D, fr.points3D, fr.camera(), fr.ransac_options(),
                                fr.bundle_options())
    assert a.best_hypothesis == b.best_hypothesis >= 0
    assert list(a.q) == list(b.q) and np.array_equal(ma, mb)
