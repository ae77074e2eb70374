// fundamental.hip - fundamental-matrix RANSAC on 2D-2D correspondences.
//
// The reference paper's Table I estimator, which its checkout does not hold
// (README.md:13); the threshold is RansacOptions_cuda::max_epipolar_error
// (utils.h:20).  One frame per call, one round of max_iterations samples:
//
//   k_f_hartley    per image: centroid and mean distance (Hartley 1997)
//   k_f_normalise  FP64 normalised coordinates; FP32 SoA for scoring
//   k_f_sample     Floyd 7-subset per lane, 7-point solver in FP64
//                  (Householder QR null space, cubic det(F2 + a D) = 0)
//   k_f_score      FP32 MSAC on the Sampson error, points on the lanes,
//                  the division gated by an exact inlier test
//   k_f_select     per-candidate fold of the chunk sums, the 8 lowest
//   k_f_rerank     FP64 Sampson MSAC of the 8 leaders
//   k_f_choose     FP64 best of the leaders ((score, id) minimum)
//   k_f_mask       inlier mask of the chosen model, its count
//   k_f_irls       Sampson-weighted IRLS 8-point refinement with rank 2,
//                  then F in pixels (one workgroup)
//
// Compiled with -ffp-contract=off like the rest of the engine: the FP64
// stages round operation by operation; the FP32 scoring spells its fused
// multiply-adds out with fmaf.
#include <float.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"

namespace tr {

namespace {

constexpr int kF7 = 7;
constexpr int kFSolutions = 3;      // 7-point solutions per sample at most
constexpr int kFSlot = 4;           // slots per sample: id = 4 g + k
constexpr int kFConst = 16;         // FP32 constants per model (15 used)
constexpr int kFRed = 1024;         // threads of the one-workgroup kernels
constexpr int kFIrls = 512;         // threads of the refinement workgroup (45 FP64 sums each)
constexpr int kFPtBlock = 256;      // points per normalise/rerank/mask block
constexpr int kFInverseIterations = 6;
constexpr uint64_t kFMaxSamples = uint64_t(1) << 22;
constexpr size_t kFUploadPiece = size_t(4) << 20;  // bytes per host-to-device copy of inputs
// Scoring shape: J points per lane, W waves, CB candidates per workgroup.
constexpr int kFJ = 8;
constexpr int kFW = 4;
constexpr int kFCB = 32;
constexpr int kFG = 8;                   // candidates per wave reduction (divides kFCB)
constexpr int kFChunk = 64 * kFJ * kFW;  // points per scoring workgroup
constexpr int kFRows = 4;                // rows of 64 points per inlier-test branch
static_assert(kFCB % kFG == 0 && (kFG & (kFG - 1)) == 0 && kFJ % kFRows == 0, "scoring shape");

struct FHeader {
    double T1[3], T2[3];  // s, cx, cy per image
    double best_score;    // FP64 MSAC of the chosen model, pixels^2
    double Fn[9];         // chosen model, normalised coordinates; refined in place
    double F[9];          // in pixels
    int64_t best_slot;    // 4 g + k, or -1
    uint64_t num_inliers;
    uint64_t best_inliers;
    uint64_t refinements;
    uint32_t n_cand;
    int32_t n_lead;
    uint32_t lead_slot[kTopK];
};

// One two-view frame of a launch as the kernels see it: its header, its
// arrays in the launch's arena and its sizes.  A launch holds one or many
// frames; every kernel takes the table and finds its frame by a grid
// dimension, so a frame runs the same work partition alone or in a batch.
struct FFrame {
    FHeader* h;
    const double* p1;   // N x 2 pixels, image 1
    const double* p2;   // N x 2 pixels, image 2
    double* x1n;        // N x 2 normalised (Hartley), FP64
    double* x2n;
    float* pts;         // FP32 scoring SoA, 4 x np: p1 - c1 (u, v), (p2 - c2) / tau (u, v)
    double* m64;        // 4 H x 9: models by slot 4 g + k, normalised
    float* m32;         // 4 H x 16: their FP32 scoring constants
    uint32_t* cslot;    // 3 H: the filled slots
    float* part;        // chunks x cap: FP32 chunk sums
    double* rra;        // rblocks x kTopK: FP64 re-rank block sums
    uint32_t* rrc;      // rblocks x kTopK: inlier counts
    uint8_t* mask;      // N
    int64_t n, np, H, cap;
    int32_t chunks, rblocks;
    double tau;
    uint64_t seed;
    uint64_t irls_max;  // IRLS iterations at most
    double step_tol;
    // Stage dumps of a one-frame debug call, else null.
    int32_t* dbg_samples;  // S x 7
    int32_t* dbg_nsol;     // S
    double* dbg_models;    // S x 3 x 9
    float* dbg_scores;     // S x 3
    int64_t dbg_max;
};

__device__ inline double blk_sum_f64(double v, double* s_red) {
    // Fixed order: xor butterfly per wave (every lane ends with the same
    // bits), then the waves' sums in order by thread 0; broadcast.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = s_red[0];
        for (int w = 1; w < nw; ++w) s += s_red[w];
        s_red[32] = s;
    }
    __syncthreads();
    return s_red[32];
}

// ---------------------------------------------------------------------------
// Hartley normalisation, one workgroup per image.  Also clears the header's
// counters (the launch's first kernel).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFRed) void k_f_hartley(const FFrame* __restrict__ fr) {
    __shared__ double s_red[33];
    const FFrame& F = fr[blockIdx.y];
    FHeader* h = F.h;
    const int64_t n = F.n;
    const double* p = blockIdx.x ? F.p2 : F.p1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        h->n_cand = 0;
        h->n_lead = 0;
        h->num_inliers = 0;
        h->best_inliers = 0;
        h->refinements = 0;
        h->best_slot = -1;
        h->best_score = DBL_MAX;
        for (int i = 0; i < 9; ++i) {
            h->Fn[i] = 0.0;
            h->F[i] = 0.0;
        }
    }
    double sx = 0.0, sy = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        sx += p[2 * i];
        sy += p[2 * i + 1];
    }
    const double cx = blk_sum_f64(sx, s_red) / (double)n;
    const double cy = blk_sum_f64(sy, s_red) / (double)n;
    double d = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double dx = p[2 * i] - cx, dy = p[2 * i + 1] - cy;
        d += sqrt(dx * dx + dy * dy);
    }
    d = blk_sum_f64(d, s_red) / (double)n;
    if (threadIdx.x == 0) {
        double* T = blockIdx.x ? h->T2 : h->T1;
        T[0] = d > 0.0 ? sqrt(2.0) / d : 1.0;
        T[1] = cx;
        T[2] = cy;
    }
}

// FP64 normalised coordinates x_n = s (x - c) of both images, and the FP32
// scoring arrays (SoA, np entries each): image 1 centred, in pixels; image 2
// centred and divided by the threshold.  Padding gets u1 = NaN, which fails
// the inlier test.
__global__ __launch_bounds__(kFPtBlock) void k_f_normalise(const FFrame* __restrict__ fr) {
    const FFrame& F = fr[blockIdx.y];
    const int64_t n = F.n, np = F.np;
    const int64_t i = (int64_t)blockIdx.x * kFPtBlock + threadIdx.x;
    if (i >= np) return;
    const double* p1 = F.p1;
    const double* p2 = F.p2;
    const FHeader* h = F.h;
    double* x1n = F.x1n;
    double* x2n = F.x2n;
    float* pts = F.pts;
    const double inv_tau = 1.0 / F.tau;
    if (i < n) {
        const double s1 = h->T1[0], c1x = h->T1[1], c1y = h->T1[2];
        const double s2 = h->T2[0], c2x = h->T2[1], c2y = h->T2[2];
        const double d1x = p1[2 * i] - c1x, d1y = p1[2 * i + 1] - c1y;
        const double d2x = p2[2 * i] - c2x, d2y = p2[2 * i + 1] - c2y;
        x1n[2 * i] = s1 * d1x;
        x1n[2 * i + 1] = s1 * d1y;
        x2n[2 * i] = s2 * d2x;
        x2n[2 * i + 1] = s2 * d2y;
        pts[i] = (float)d1x;
        pts[np + i] = (float)d1y;
        pts[2 * np + i] = (float)(d2x * inv_tau);
        pts[3 * np + i] = (float)(d2y * inv_tau);
    } else {
        pts[i] = __builtin_nanf("");
        pts[np + i] = 0.0f;
        pts[2 * np + i] = 0.0f;
        pts[3 * np + i] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// 7-point solver (Hartley and Zisserman, 2nd ed., 11.1.2), FP64.
// ---------------------------------------------------------------------------
__device__ inline double det3d(const double* M) {
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
           M[2] * (M[3] * M[7] - M[4] * M[6]);
}

__device__ inline void cofactor3d(const double* M, double* C) {
    C[0] = M[4] * M[8] - M[5] * M[7];
    C[1] = M[5] * M[6] - M[3] * M[8];
    C[2] = M[3] * M[7] - M[4] * M[6];
    C[3] = M[2] * M[7] - M[1] * M[8];
    C[4] = M[0] * M[8] - M[2] * M[6];
    C[5] = M[1] * M[6] - M[0] * M[7];
    C[6] = M[1] * M[5] - M[2] * M[4];
    C[7] = M[2] * M[3] - M[0] * M[5];
    C[8] = M[0] * M[4] - M[1] * M[3];
}

// Real roots of c3 x^3 + c2 x^2 + c1 x + c0: one root from the depressed
// form (Cardano without cancellation when there is one real root, the
// largest-magnitude root of Viete's trigonometric form otherwise), polished;
// the quadratic factor of the deflation from whichever coefficient formula
// rounds less, its discriminant deciding the other two roots; every root
// polished by two Newton steps on the original coefficients.
__device__ int f_solve_cubic(double c3, double c2, double c1, double c0, double x[3]) {
    const double scale = fmax(fmax(fabs(c3), fabs(c2)), fmax(fabs(c1), fabs(c0)));
    if (!(scale > 0.0)) return 0;
    int nr = 0;
    if (fabs(c3) <= 1e-14 * scale) {
        if (fabs(c2) <= 1e-14 * scale) {
            if (fabs(c1) <= 1e-14 * scale) return 0;
            x[0] = -c0 / c1;
            return 1;
        }
        const double disc = c1 * c1 - 4.0 * c2 * c0;
        if (disc < 0.0) return 0;
        const double sq = sqrt(disc);
        const double qq = -0.5 * (c1 + (c1 >= 0.0 ? sq : -sq));
        x[nr++] = qq / c2;
        if (qq != 0.0) x[nr++] = c0 / qq;
        return nr;
    }
    const double a = c2 / c3, b = c1 / c3, c = c0 / c3;
    const double p = b - a * a / 3.0;
    const double q = 2.0 * a * a * a / 27.0 - a * b / 3.0 + c;
    const double disc = 0.25 * q * q + p * p * p / 27.0;
    double r1;
    if (disc > 0.0) {
        const double sd = sqrt(disc);
        const double A = -cbrt(0.5 * fabs(q) + sd) * (q >= 0.0 ? 1.0 : -1.0);
        const double B = A != 0.0 ? -p / (3.0 * A) : 0.0;
        r1 = A + B - a / 3.0;
    } else {
        const double r = sqrt(fmax(0.0, -p / 3.0));
        const double cosv = (r > 0.0) ? fmin(1.0, fmax(-1.0, -0.5 * q / (r * r * r))) : 1.0;
        const double phi = acos(cosv);
        r1 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double xk = 2.0 * r * cos((phi + 2.0 * M_PI * (double)k) / 3.0) - a / 3.0;
            if (fabs(xk) > fabs(r1)) r1 = xk;
        }
    }
    for (int it = 0; it < 2; ++it) {
        const double f = ((r1 + a) * r1 + b) * r1 + c;
        const double df = (3.0 * r1 + 2.0 * a) * r1 + b;
        if (df != 0.0) r1 -= f / df;
    }
    x[nr++] = r1;
    double e, f;
    if (r1 != 0.0) {
        f = -c / r1;
        const double err_a = fmax(fabs(a), fabs(r1));
        const double err_b = (fabs(f) + fabs(b)) / fabs(r1);
        e = err_a <= err_b ? a + r1 : (f - b) / r1;
    } else {
        f = b;
        e = a;
    }
    const double dq = e * e - 4.0 * f;
    if (dq >= 0.0) {
        const double sq = sqrt(dq);
        const double qq = -0.5 * (e + (e >= 0.0 ? sq : -sq));
        if (qq != 0.0) {
            x[nr++] = qq;
            x[nr++] = f / qq;
        } else {
            x[nr++] = 0.0;
            x[nr++] = 0.0;
        }
    }
    for (int i = 0; i < nr; ++i)
        for (int it = 0; it < 2; ++it) {
            const double fv = ((c3 * x[i] + c2) * x[i] + c1) * x[i] + c0;
            const double df = (3.0 * c3 * x[i] + 2.0 * c2) * x[i] + c1;
            if (df != 0.0) x[i] -= fv / df;
        }
    return nr;
}

__device__ inline bool f_unit(double* F) {
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += F[i] * F[i];
    if (!(s > 1e-300)) return false;
    const double inv = 1.0 / sqrt(s);
    for (int i = 0; i < 9; ++i) F[i] *= inv;
    return true;
}

// The 7 correspondences (u1, v1) <-> (u2, v2) (normalised coordinates):
// the null space {F1, F2} of their 7 x 9 epipolar system A, the last two
// columns of Q in the Householder QR of A^T (reflection k zeroes column k
// below the diagonal), and the real roots of det(F2 + x D) = 0 with
// D = F1 - F2, by the cofactor expansion det F2 + x <cof F2, D> +
// x^2 <F2, cof D> + x^3 det D.  Returns the root count (0 for a rank-
// deficient sample); solution r is F2 + roots[r] D, unit norm.
__device__ int f_seven_point(const double* u1, const double* v1, const double* u2,
                             const double* v2, double F2[9], double Dout[9], double roots[3]) {
    double M[9][kF7];
#pragma unroll
    for (int i = 0; i < kF7; ++i) {
        M[0][i] = u2[i] * u1[i];
        M[1][i] = u2[i] * v1[i];
        M[2][i] = u2[i];
        M[3][i] = v2[i] * u1[i];
        M[4][i] = v2[i] * v1[i];
        M[5][i] = v2[i];
        M[6][i] = u1[i];
        M[7][i] = v1[i];
        M[8][i] = 1.0;
    }
    double beta[kF7];
#pragma unroll
    for (int k = 0; k < kF7; ++k) {
        double nrm = 0.0;
#pragma unroll
        for (int r = k; r < 9; ++r) nrm += M[r][k] * M[r][k];
        nrm = sqrt(nrm);
        if (!(nrm > 0.0)) return 0;
        const double alpha = M[k][k] > 0.0 ? -nrm : nrm;
        M[k][k] -= alpha;
        double vtv = 0.0;
#pragma unroll
        for (int r = k; r < 9; ++r) vtv += M[r][k] * M[r][k];
        if (!(vtv > 0.0)) return 0;
        beta[k] = 2.0 / vtv;
#pragma unroll
        for (int c = k + 1; c < kF7; ++c) {
            double d = 0.0;
#pragma unroll
            for (int r = k; r < 9; ++r) d += M[r][k] * M[r][c];
            d *= beta[k];
#pragma unroll
            for (int r = k; r < 9; ++r) M[r][c] -= d * M[r][k];
        }
    }
    double N[2][9];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int r = 0; r < 9; ++r) N[j][r] = (r == kF7 + j) ? 1.0 : 0.0;
#pragma unroll
        for (int k = kF7 - 1; k >= 0; --k) {
            double d = 0.0;
#pragma unroll
            for (int r = k; r < 9; ++r) d += M[r][k] * N[j][r];
            d *= beta[k];
#pragma unroll
            for (int r = k; r < 9; ++r) N[j][r] -= d * M[r][k];
        }
    }
    double D[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) D[i] = N[0][i] - N[1][i];
    double C2[9], CD[9];
    cofactor3d(N[1], C2);
    cofactor3d(D, CD);
    double k1 = 0.0, k2 = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        k1 += C2[i] * D[i];
        k2 += N[1][i] * CD[i];
    }
    const int nr = f_solve_cubic(det3d(D), k2, k1, det3d(N[1]), roots);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        F2[i] = N[1][i];
        Dout[i] = D[i];
    }
    return nr;
}

// ---------------------------------------------------------------------------
// Sampler + 7-point: one sample per lane.  Solution k of sample g is slot
// 4 g + k: FP64 model (normalised) and the FP32 constants of the scoring
// kernel.  The wave appends its slots to the candidate list with one atomic.
//
// FP32 constants: with S_i = diag(s_i, s_i, 1), F_c = S2 F_n S1 is F on the
// centred pixel coordinates p1, p2.  The scoring arrays hold p1 and
// q2 = p2 / tau; then a = A (p1, 1) with A = F_c whose last row is divided
// by tau, b_j = B_j . (q2, 1) with B_j = (tau F_c[0][j], tau F_c[1][j],
// F_c[2][j]), so that num / tau = q2 . (a0, a1) + a2 and the Sampson
// denominator in pixels is a0^2 + a1^2 + b0^2 + b1^2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_f_sample(const FFrame* __restrict__ fr) {
    const FFrame& a = fr[blockIdx.y];
    const int64_t g = (int64_t)blockIdx.x * 64 + threadIdx.x;
    int ns = 0;
    if (g < a.H) {
        SampleStream st(a.seed, (uint64_t)g);
        int32_t idx[kF7];
        const uint32_t n = (uint32_t)a.n;
#pragma unroll
        for (int m = 0; m < kF7; ++m) {
            const uint32_t j = n - (uint32_t)kF7 + (uint32_t)m;
            uint32_t t = st.bounded(j + 1u);
#pragma unroll
            for (int q = 0; q < m; ++q)
                if (idx[q] == (int32_t)t) t = j;
            idx[m] = (int32_t)t;
        }
        double u1[kF7], v1[kF7], u2[kF7], v2[kF7];
#pragma unroll
        for (int k = 0; k < kF7; ++k) {
            u1[k] = a.x1n[2 * idx[k]];
            v1[k] = a.x1n[2 * idx[k] + 1];
            u2[k] = a.x2n[2 * idx[k]];
            v2[k] = a.x2n[2 * idx[k] + 1];
        }
        double F2[9], D[9], roots[3];
        const int nr = f_seven_point(u1, v1, u2, v2, F2, D, roots);
        const double s1 = a.h->T1[0], s2 = a.h->T2[0];
        const bool dbg = a.dbg_samples && g < a.dbg_max;
#pragma unroll
        for (int r = 0; r < kFSolutions; ++r) {
            if (r >= nr) break;
            double G[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) G[i] = F2[i] + roots[r] * D[i];
            if (!f_unit(G)) continue;
            const int64_t k = ns++;
            double* o64 = a.m64 + 9 * (4 * g + k);
            for (int i = 0; i < 9; ++i) o64[i] = G[i];
            if (dbg)
                for (int i = 0; i < 9; ++i) a.dbg_models[27 * g + 9 * k + i] = G[i];
            double Fc[9];
#pragma unroll
            for (int rr = 0; rr < 3; ++rr)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    Fc[3 * rr + c] = G[3 * rr + c] * (rr < 2 ? s2 : 1.0) * (c < 2 ? s1 : 1.0);
            float cst[kFConst];
            for (int c = 0; c < 3; ++c) {
                cst[c] = (float)Fc[c];
                cst[3 + c] = (float)Fc[3 + c];
                cst[6 + c] = (float)(Fc[6 + c] / a.tau);
            }
            for (int j = 0; j < 2; ++j) {
                cst[9 + 3 * j] = (float)(a.tau * Fc[j]);
                cst[10 + 3 * j] = (float)(a.tau * Fc[3 + j]);
                cst[11 + 3 * j] = (float)Fc[6 + j];
            }
            cst[15] = 0.0f;
            float4* o32 = reinterpret_cast<float4*>(a.m32 + kFConst * (4 * g + k));
            for (int v = 0; v < 4; ++v)
                o32[v] = make_float4(cst[4 * v], cst[4 * v + 1], cst[4 * v + 2], cst[4 * v + 3]);
        }
        if (dbg) {
            for (int k = 0; k < kF7; ++k) a.dbg_samples[kF7 * g + k] = idx[k];
            a.dbg_nsol[g] = ns;
            for (int k = ns; k < kFSolutions; ++k)
                for (int i = 0; i < 9; ++i) a.dbg_models[27 * g + 9 * k + i] = __builtin_nan("");
        }
    }
    const uint64_t lt = threadIdx.x ? (~0ull >> (64 - threadIdx.x)) : 0ull;
    const uint64_t b0 = __ballot(ns & 1), b1 = __ballot(ns & 2);
    const uint32_t total = (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1);
    const uint32_t pre = (uint32_t)__popcll(b0 & lt) + 2u * (uint32_t)__popcll(b1 & lt);
    uint32_t base = 0;
    if (total) {
        if (threadIdx.x == 0) base = atomicAdd(&a.h->n_cand, total);
        base = __shfl(base, 0, 64);
    }
    for (int k = 0; k < ns; ++k) a.cslot[base + pre + (uint32_t)k] = 4u * (uint32_t)g + (uint32_t)k;
}

// Sums N per-lane values (N a power of two, at most 64) over the wave by
// reduce-scatter: each step trades half of the values with the partner
// lane (N - 1 shuffles in all, against 6 N for a butterfly per value), the
// last log2(64 / N) steps add the one value left; afterwards lane l holds
// the total of value l / (64 / N).  Fixed order, deterministic.
template <int N, typename T>
__device__ __forceinline__ T f_wave_reduce_scatter(T (&v)[N]) {
    const int lane = threadIdx.x & 63;
    int half = N / 2;
#pragma unroll
    for (int bit = 5; bit >= 0; --bit) {
        if (half >= 1) {
            const bool hi = (lane >> bit) & 1;
#pragma unroll
            for (int i = 0; i < N / 2; ++i) {
                if (i < half) {
                    const T a = v[i], c = v[half + i];
                    const T send = hi ? a : c;
                    const T keep = hi ? c : a;
                    v[i] = keep + __shfl_xor(send, 1 << bit, 64);
                }
            }
            half >>= 1;
        } else {
            v[0] += __shfl_xor(v[0], 1 << bit, 64);
        }
    }
    return v[0];
}

// ---------------------------------------------------------------------------
// FP32 MSAC on the Sampson error.  Workgroup (candidate block, chunk): W
// waves, each holding 64 J consecutive points of the chunk (J per lane), go
// through the block's CB candidates (constants staged in LDS), G at a time.
// Per (candidate, point): num' = num / tau and den as above; the point is an
// inlier iff num'^2 < den (exact, division-free; false for den = 0 and for
// the NaN padding), and its weight 1 - num'^2 / den is computed only for
// groups of rows holding an inlier.  A wave goes only through its groups of
// rows that hold points of the frame (the rest, padding, would add +0), and
// a wave with none writes 0.  A candidate's chunk sum is the sum of its
// waves' sums in wave order, each the reduce-scatter tree over lanes of
// per-lane sums in row order: the result does not depend on the launch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64 * kFW) void k_f_score(const FFrame* __restrict__ fr) {
    __shared__ float4 s_m[kFCB * 4];
    __shared__ float s_sum[kFW][kFCB];
    const FFrame& F = fr[blockIdx.z];
    if ((int)blockIdx.y >= F.chunks) return;  // uniform: a chunk past this frame's points
    const int ncand = (int)F.h->n_cand;
    const int cbase = blockIdx.x * kFCB;
    if (cbase >= ncand) return;
    const float* pts = F.pts;
    const int64_t np = F.np, cap = F.cap;
    const uint32_t* cand_slot = F.cslot;
    const float* models32 = F.m32;
    float* partial = F.part;
    const int cn = min(kFCB, ncand - cbase);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x < 4 * cn) {
        const uint32_t slot = cand_slot[cbase + (threadIdx.x >> 2)];
        s_m[threadIdx.x] = reinterpret_cast<const float4*>(models32)[4 * (size_t)slot + (threadIdx.x & 3)];
    }
    const int64_t p0 = (int64_t)blockIdx.y * kFChunk + (int64_t)wv * 64 * kFJ;
    // rows of this wave up to the last group holding a point (< n): uniform
    const int64_t left = F.n - p0;
    const int rows = left <= 0 ? 0
                               : (int)min<int64_t>(kFJ, (left + 64 * kFRows - 1) / (64 * kFRows) * kFRows);
    float p1u[kFJ], p1v[kFJ], q2u[kFJ], q2v[kFJ];
#pragma unroll
    for (int k = 0; k < kFJ; ++k) {
        const int64_t i = p0 + 64 * k + lane;  // < np: the chunk lies inside the padded arrays
        p1u[k] = k < rows ? pts[i] : 0.0f;
        p1v[k] = k < rows ? pts[np + i] : 0.0f;
        q2u[k] = k < rows ? pts[2 * np + i] : 0.0f;
        q2v[k] = k < rows ? pts[3 * np + i] : 0.0f;
    }
    __syncthreads();
    if (rows == 0) {
        if (lane < cn) s_sum[wv][lane] = 0.0f;
    } else {
        for (int c0 = 0; c0 < cn; c0 += kFG) {
            float acc[kFG];
#pragma unroll
            for (int j = 0; j < kFG; ++j) {
                acc[j] = 0.0f;
                const int c = c0 + j;
                if (c >= cn) continue;  // uniform
                const float4 m0 = s_m[4 * c], m1 = s_m[4 * c + 1], m2 = s_m[4 * c + 2],
                             m3 = s_m[4 * c + 3];
                // A rows: (m0.x m0.y m0.z) (m0.w m1.x m1.y) (m1.z m1.w m2.x)
                // B_0: (m2.y m2.z m2.w), B_1: (m3.x m3.y m3.z)
#pragma unroll
                for (int k0 = 0; k0 < kFJ; k0 += kFRows) {
                    if (k0 >= rows) break;  // uniform
                    float nn[kFRows], dd[kFRows];
                    uint64_t any = 0;
#pragma unroll
                    for (int r = 0; r < kFRows; ++r) {
                        const int k = k0 + r;
                        const float a0 = fmaf(m0.x, p1u[k], fmaf(m0.y, p1v[k], m0.z));
                        const float a1 = fmaf(m0.w, p1u[k], fmaf(m1.x, p1v[k], m1.y));
                        const float a2 = fmaf(m1.z, p1u[k], fmaf(m1.w, p1v[k], m2.x));
                        const float b0 = fmaf(m2.y, q2u[k], fmaf(m2.z, q2v[k], m2.w));
                        const float b1 = fmaf(m3.x, q2u[k], fmaf(m3.y, q2v[k], m3.z));
                        const float num = fmaf(q2u[k], a0, fmaf(q2v[k], a1, a2));
                        nn[r] = num * num;
                        dd[r] = fmaf(a0, a0, fmaf(a1, a1, fmaf(b0, b0, b1 * b1)));
                        any |= __ballot(nn[r] < dd[r]);
                    }
                    if (__builtin_expect(any != 0, 0)) {
                        asm volatile("");
#pragma unroll
                        for (int r = 0; r < kFRows; ++r) {
                            const float w = __builtin_amdgcn_fmed3f(
                                fmaf(-nn[r], __builtin_amdgcn_rcpf(dd[r]), 1.0f), 0.0f, 1.0f);
                            acc[j] += nn[r] < dd[r] ? w : 0.0f;
                        }
                    }
                }
            }
            // lane l: the wave's sum of candidate c0 + l / (64 / G)
            const float ws = f_wave_reduce_scatter<kFG>(acc);
            const int c = c0 + lane / (64 / kFG);
            if ((lane & (64 / kFG - 1)) == 0 && c < cn) s_sum[wv][c] = ws;
        }
    }
    __syncthreads();
    if (threadIdx.x < cn) {
        float s = s_sum[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < kFW; ++w) s += s_sum[w][threadIdx.x];
        partial[(size_t)blockIdx.y * cap + cbase + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Fold + select, one workgroup: candidate c's FP32 score (units of tau^2)
// is n - (sum of its chunk sums in chunk order); the 8 lowest
// (score, slot) keys become the leaders.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kFRed) void k_f_select(const FFrame* __restrict__ fr) {
    __shared__ uint64_t s_top[kFRed / 64][kTopK];
    const FFrame& F = fr[blockIdx.x];
    FHeader* h = F.h;
    const uint32_t* cand_slot = F.cslot;
    const float* partial = F.part;
    const int64_t cap = F.cap, n = F.n;
    const int chunks = F.chunks;
    const float tau2f = (float)(F.tau * F.tau);
    float* dbg_scores = F.dbg_scores;
    const int64_t dbg_max = F.dbg_max;
    const int ncand = (int)h->n_cand;
    uint64_t top[kTopK];
#pragma unroll
    for (int k = 0; k < kTopK; ++k) top[k] = ~0ull;
    for (int c = threadIdx.x; c < ncand; c += kFRed) {
        float s = 0.0f;
        for (int ch = 0; ch < chunks; ++ch) s += partial[(size_t)ch * cap + c];
        const float score = (float)n - s;
        const uint32_t slot = cand_slot[c];
        if (dbg_scores && (int64_t)(slot >> 2) < dbg_max)
            dbg_scores[3 * (slot >> 2) + (slot & 3)] = score * tau2f;
        uint64_t key = ((uint64_t)float_order_key(score) << 32) | slot;
        // insertion into the sorted registers
#pragma unroll
        for (int k = 0; k < kTopK; ++k) {
            if (key < top[k]) {
                const uint64_t t = top[k];
                top[k] = key;
                key = t;
            }
        }
    }
    // Two-level merge of the lanes' sorted lists: each wave takes its kTopK
    // lowest keys by kTopK rounds of a wave minimum (the lane holding the
    // minimum drops it; keys are unique, the slot is in the key), then the
    // first wave does the same over the waves' lists: one barrier.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < kTopK; ++r) {
        uint64_t v = top[0];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = __shfl_xor(v, off, 64);
            v = o < v ? o : v;
        }
        if (top[0] == v) {
#pragma unroll
            for (int k = 0; k + 1 < kTopK; ++k) top[k] = top[k + 1];
            top[kTopK - 1] = ~0ull;
        }
        if (lane == 0) s_top[wave][r] = v;
    }
    __syncthreads();
    if (wave == 0) {
        // lane l holds the sorted list of wave l (the lists are sorted: each
        // round above took the wave's next minimum).
#pragma unroll
        for (int k = 0; k < kTopK; ++k) top[k] = lane < kFRed / 64 ? s_top[lane][k] : ~0ull;
        int nl = 0;
#pragma unroll
        for (int r = 0; r < kTopK; ++r) {
            uint64_t v = top[0];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint64_t o = __shfl_xor(v, off, 64);
                v = o < v ? o : v;
            }
            if (v == ~0ull) break;
            if (top[0] == v) {
#pragma unroll
                for (int k = 0; k + 1 < kTopK; ++k) top[k] = top[k + 1];
                top[kTopK - 1] = ~0ull;
            }
            if (lane == 0) h->lead_slot[r] = (uint32_t)v;
            ++nl;
        }
        if (lane == 0) h->n_lead = nl;
    }
}

// Sampson distance squared in pixels of (x1n, x2n) under F_n (the oracle's
// operation order).
__device__ inline double f_sampson2(const double* F, double u1, double v1, double u2, double v2,
                                    double s1sq, double s2sq) {
    const double a1 = F[0] * u1 + F[1] * v1 + F[2];
    const double a2 = F[3] * u1 + F[4] * v1 + F[5];
    const double a3 = F[6] * u1 + F[7] * v1 + F[8];
    const double b1 = F[0] * u2 + F[3] * v2 + F[6];
    const double b2 = F[1] * u2 + F[4] * v2 + F[7];
    const double num = u2 * a1 + v2 * a2 + a3;
    const double den = s2sq * (a1 * a1 + a2 * a2) + s1sq * (b1 * b1 + b2 * b2);
    return num * num / den;
}

// FP64 MSAC of the leaders: per block of points, per leader, the sum of
// min(e^2, tau^2) (inliers: e^2 < tau^2) and the inlier count.
__global__ __launch_bounds__(kFPtBlock) void k_f_rerank(const FFrame* __restrict__ fr) {
    __shared__ double s_F[kTopK][9];
    __shared__ double s_acc[kFPtBlock / 64][kTopK];
    __shared__ uint32_t s_cnt[kFPtBlock / 64][kTopK];
    const FFrame& F = fr[blockIdx.y];
    if ((int)blockIdx.x >= F.rblocks) return;  // uniform: a block past this frame's points
    const FHeader* h = F.h;
    const double* x1n = F.x1n;
    const double* x2n = F.x2n;
    const int64_t n = F.n;
    const double* models64 = F.m64;
    const double tau2 = F.tau * F.tau;
    double* rr_acc = F.rra;
    uint32_t* rr_cnt = F.rrc;
    const int nl = h->n_lead;
    if (threadIdx.x < 9 * nl) {
        const int l = threadIdx.x / 9, e = threadIdx.x % 9;
        s_F[l][e] = models64[9 * (size_t)h->lead_slot[l] + e];
    }
    __syncthreads();
    const double s1sq = h->T1[0] * h->T1[0], s2sq = h->T2[0] * h->T2[0];
    const int64_t i = (int64_t)blockIdx.x * kFPtBlock + threadIdx.x;
    double u1 = 0.0, v1 = 0.0, u2 = 0.0, v2 = 0.0;
    if (i < n) {
        u1 = x1n[2 * i];
        v1 = x1n[2 * i + 1];
        u2 = x2n[2 * i];
        v2 = x2n[2 * i + 1];
    }
    for (int l = 0; l < nl; ++l) {
        double v = 0.0;
        bool in = false;
        if (i < n) {
            const double e2 = f_sampson2(s_F[l], u1, v1, u2, v2, s1sq, s2sq);
            in = e2 < tau2;
            v = in ? e2 : tau2;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        const uint32_t c = (uint32_t)__popcll(__ballot(in));
        if ((threadIdx.x & 63) == 0) {
            s_acc[threadIdx.x >> 6][l] = v;
            s_cnt[threadIdx.x >> 6][l] = c;
        }
    }
    __syncthreads();
    if (threadIdx.x < nl) {
        double s = s_acc[0][threadIdx.x];
        uint32_t c = s_cnt[0][threadIdx.x];
        for (int w = 1; w < kFPtBlock / 64; ++w) {
            s += s_acc[w][threadIdx.x];
            c += s_cnt[w][threadIdx.x];
        }
        rr_acc[(size_t)blockIdx.x * kTopK + threadIdx.x] = s;
        rr_cnt[(size_t)blockIdx.x * kTopK + threadIdx.x] = c;
    }
}

// The leaders' FP64 scores (block sums in block order) and the (score, id)
// minimum; its model becomes F_n.
__global__ __launch_bounds__(64) void k_f_choose(const FFrame* __restrict__ fr) {
    __shared__ double s_sc[kTopK];
    __shared__ uint64_t s_in[kTopK];
    const FFrame& F = fr[blockIdx.x];
    FHeader* h = F.h;
    const double* rr_acc = F.rra;
    const uint32_t* rr_cnt = F.rrc;
    const int blocks = F.rblocks;
    const double* models64 = F.m64;
    const int nl = h->n_lead;
    const int l = threadIdx.x;
    if (l < nl) {
        double s = 0.0;
        uint64_t c = 0;
        for (int b = 0; b < blocks; ++b) {
            s += rr_acc[(size_t)b * kTopK + l];
            c += rr_cnt[(size_t)b * kTopK + l];
        }
        s_sc[l] = s;
        s_in[l] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int best = -1;
        for (int k = 0; k < nl; ++k) {
            if (best < 0 || s_sc[k] < s_sc[best] ||
                (s_sc[k] == s_sc[best] && h->lead_slot[k] < h->lead_slot[best]))
                best = k;
        }
        if (best >= 0) {
            const uint32_t slot = h->lead_slot[best];
            h->best_slot = (int64_t)slot;
            h->best_score = s_sc[best];
            h->best_inliers = s_in[best];
            for (int i = 0; i < 9; ++i) h->Fn[i] = models64[9 * (size_t)slot + i];
        }
    }
}

__global__ __launch_bounds__(kFPtBlock) void k_f_mask(const FFrame* __restrict__ fr) {
    const FFrame& F = fr[blockIdx.y];
    if ((int)blockIdx.x >= F.rblocks) return;  // uniform: a block past this frame's points
    FHeader* h = F.h;
    const double* x1n = F.x1n;
    const double* x2n = F.x2n;
    const int64_t n = F.n;
    const double tau2 = F.tau * F.tau;
    uint8_t* mask = F.mask;
    const int64_t i = (int64_t)blockIdx.x * kFPtBlock + threadIdx.x;
    const bool have = h->best_slot >= 0;
    bool in = false;
    if (i < n) {
        if (have) {
            const double s1sq = h->T1[0] * h->T1[0], s2sq = h->T2[0] * h->T2[0];
            in = f_sampson2(h->Fn, x1n[2 * i], x1n[2 * i + 1], x2n[2 * i], x2n[2 * i + 1], s1sq,
                            s2sq) < tau2;
        }
        mask[i] = in ? 1 : 0;
    }
    const uint64_t b = __ballot(in);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd((unsigned long long*)&h->num_inliers, (unsigned long long)__popcll(b));
}

// ---------------------------------------------------------------------------
// IRLS refinement (Torr and Murray 1997; Hartley and Zisserman 11.4), one
// workgroup.  Per iteration: the 45 entries of the lower triangle of
// M = sum over inliers of a a^T / den(F) (a the epipolar row, den the
// Sampson denominator in pixels; points with den <= 0 skipped), reduced in a
// fixed order; thread 0 then takes the smallest eigenvector of M by 6 steps
// of inverse iteration from F (Cholesky of M + mu I, mu = 1e-12 trace M,
// non-positive pivots floored), the closest rank-2 matrix, unit norm, the
// sign of F, and stops when the change is below step_tol.  Finally F in
// pixels: T2^T F_n T1, unit norm, the largest-magnitude entry of the last
// row positive.
// ---------------------------------------------------------------------------
__device__ void f_smallest_eigvec3(const double S[9], double v[3]) {
    const double p1 = S[1] * S[1] + S[2] * S[2] + S[5] * S[5];
    const double tr = S[0] + S[4] + S[8];
    const double q = tr / 3.0;
    const double p2 = (S[0] - q) * (S[0] - q) + (S[4] - q) * (S[4] - q) + (S[8] - q) * (S[8] - q) +
                      2.0 * p1;
    const double p = sqrt(p2 / 6.0);
    double lambda = q;
    if (p > 0.0) {
        double B[9];
        for (int i = 0; i < 9; ++i) B[i] = (S[i] - ((i % 4 == 0) ? q : 0.0)) / p;
        const double r = fmin(1.0, fmax(-1.0, det3d(B) / 2.0));
        const double phi = acos(r) / 3.0;
        lambda = q + 2.0 * p * cos(phi + 2.0 * M_PI / 3.0);
    }
    const double R[3][3] = {{S[0] - lambda, S[1], S[2]},
                            {S[3], S[4] - lambda, S[5]},
                            {S[6], S[7], S[8] - lambda}};
    double best[3] = {0.0, 0.0, 1.0}, bn = -1.0;
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) {
            const double* x = R[a];
            const double* y = R[b];
            const double c[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2],
                                 x[0] * y[1] - x[1] * y[0]};
            const double nn = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
            if (nn > bn) {
                bn = nn;
                best[0] = c[0];
                best[1] = c[1];
                best[2] = c[2];
            }
        }
    const double inv = bn > 0.0 ? 1.0 / sqrt(bn) : 1.0;
    v[0] = best[0] * inv;
    v[1] = best[1] * inv;
    v[2] = best[2] * inv;
}

__device__ void f_rank2(double F[9]) {
    double S[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            S[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
    double v[3];
    f_smallest_eigvec3(S, v);
    for (int r = 0; r < 3; ++r) {
        const double fv = F[3 * r] * v[0] + F[3 * r + 1] * v[1] + F[3 * r + 2] * v[2];
        for (int c = 0; c < 3; ++c) F[3 * r + c] -= fv * v[c];
    }
    f_unit(F);
}

// Cholesky of A + mu I (A: 9 x 9, lower triangle packed row by row, 45).
// Each pivot is inverted once (inv[j] = 1 / L[j][j]) and the column and
// the substitutions multiply by it: 9 FP64 divisions per factorisation
// instead of 36, and none per solve instead of 18, on the single thread
// that takes the step.
// Every loop is unrolled, so a factor the caller keeps in a local array
// stays in registers.
__device__ __forceinline__ void f_chol9(const double* A, double mu, double (&L)[9][9],
                                        double (&inv)[9]) {
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        double d = A[j * (j + 1) / 2 + j] + mu;
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        L[j][j] = sqrt(d > 1e-300 ? d : 1e-300);
        inv[j] = 1.0 / L[j][j];
#pragma unroll
        for (int i = j + 1; i < 9; ++i) {
            double s = A[i * (i + 1) / 2 + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            L[i][j] = s * inv[j];
        }
    }
}

__device__ __forceinline__ void f_chol9_solve(const double (&L)[9][9], const double (&inv)[9],
                                              const double (&x)[9], double (&y)[9]) {
    double z[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double s = x[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= L[i][k] * z[k];
        z[i] = s * inv[i];
    }
#pragma unroll
    for (int i = 8; i >= 0; --i) {
        double s = z[i];
#pragma unroll
        for (int k = i + 1; k < 9; ++k) s -= L[k][i] * y[k];
        y[i] = s * inv[i];
    }
}

__global__ __launch_bounds__(kFIrls) void k_f_irls(const FFrame* __restrict__ fr) {
    const FFrame& Fr = fr[blockIdx.x];
    FHeader* h = Fr.h;
    const double* x1n = Fr.x1n;
    const double* x2n = Fr.x2n;
    const uint8_t* mask = Fr.mask;
    const int64_t n = Fr.n;
    const uint64_t max_it = Fr.irls_max;
    const double step_tol = Fr.step_tol;
    __shared__ double s_red[kFIrls / 64][46];
    __shared__ double s_A[45];
    __shared__ double s_F[9];
    __shared__ int s_done;
    if (h->best_slot < 0) return;  // uniform: no model, F stays zero
    const double s1sq = h->T1[0] * h->T1[0], s2sq = h->T2[0] * h->T2[0];
    if (threadIdx.x < 9) s_F[threadIdx.x] = h->Fn[threadIdx.x];
    if (threadIdx.x == 0) s_done = 0;
    __syncthreads();
    uint64_t it = 0;
    if (h->num_inliers >= 8) {
        for (; it < max_it; ++it) {
            double F[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) F[i] = s_F[i];
            double acc[45];
#pragma unroll
            for (int e = 0; e < 45; ++e) acc[e] = 0.0;
            for (int64_t i = threadIdx.x; i < n; i += kFIrls) {
                if (!mask[i]) continue;
                const double u1 = x1n[2 * i], v1 = x1n[2 * i + 1];
                const double u2 = x2n[2 * i], v2 = x2n[2 * i + 1];
                const double a1 = F[0] * u1 + F[1] * v1 + F[2];
                const double a2 = F[3] * u1 + F[4] * v1 + F[5];
                const double b1 = F[0] * u2 + F[3] * v2 + F[6];
                const double b2 = F[1] * u2 + F[4] * v2 + F[7];
                const double den = s2sq * (a1 * a1 + a2 * a2) + s1sq * (b1 * b1 + b2 * b2);
                if (!(den > 0.0)) continue;
                const double w = 1.0 / den;
                const double a[9] = {u2 * u1, u2 * v1, u2, v2 * u1, v2 * v1, v2, u1, v1, 1.0};
#pragma unroll
                for (int r = 0; r < 9; ++r) {
                    const double wr = w * a[r];
#pragma unroll
                    for (int c = 0; c <= r; ++c) acc[r * (r + 1) / 2 + c] += wr * a[c];
                }
            }
            {
                // The 45 sums over the wave: 32 of them scattered over
                // lane pairs, the other 13 (padded to 16) over lane quads.
                double lo[32], hi[16];
#pragma unroll
                for (int e = 0; e < 32; ++e) lo[e] = acc[e];
#pragma unroll
                for (int e = 0; e < 16; ++e) hi[e] = 32 + e < 45 ? acc[32 + e] : 0.0;
                const double vlo = f_wave_reduce_scatter<32>(lo);
                const double vhi = f_wave_reduce_scatter<16>(hi);
                const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
                if (!(lane & 1)) s_red[wave][lane >> 1] = vlo;
                if (!(lane & 3) && 32 + (lane >> 2) < 45) s_red[wave][32 + (lane >> 2)] = vhi;
            }
            __syncthreads();
            if (threadIdx.x < 45) {
                double s = s_red[0][threadIdx.x];
                for (int w = 1; w < kFIrls / 64; ++w) s += s_red[w][threadIdx.x];
                s_A[threadIdx.x] = s;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                double trace = 0.0;
                for (int r = 0; r < 9; ++r) trace += s_A[r * (r + 1) / 2 + r];
                const double mu = 1e-12 * trace;
                double L[9][9], inv[9];  // in registers: every index is a constant
                f_chol9(s_A, mu, L, inv);
                double f[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) f[i] = F[i];
                bool ok = true;
                for (int k = 0; k < kFInverseIterations && ok; ++k) {
                    double y[9];
                    f_chol9_solve(L, inv, f, y);
#pragma unroll
                    for (int i = 0; i < 9; ++i) f[i] = y[i];
                    ok = f_unit(f);
                }
                if (!ok) {
                    s_done = 2;  // F unchanged, this iteration not counted
                } else {
                    f_rank2(f);
                    double dot = 0.0;
                    for (int i = 0; i < 9; ++i) dot += f[i] * F[i];
                    if (dot < 0.0)
                        for (int i = 0; i < 9; ++i) f[i] = -f[i];
                    double step = 0.0;
                    for (int i = 0; i < 9; ++i) step += (f[i] - F[i]) * (f[i] - F[i]);
                    for (int i = 0; i < 9; ++i) s_F[i] = f[i];
                    if (sqrt(step) < step_tol) s_done = 1;
                }
            }
            __syncthreads();
            const int done = s_done;
            if (done == 2) break;
            if (done == 1) {
                ++it;
                break;
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        h->refinements = it;
        const double* Fn = s_F;
        for (int i = 0; i < 9; ++i) h->Fn[i] = Fn[i];
        const double s1 = h->T1[0], s2 = h->T2[0];
        double G[9], F[9];
        for (int r = 0; r < 3; ++r) {
            G[3 * r] = Fn[3 * r] * s1;
            G[3 * r + 1] = Fn[3 * r + 1] * s1;
            G[3 * r + 2] = Fn[3 * r + 2] - Fn[3 * r] * s1 * h->T1[1] - Fn[3 * r + 1] * s1 * h->T1[2];
        }
        for (int c = 0; c < 3; ++c) {
            F[c] = s2 * G[c];
            F[3 + c] = s2 * G[3 + c];
            F[6 + c] = G[6 + c] - s2 * h->T2[1] * G[c] - s2 * h->T2[2] * G[3 + c];
        }
        f_unit(F);
        int big = 6;
        for (int c = 7; c < 9; ++c)
            if (fabs(F[c]) > fabs(F[big])) big = c;
        if (F[big] < 0.0)
            for (int i = 0; i < 9; ++i) F[i] = -F[i];
        for (int i = 0; i < 9; ++i) h->F[i] = F[i];
    }
}

size_t cdiv_sz(size_t a, size_t b) { return (a + b - 1) / b; }
size_t align256(size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

// ---------------------------------------------------------------------------
// Host: tr_estimate_fundamental.
// ---------------------------------------------------------------------------
// Checks one frame as the one-frame call does; sets res->status where the
// reference's result would carry it.
static Status f_validate(const tr_pair_frame& f, tr_fundamental_result* res) {
    const tr_ransac_options* ro = &f.ransac;
    const tr_bundle_options* bo = &f.bundle;
    const int64_t n = f.num_points;
    if (!f.points1 || !f.points2) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null argument");
    if (n < kF7) {
        res->status = TR_ERR_TOO_FEW_POINTS;
        return Status::Err(TR_ERR_TOO_FEW_POINTS, "fewer than 7 correspondences");
    }
    if (n >= (int64_t(1) << 28)) return Status::Err(TR_ERR_CAPACITY, "2^28 or more correspondences");
    if (ro->max_iterations == 0 || ro->max_iterations > kFMaxSamples)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "max_iterations must be in [1, 2^22]");
    if (!(ro->max_epipolar_error > 0.0) || !std::isfinite(ro->max_epipolar_error))
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "max_epipolar_error must be positive");
    if (bo->max_iterations > 100000)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "bundle max_iterations above 100000");
    if (cdiv_sz((size_t)n, kFChunk) > 65535)
        return Status::Err(TR_ERR_CAPACITY, "more than 65535 x 2048 correspondences");
    return Status::Ok();  // finite values: checked while staging (run_fundamental)
}

Status Context::estimate_fundamental(const double* p1, const double* p2, int64_t n,
                                     const tr_ransac_options* ro, const tr_bundle_options* bo,
                                     tr_fundamental_result* res, uint8_t* mask,
                                     tr_fundamental_debug* dbg) {
    if (!p1 || !p2 || !ro || !bo || !res) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null argument");
    if (dbg && dbg->max_samples < 0) return Status::Err(TR_ERR_INVALID_ARGUMENT, "debug max_samples < 0");
    tr_pair_frame f;
    f.points1 = p1;
    f.points2 = p2;
    f.num_points = n;
    f.ransac = *ro;
    f.bundle = *bo;
    uint8_t* const masks[1] = {mask};
    return run_fundamental(&f, 1, res, masks, dbg);
}

Status Context::estimate_fundamental_batch(const tr_pair_frame* frames, int nf,
                                           tr_fundamental_result* res, uint8_t* const* masks) {
    if (!frames || !res) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null argument");
    if (nf < 1 || nf > 65535)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "num_frames must be in [1, 65535]");
    return run_fundamental(frames, nf, res, masks, nullptr);
}

// Device arena of a launch: [frame table | inputs | headers | masks] (the
// part mirrored in pinned host memory: the first two go up in one copy,
// the last two come down in one), then each frame's work arrays, then the
// stage dumps of a one-frame debug call.
Status Context::run_fundamental(const tr_pair_frame* frames, int nf, tr_fundamental_result* res,
                                uint8_t* const* masks, tr_fundamental_debug* dbg) {
    for (int f = 0; f < nf; ++f) {
        std::memset(&res[f], 0, sizeof(res[f]));
        res[f].best_hypothesis = -1;
    }
    for (int f = 0; f < nf; ++f) {
        const Status st = f_validate(frames[f], &res[f]);
        if (!st.ok())
            return nf == 1 ? st : Status::Err(st.code, "frame " + std::to_string(f) + ": " + st.msg);
    }
    const int64_t S = (dbg && nf == 1)
                          ? std::min<int64_t>(dbg->max_samples, (int64_t)frames[0].ransac.max_iterations)
                          : 0;

    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += align256(bytes);
        return o;
    };
    std::vector<FFrame> tab(nf);
    const size_t o_tab = take(sizeof(FFrame) * (size_t)nf);
    std::vector<size_t> o_in(nf), o_mask(nf);
    for (int f = 0; f < nf; ++f) o_in[f] = take(32 * (size_t)frames[f].num_points);
    const size_t o_hdr = take(sizeof(FHeader) * (size_t)nf);
    for (int f = 0; f < nf; ++f) o_mask[f] = take((size_t)frames[f].num_points);
    const size_t mirror = off;  // host staging covers [0, mirror)
    int64_t max_np = 0, max_H = 0, max_cap = 0;
    int max_chunks = 0, max_rblocks = 0;
    std::vector<size_t> o_x1(nf), o_x2(nf), o_pts(nf), o_m64(nf), o_m32(nf), o_cs(nf), o_part(nf),
        o_rra(nf), o_rrc(nf);
    for (int f = 0; f < nf; ++f) {
        const int64_t n = frames[f].num_points;
        const int64_t H = (int64_t)frames[f].ransac.max_iterations;
        const int64_t np = (int64_t)cdiv_sz((size_t)n, kFChunk) * kFChunk;
        FFrame& F = tab[f];
        F.n = n;
        F.np = np;
        F.H = H;
        F.cap = kFSolutions * H;
        F.chunks = (int32_t)(np / kFChunk);
        F.rblocks = (int32_t)cdiv_sz((size_t)n, kFPtBlock);
        F.tau = frames[f].ransac.max_epipolar_error;
        F.seed = frames[f].ransac.seed;
        F.irls_max = (uint64_t)frames[f].bundle.max_iterations;
        F.step_tol = frames[f].bundle.step_tol;
        o_x1[f] = take(16 * (size_t)n);
        o_x2[f] = take(16 * (size_t)n);
        o_pts[f] = take(16 * (size_t)np);
        o_m64[f] = take(sizeof(double) * 9 * kFSlot * (size_t)H);
        o_m32[f] = take(sizeof(float) * kFConst * kFSlot * (size_t)H);
        o_cs[f] = take(sizeof(uint32_t) * (size_t)F.cap);
        o_part[f] = take(sizeof(float) * (size_t)F.chunks * (size_t)F.cap);
        o_rra[f] = take(sizeof(double) * kTopK * (size_t)F.rblocks);
        o_rrc[f] = take(sizeof(uint32_t) * kTopK * (size_t)F.rblocks);
        max_np = std::max(max_np, np);
        max_H = std::max(max_H, H);
        max_cap = std::max(max_cap, F.cap);
        max_chunks = std::max(max_chunks, (int)F.chunks);
        max_rblocks = std::max(max_rblocks, (int)F.rblocks);
    }
    const size_t o_ds = take(sizeof(int32_t) * kF7 * (size_t)S);
    const size_t o_dn = take(sizeof(int32_t) * (size_t)S);
    const size_t o_dm = take(sizeof(double) * 27 * (size_t)S);
    const size_t o_dsc = take(sizeof(float) * 3 * (size_t)S);
    TR_HIP(hipSetDevice(device));
    TR_TRY(f_dev_.reserve(off));
    TR_TRY(f_host_.reserve(mirror));
    char* D = static_cast<char*>(f_dev_.ptr);
    char* Hs = static_cast<char*>(f_host_.ptr);
    for (int f = 0; f < nf; ++f) {
        const int64_t n = frames[f].num_points;
        FFrame& F = tab[f];
        F.h = reinterpret_cast<FHeader*>(D + o_hdr) + f;
        F.p1 = reinterpret_cast<const double*>(D + o_in[f]);
        F.p2 = reinterpret_cast<const double*>(D + o_in[f] + 16 * (size_t)n);
        F.x1n = reinterpret_cast<double*>(D + o_x1[f]);
        F.x2n = reinterpret_cast<double*>(D + o_x2[f]);
        F.pts = reinterpret_cast<float*>(D + o_pts[f]);
        F.m64 = reinterpret_cast<double*>(D + o_m64[f]);
        F.m32 = reinterpret_cast<float*>(D + o_m32[f]);
        F.cslot = reinterpret_cast<uint32_t*>(D + o_cs[f]);
        F.part = reinterpret_cast<float*>(D + o_part[f]);
        F.rra = reinterpret_cast<double*>(D + o_rra[f]);
        F.rrc = reinterpret_cast<uint32_t*>(D + o_rrc[f]);
        F.mask = reinterpret_cast<uint8_t*>(D + o_mask[f]);
        F.dbg_samples = S ? reinterpret_cast<int32_t*>(D + o_ds) : nullptr;
        F.dbg_nsol = S ? reinterpret_cast<int32_t*>(D + o_dn) : nullptr;
        F.dbg_models = S ? reinterpret_cast<double*>(D + o_dm) : nullptr;
        F.dbg_scores = S ? reinterpret_cast<float*>(D + o_dsc) : nullptr;
        F.dbg_max = S;
    }
    // The table and the inputs go up in pieces of about kFUploadPiece bytes
    // as the frames are staged, so that the copy engine works while later
    // frames are staged.  A frame found not finite ends the call after the
    // pieces already queued have been read (they read the staging buffer).
    std::memcpy(Hs + o_tab, tab.data(), sizeof(FFrame) * (size_t)nf);
    size_t sent = 0;
    for (int f = 0; f < nf; ++f) {
        const int64_t n = frames[f].num_points;
        double* dst = reinterpret_cast<double*>(Hs + o_in[f]);
        if (!copy_finite(dst, frames[f].points1, 2 * (size_t)n) ||
            !copy_finite(dst + 2 * n, frames[f].points2, 2 * (size_t)n)) {
            TR_HIP(hipStreamSynchronize(stream));
            return nf == 1 ? Status::Err(TR_ERR_NOT_FINITE, "points not finite")
                           : Status::Err(TR_ERR_NOT_FINITE,
                                         "frame " + std::to_string(f) + ": points not finite");
        }
        const size_t end = f + 1 < nf ? o_in[f + 1] : o_hdr;
        if (end - sent >= kFUploadPiece || f + 1 == nf) {
            TR_HIP(hipMemcpyAsync(D + sent, Hs + sent, end - sent, hipMemcpyHostToDevice, stream));
            sent = end;
        }
    }
    if (S) {
        // unsolved samples' debug rows read as NaN; scores of empty slots too
        TR_HIP(hipMemsetAsync(D + o_dsc, 0xff, sizeof(float) * 3 * (size_t)S, stream));
    }
    const FFrame* dt = reinterpret_cast<const FFrame*>(D + o_tab);
    const unsigned fy = (unsigned)nf;
    hipLaunchKernelGGL(k_f_hartley, dim3(2, fy), dim3(kFRed), 0, stream, dt);
    hipLaunchKernelGGL(k_f_normalise, dim3((unsigned)cdiv_sz((size_t)max_np, kFPtBlock), fy),
                       dim3(kFPtBlock), 0, stream, dt);
    hipLaunchKernelGGL(k_f_sample, dim3((unsigned)cdiv_sz((size_t)max_H, 64), fy), dim3(64), 0,
                       stream, dt);
    hipLaunchKernelGGL(k_f_score,
                       dim3((unsigned)cdiv_sz((size_t)max_cap, kFCB), (unsigned)max_chunks, fy),
                       dim3(64 * kFW), 0, stream, dt);
    hipLaunchKernelGGL(k_f_select, dim3(fy), dim3(kFRed), 0, stream, dt);
    hipLaunchKernelGGL(k_f_rerank, dim3((unsigned)max_rblocks, fy), dim3(kFPtBlock), 0, stream, dt);
    hipLaunchKernelGGL(k_f_choose, dim3(fy), dim3(64), 0, stream, dt);
    hipLaunchKernelGGL(k_f_mask, dim3((unsigned)max_rblocks, fy), dim3(kFPtBlock), 0, stream, dt);
    hipLaunchKernelGGL(k_f_irls, dim3(fy), dim3(kFIrls), 0, stream, dt);
    TR_HIP(hipGetLastError());
    TR_HIP(hipMemcpyAsync(Hs + o_hdr, D + o_hdr, mirror - o_hdr, hipMemcpyDeviceToHost, stream));
    if (S) {
        if (dbg->samples)
            TR_HIP(hipMemcpyAsync(dbg->samples, D + o_ds, sizeof(int32_t) * kF7 * (size_t)S,
                                  hipMemcpyDeviceToHost, stream));
        if (dbg->num_solutions)
            TR_HIP(hipMemcpyAsync(dbg->num_solutions, D + o_dn, sizeof(int32_t) * (size_t)S,
                                  hipMemcpyDeviceToHost, stream));
        if (dbg->models)
            TR_HIP(hipMemcpyAsync(dbg->models, D + o_dm, sizeof(double) * 27 * (size_t)S,
                                  hipMemcpyDeviceToHost, stream));
        if (dbg->scores32)
            TR_HIP(hipMemcpyAsync(dbg->scores32, D + o_dsc, sizeof(float) * 3 * (size_t)S,
                                  hipMemcpyDeviceToHost, stream));
    }
    TR_HIP(hipStreamSynchronize(stream));
    const FHeader* hh = reinterpret_cast<const FHeader*>(Hs + o_hdr);
    for (int f = 0; f < nf; ++f) {
        const int64_t n = frames[f].num_points;
        tr_fundamental_result& r = res[f];
        for (int i = 0; i < 9; ++i) r.F[i] = hh[f].F[i];
        r.iterations = frames[f].ransac.max_iterations;
        r.refinements = hh[f].refinements;
        r.num_inliers = hh[f].num_inliers;
        r.inlier_ratio = (double)hh[f].num_inliers / (double)n;
        r.model_score = hh[f].best_slot >= 0 ? hh[f].best_score : DBL_MAX;
        r.num_candidates = hh[f].n_cand;
        r.best_hypothesis = hh[f].best_slot;
        r.best_inliers = hh[f].best_slot >= 0 ? hh[f].best_inliers : 0;
        r.status = TR_OK;
        if (masks && masks[f]) std::memcpy(masks[f], Hs + o_mask[f], (size_t)n);
    }
    return Status::Ok();
}

}  // namespace tr
