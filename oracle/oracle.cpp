// oracle.cpp - scalar FP64 host oracle of the absolute-pose RANSAC path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the
// CPU-baseline leg of bench.py as the checker; the engine never uses it.
//
// A plain double-precision restatement of Module/TbrCudaPnP (the reference
// library); each function cites the reference lines it follows.  Built with
// g++ without -ffast-math and without FMA contraction, so every operation is
// one IEEE-754 double operation.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 scale(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// 3x3 row-major
struct M3 {
    double m[3][3];
};

// --------------------------------------------------------------------------
// Sampler (replaces curand XORWOW of ransac_utils.cuh:14-19 and the
// rejection loop of ransac.cuh:176-208).
// --------------------------------------------------------------------------
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

void philox(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Word stream of one minimal sample: block b of sample g is
// Philox(ctr = (g_lo, g_hi, b, 0), key = (seed_lo, seed_hi)).
struct WordStream {
    uint32_t key[2];
    uint64_t sample;
    uint32_t block = 0;
    uint32_t buf[4];
    int used = 4;
    int drawn = 0;
    uint32_t next() {
        if (used == 4) {
            const uint32_t ctr[4] = {(uint32_t)sample, (uint32_t)(sample >> 32), block, 0u};
            philox(ctr, key, buf);
            ++block;
            used = 0;
        }
        ++drawn;
        return buf[used++];
    }
};

// Words a sample may consume; past it the last candidate is accepted.
constexpr int kMaxWordsPerSample = 64;

// Lemire, "Fast random integer generation in an interval" (2019): the high
// word of x * bound, rejecting the low word below (2^32 - bound) mod bound.
template <class Next>
uint32_t bounded(Next&& next, uint32_t bound, int* drawn) {
    uint32_t x = next();
    uint64_t m = (uint64_t)x * bound;
    uint32_t lo = (uint32_t)m;
    if (lo < bound) {
        const uint32_t thr = (uint32_t)(0u - bound) % bound;
        while (lo < thr && *drawn < kMaxWordsPerSample) {
            x = next();
            m = (uint64_t)x * bound;
            lo = (uint32_t)m;
        }
    }
    return (uint32_t)(m >> 32);
}

// Floyd's algorithm for a uniformly random 3-subset of [0, n).
void sample_triple(uint64_t seed, int64_t n, uint64_t g, int32_t out[3]) {
    WordStream ws;
    ws.key[0] = (uint32_t)seed;
    ws.key[1] = (uint32_t)(seed >> 32);
    ws.sample = g;
    auto next = [&]() { return ws.next(); };
    const uint32_t un = (uint32_t)n;
    uint32_t i0 = bounded(next, un - 2u, &ws.drawn);
    uint32_t i1 = bounded(next, un - 1u, &ws.drawn);
    if (i1 == i0) i1 = un - 2u;
    uint32_t i2 = bounded(next, un, &ws.drawn);
    if (i2 == i0 || i2 == i1) i2 = un - 1u;
    out[0] = (int32_t)i0;
    out[1] = (int32_t)i1;
    out[2] = (int32_t)i2;
}

// Floyd's algorithm for a uniformly random 2-subset of [0, n).
void sample_pair(uint64_t seed, int64_t n, uint64_t g, int32_t out[2]) {
    WordStream ws;
    ws.key[0] = (uint32_t)seed;
    ws.key[1] = (uint32_t)(seed >> 32);
    ws.sample = g;
    auto next = [&]() { return ws.next(); };
    const uint32_t un = (uint32_t)n;
    uint32_t i0 = bounded(next, un - 1u, &ws.drawn);
    uint32_t i1 = bounded(next, un, &ws.drawn);
    if (i1 == i0) i1 = un - 1u;
    out[0] = (int32_t)i0;
    out[1] = (int32_t)i1;
}

// PROSAC (Chum and Matas, "Matching with PROSAC", CVPR 2005, sections 2-3)
// for samples of m = 3 over N correspondences sorted best first, with T_N =
// max_prosac samples.  T_n = T_N prod_{i<m} (n-i)/(N-i) is the expected
// number of the T_N samples drawn from the top n, T_{n+1} = T_n (n+1) /
// (n+1-m), and T'_m = 1, T'_{n+1} = T'_n + ceil(T_{n+1} - T_n).  Sample t
// (t = 1, 2, ...): if t = T'_n and n < N, n grows by one; then, if T'_n < t,
// the sample is 3 points of the top n, else the n-th point and 2 of the top
// n - 1.  The events (t where n grows) are the successive T' values, so the
// schedule is the list E_0 = 1, E_1, ...: after k events n = 3 + k and
// T'_n = E_k.  E holds the events up to t_max and the T' that follows.
std::vector<uint64_t> prosac_events(int64_t N, uint64_t max_prosac, uint64_t t_max) {
    const int m = 3;
    double Tn = (double)max_prosac;
    for (int i = 0; i < m; ++i) Tn *= (double)(m - i) / (double)(N - i);
    std::vector<uint64_t> E{1};
    int64_t n = m;
    while (n < N && E.back() <= t_max) {
        const double Tn1 = Tn * (double)(n + 1) / (double)(n + 1 - m);
        E.push_back(E.back() + (uint64_t)std::ceil(Tn1 - Tn));
        Tn = Tn1;
        ++n;
    }
    return E;
}

// The triple of global sample g (t = g + 1) under the PROSAC schedule E
// (all events up to t_max recorded; samples past max_prosac are uniform).
void sample_prosac(uint64_t seed, int64_t N, uint64_t g, uint64_t max_prosac,
                   const std::vector<uint64_t>& E, int32_t out[3]) {
    const uint64_t t = g + 1;
    if (t > max_prosac) {
        sample_triple(seed, N, g, out);
        return;
    }
    // events so far: E_k <= t for k < events recorded (E.size() - 1)
    const size_t L = E.size() - 1;
    size_t k = 0;
    while (k < L && E[k] <= t) ++k;
    const int64_t n = 3 + (int64_t)k;
    if (E[k] < t) {
        sample_triple(seed, n, g, out);
    } else {
        int32_t p[2];
        sample_pair(seed, n - 1, g, p);
        out[0] = p[0];
        out[1] = p[1];
        out[2] = (int32_t)(n - 1);
    }
}

// --------------------------------------------------------------------------
// Lambda-Twist P3P: p3p.cuh:9-222 with p3p_utils.cuh:193-422.
// --------------------------------------------------------------------------

// Threshold constants of config.h:20-22: float literals even though P3P
// runs in double (USE_FLOAT, config.h:5).
const double kEpsRoot2Real = (double)1e-6f;
const double kEpsRefineLambda = (double)1e-6f;

// One real root of s^3 + c2 s^2 + c1 s + c0; returns true when the cubic
// has a single real root (p3p_utils.cuh:393-422).
bool cubic_one_root(double c2, double c1, double c0, double* root) {
    const double a = c1 - c2 * c2 / 3.0;
    double b = (2.0 * c2 * c2 * c2 - 9.0 * c2 * c1) / 27.0 + c0;
    double c = b * b / 4.0 + a * a * a / 27.0;
    if (c != 0.0) {
        if (c > 0.0) {
            c = std::sqrt(c);
            b *= -0.5;
            *root = std::cbrt(b + c) + std::cbrt(b - c) - c2 / 3.0;
            return true;
        }
        c = 3.0 * b / (2.0 * a) * std::sqrt(-3.0 / a);
        *root = 2.0 * std::sqrt(-a / 3.0) * std::cos(std::acos(c) / 3.0) - c2 / 3.0;
    } else {
        *root = -c2 / 3.0 + ((a != 0.0) ? (3.0 * b / a) : 0.0);
    }
    return false;
}

// Real roots of t^2 + b t + c with the positive discriminant floor of
// p3p_utils.cuh:295-327 (discriminants below the threshold count as none).
bool quadratic_roots(double b, double c, double* r1, double* r2) {
    const double v = b * b - 4.0 * c;
    if (v < kEpsRoot2Real) {
        *r1 = *r2 = -0.5 * b;
        return false;
    }
    const double y = std::sqrt(v);
    if (b < 0.0) {
        *r1 = 0.5 * (-b + y);
        *r2 = 0.5 * (-b - y);
    } else {
        *r1 = 2.0 * c / (-b + y);
        *r2 = 2.0 * c / (-b - y);
    }
    return true;
}

// Five Gauss-Newton steps on the three depth constraints
// l_i^2 - 2 l_i l_j b_ij + l_j^2 = a_ij (p3p_utils.cuh:329-360).
void refine_depths(double* l1, double* l2, double* l3, double a12, double a13, double a23,
                   double b12, double b13, double b23) {
    for (int it = 0; it < 5; ++it) {
        const double u = *l1, v = *l2, w = *l3;
        const double r1 = u * u - 2.0 * u * v * b12 + v * v - a12;
        const double r2 = u * u - 2.0 * u * w * b13 + w * w - a13;
        const double r3 = v * v - 2.0 * v * w * b23 + w * w - a23;
        if (std::fabs(r1) + std::fabs(r2) + std::fabs(r3) < kEpsRefineLambda) return;
        const double x11 = u - v * b12, x12 = v - u * b12;
        const double x21 = u - w * b13, x23 = w - u * b13;
        const double x32 = v - w * b23, x33 = w - v * b23;
        const double det = 0.5 / (x11 * x23 * x32 + x12 * x21 * x33);
        *l1 += (-x23 * x32 * r1 - x12 * x33 * r2 + x12 * x23 * r3) * det;
        *l2 += (-x21 * x33 * r1 + x11 * x33 * r2 - x11 * x23 * r3) * det;
        *l3 += (x21 * x32 * r1 - x11 * x32 * r2 - x12 * x21 * r3) * det;
    }
}

// Inverse of a 3x3 matrix given by columns; identity when singular
// (p3p_utils.cuh:259-293).
M3 inverse_cols(V3 c0, V3 c1, V3 c2) {
    const double m00 = c0.x, m01 = c1.x, m02 = c2.x;
    const double m10 = c0.y, m11 = c1.y, m12 = c2.y;
    const double m20 = c0.z, m21 = c1.z, m22 = c2.z;
    const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) +
                       m02 * (m10 * m21 - m11 * m20);
    M3 r;
    if (det == 0.0) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[i][j] = (i == j) ? 1.0 : 0.0;
        return r;
    }
    r.m[0][0] = (m11 * m22 - m12 * m21) / det;
    r.m[0][1] = -(m01 * m22 - m02 * m21) / det;
    r.m[0][2] = (m01 * m12 - m02 * m11) / det;
    r.m[1][0] = -(m10 * m22 - m12 * m20) / det;
    r.m[1][1] = (m00 * m22 - m02 * m20) / det;
    r.m[1][2] = -(m00 * m12 - m02 * m10) / det;
    r.m[2][0] = (m10 * m21 - m11 * m20) / det;
    r.m[2][1] = -(m00 * m21 - m01 * m20) / det;
    r.m[2][2] = (m00 * m11 - m01 * m10) / det;
    return r;
}

// Splits the degenerate conic C into two lines p, q (p3p_utils.cuh:193-241):
// the largest diagonal entry of -adj(C) gives the cross-product vector v.
void conic_lines(const double C[3][3], V3* p, V3* q) {
    double A[3][3];
    A[0][0] = C[1][2] * C[2][1] - C[1][1] * C[2][2];
    A[1][1] = C[0][2] * C[2][0] - C[0][0] * C[2][2];
    A[2][2] = C[0][1] * C[1][0] - C[0][0] * C[1][1];
    A[0][1] = C[0][1] * C[2][2] - C[0][2] * C[2][1];
    A[0][2] = C[0][2] * C[1][1] - C[0][1] * C[1][2];
    A[1][2] = C[0][0] * C[1][2] - C[0][2] * C[1][0];
    A[1][0] = A[0][1];
    A[2][0] = A[0][2];
    A[2][1] = A[1][2];
    int col;
    if (A[0][0] > A[1][1])
        col = (A[0][0] > A[2][2]) ? 0 : 2;
    else
        col = (A[1][1] > A[2][2]) ? 1 : 2;
    const double s = std::sqrt(A[col][col]);
    const V3 v = {A[0][col] / s, A[1][col] / s, A[2][col] / s};
    double M[3][3];
    std::memcpy(M, C, sizeof(M));
    M[0][1] -= v.z;
    M[0][2] += v.y;
    M[1][2] -= v.x;
    M[1][0] += v.z;
    M[2][0] -= v.y;
    M[2][1] += v.x;
    *p = {M[0][0], M[1][0], M[2][0]};
    *q = {M[0][0], M[0][1], M[0][2]};
}

struct Pose34 {
    double R[3][3];
    double t[3];
};

void emit_pose(double d0, double d1, double d2, const V3 x[3], const V3 X[3], const M3& XXi,
               Pose34* out) {
    const V3 v1 = scale(d0, x[0]) - scale(d1, x[1]);
    const V3 v2 = scale(d0, x[0]) - scale(d2, x[2]);
    const V3 v3 = cross(v1, v2);
    const double Y[3][3] = {{v1.x, v2.x, v3.x}, {v1.y, v2.y, v3.y}, {v1.z, v2.z, v3.z}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            out->R[i][j] = Y[i][0] * XXi.m[0][j] + Y[i][1] * XXi.m[1][j] + Y[i][2] * XXi.m[2][j];
    const double xc[3] = {x[0].x, x[0].y, x[0].z};
    const double Xc[3] = {X[0].x, X[0].y, X[0].z};
    for (int i = 0; i < 3; ++i)
        out->t[i] = d0 * xc[i] -
                    (out->R[i][0] * Xc[0] + out->R[i][1] * Xc[1] + out->R[i][2] * Xc[2]);
}

int p3p(const V3 xin[3], const V3 Xin[3], Pose34 out[4]) {
    V3 x[3] = {xin[0], xin[1], xin[2]};
    V3 X[3] = {Xin[0], Xin[1], Xin[2]};
    V3 d01 = X[0] - X[1], d02 = X[0] - X[2];
    const V3 d12 = X[1] - X[2];
    double a01 = dot(d01, d01), a02 = dot(d02, d02), a12 = dot(d12, d12);

    // Reorder so that the side between points 1 and 2 is the longest
    // (p3p.cuh:34-53).
    if (a01 > a02) {
        if (a01 > a12) {
            std::swap(x[0], x[2]);
            std::swap(X[0], X[2]);
            std::swap(a01, a12);
            d01 = -d12;
            d02 = -d02;
        }
    } else if (a02 > a12) {
        std::swap(x[0], x[1]);
        std::swap(X[0], X[1]);
        std::swap(a02, a12);
        d01 = -d01;
        d02 = d12;
    }

    // Cubic of the degenerate conic (p3p.cuh:56-82).
    const double inv12 = 1.0 / a12;
    const double a = a01 * inv12, b = a02 * inv12;
    const double m01 = dot(x[0], x[1]), m02 = dot(x[0], x[2]), m12 = dot(x[1], x[2]);
    const double m12sq = -m12 * m12 + 1.0;
    const double m02sq = -1.0 + m02 * m02;
    const double m01sq = -1.0 + m01 * m01;
    const double ab = a * b, bsq = b * b, asq = a * a;
    const double m013 = -2.0 + 2.0 * m01 * m02 * m12;
    const double bsqm12sq = bsq * m12sq, asqm12sq = asq * m12sq, abm12sq = 2.0 * ab * m12sq;
    const double k3_inv = 1.0 / (bsqm12sq + b * m02sq);
    const double k2 = k3_inv * ((-1.0 + a) * m02sq + abm12sq + bsqm12sq + b * m013);
    const double k1 = k3_inv * (asqm12sq + abm12sq + a * m013 + (-1.0 + b) * m01sq);
    const double k0 = k3_inv * (asqm12sq + a * m01sq);
    double s;
    const bool single_root = cubic_one_root(k2, k1, k0, &s);

    double C[3][3];
    C[0][0] = -a + s * (1.0 - b);
    C[0][1] = -m02 * s;
    C[0][2] = a * m12 + b * m12 * s;
    C[1][0] = C[0][1];
    C[1][1] = s + 1.0;
    C[1][2] = -m01;
    C[2][0] = C[0][2];
    C[2][1] = C[1][2];
    C[2][2] = -a - b * s + 1.0;
    V3 lines[2];
    conic_lines(C, &lines[0], &lines[1]);

    const M3 XXi = inverse_cols(d01, d02, cross(d01, d02));

    int n = 0;
    for (int i = 0; i < 2; ++i) {
        const double p0 = lines[i].x, p1 = lines[i].y, p2 = lines[i].z;
        // The line is p0 d2 + p1 d0 + p2 d1 = 0; eliminate the depth with
        // the larger coefficient (p3p.cuh:118-214).
        if (std::fabs(p0) <= std::fabs(p1)) {
            const double w0 = -p0 / p1, w1 = -p2 / p1;
            const double ca = 1.0 / (w1 * w1 - b);
            const double cb = 2.0 * (b * m12 - m02 * w1 + w0 * w1) * ca;
            const double cc = (w0 * w0 - 2.0 * m02 * w0 - b + 1.0) * ca;
            double taus[2];
            if (!quadratic_roots(cb, cc, &taus[0], &taus[1])) continue;
            for (double tau : taus) {
                if (tau <= 0.0) continue;
                double d2 = std::sqrt(a12 / (tau * (tau - 2.0 * m12) + 1.0));
                double d1 = tau * d2;
                double d0 = w0 * d2 + w1 * d1;
                if (d0 < 0.0) continue;
                refine_depths(&d0, &d1, &d2, a01, a02, a12, m01, m02, m12);
                emit_pose(d0, d1, d2, x, X, XXi, &out[n++]);
            }
        } else {
            const double w0 = -p1 / p0, w1 = -p2 / p0;
            const double ca = 1.0 / (-a * w1 * w1 + 2.0 * a * m12 * w1 - a + 1.0);
            const double cb = 2.0 * (a * m12 * w0 - m01 - a * w0 * w1) * ca;
            const double cc = (1.0 - a * w0 * w0) * ca;
            double taus[2];
            if (!quadratic_roots(cb, cc, &taus[0], &taus[1])) continue;
            for (double tau : taus) {
                if (tau <= 0.0) continue;
                double d0 = std::sqrt(a01 / (tau * (tau - 2.0 * m01) + 1.0));
                double d1 = tau * d0;
                double d2 = w0 * d0 + w1 * d1;
                if (d2 < 0.0) continue;
                refine_depths(&d0, &d1, &d2, a01, a02, a12, m01, m02, m12);
                emit_pose(d0, d1, d2, x, X, XXi, &out[n++]);
            }
        }
        if (n > 0 && single_root) break;
    }
    return n;
}

V3 unit(V3 v) {
    const double nrm = std::sqrt(dot(v, v));
    return {v.x / nrm, v.y / nrm, v.z / nrm};
}

// --------------------------------------------------------------------------
// MSAC (msac.cuh:18-56 and the per-candidate total of ransac.cuh:322).
// --------------------------------------------------------------------------
double msac(const double* x, const double* X, int64_t n, const double* P, double tau2,
            int64_t* num_in) {
    double acc = 0.0;
    int64_t in = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double X0 = X[3 * i], X1 = X[3 * i + 1], X2 = X[3 * i + 2];
        const double z0 = P[0] * X0 + P[1] * X1 + P[2] * X2 + P[3];
        const double z1 = P[4] * X0 + P[5] * X1 + P[6] * X2 + P[7];
        const double z2 = P[8] * X0 + P[9] * X1 + P[10] * X2 + P[11];
        const double iz = 1.0 / z2;
        const double r0 = z0 * iz - x[2 * i];
        const double r1 = z1 * iz - x[2 * i + 1];
        const double r2 = r0 * r0 + r1 * r1;
        if (r2 < tau2 && z2 > 0.0) {
            acc += r2;
            ++in;
        }
    }
    if (num_in) *num_in = in;
    return acc + (double)(n - in) * tau2;
}

// --------------------------------------------------------------------------
// Quaternions (pose_utils.cuh:13-188), w first.
// --------------------------------------------------------------------------
void rotmat_to_quat(const double R[9], double q[4]) {
    // Shepperd's method: pick the largest of trace and diagonal entries.
    const double tr = R[0] + R[4] + R[8];
    if (tr > 0.0) {
        const double s = 0.5 / std::sqrt(tr + 1.0);
        q[0] = 0.25 / s;
        q[1] = (R[7] - R[5]) * s;
        q[2] = (R[2] - R[6]) * s;
        q[3] = (R[3] - R[1]) * s;
    } else if (R[0] > R[4] && R[0] > R[8]) {
        const double s = 2.0 * std::sqrt(1.0 + R[0] - R[4] - R[8]);
        q[0] = (R[7] - R[5]) / s;
        q[1] = 0.25 * s;
        q[2] = (R[1] + R[3]) / s;
        q[3] = (R[2] + R[6]) / s;
    } else if (R[4] > R[8]) {
        const double s = 2.0 * std::sqrt(1.0 + R[4] - R[0] - R[8]);
        q[0] = (R[2] - R[6]) / s;
        q[1] = (R[1] + R[3]) / s;
        q[2] = 0.25 * s;
        q[3] = (R[5] + R[7]) / s;
    } else {
        const double s = 2.0 * std::sqrt(1.0 + R[8] - R[0] - R[4]);
        q[0] = (R[3] - R[1]) / s;
        q[1] = (R[2] + R[6]) / s;
        q[2] = (R[5] + R[7]) / s;
        q[3] = 0.25 * s;
    }
    const double nrm = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i) q[i] /= nrm;
}

void quat_to_rotmat(const double q[4], double R[9]) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = 1.0 - 2.0 * (y * y + z * z);
    R[1] = 2.0 * (x * y - z * w);
    R[2] = 2.0 * (x * z + y * w);
    R[3] = 2.0 * (x * y + z * w);
    R[4] = 1.0 - 2.0 * (x * x + z * z);
    R[5] = 2.0 * (y * z - x * w);
    R[6] = 2.0 * (x * z - y * w);
    R[7] = 2.0 * (y * z + x * w);
    R[8] = 1.0 - 2.0 * (x * x + y * y);
}

void quat_mul(const double a[4], const double b[4], double o[4]) {
    const double r0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    const double r1 = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    const double r2 = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    const double r3 = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3;
}

// q (0, p) q^*, as two Hamilton products.
void quat_rot(const double q[4], const double p[3], double o[3]) {
    const double qp[4] = {-q[1] * p[0] - q[2] * p[1] - q[3] * p[2],
                          q[0] * p[0] + q[2] * p[2] - q[3] * p[1],
                          q[0] * p[1] + q[3] * p[0] - q[1] * p[2],
                          q[0] * p[2] + q[1] * p[1] - q[2] * p[0]};
    o[0] = -qp[0] * q[1] + qp[1] * q[0] - qp[2] * q[3] + qp[3] * q[2];
    o[1] = -qp[0] * q[2] + qp[2] * q[0] - qp[3] * q[1] + qp[1] * q[3];
    o[2] = -qp[0] * q[3] + qp[3] * q[0] - qp[1] * q[2] + qp[2] * q[1];
}

// exp of a rotation vector; Taylor series below 1e-6 rad (pose_utils.cuh:142-178).
void quat_exp(const double w[3], double q[4]) {
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double th = std::sqrt(th2);
    double re, im;
    if (th > 1e-6) {
        re = std::cos(0.5 * th);
        im = std::sin(0.5 * th) / th;
    } else {
        const double th4 = th2 * th2;
        re = 1.0 - th2 / 8.0 + th4 / 384.0;
        im = 0.5 - th2 / 48.0 + th4 / 3840.0;
        const double s = std::sqrt(re * re + im * im * th2);
        re /= s;
        im /= s;
    }
    q[0] = re;
    q[1] = im * w[0];
    q[2] = im * w[1];
    q[3] = im * w[2];
}

// --------------------------------------------------------------------------
// Robust losses (robust_loss.cuh:10-38; Huber as in PoseLib, which the
// reference's option enum names but does not implement).
// --------------------------------------------------------------------------
// TR-IRLS weight of Le and Zach ("Robust Fitting with Truncated Least
// Squares: A Bilevel Optimization Approach", 3DV 2021) in the form PoseLib
// publishes (TruncatedLossLeZach), which the reference's enum and comment
// name (utils.h:50-51), at x = r^2 / s^2: 1/2 inside the threshold; beyond
// it (z* - zbar) / rho with z* = 1, zbar = clamp(a, 0, 1) the minimiser over
// [0, 1] of (1 + mu rho) z^2 / 2 - (x + mu rho z* - rho x / 2) z, and
// rho = (2 (x - 1) + sqrt(4 (x - 1)^2 mu^2 + 2 mu (x - 1))) / mu, mu = 1/2.
// At x = 1 the published form is 0 / 0; its limit from above, 1/2, is used.
double le_zach_weight(double x) {
    if (x <= 1.0) return 0.5;
    const double mu = 0.5;
    const double d = x - 1.0;
    const double rho = (2.0 * d + std::sqrt(4.0 * d * d * mu * mu + 2.0 * mu * d)) / mu;
    const double a = (x + mu * rho * 1.0 - 0.5 * rho * x) / (1.0 + mu * rho);
    const double zbar = std::max(0.0, std::min(a, 1.0));
    return (1.0 - zbar) / rho;
}

struct Loss {
    int type;
    double s, s2, inv_s2;
    explicit Loss(int t, double scale) : type(t), s(scale), s2(scale * scale), inv_s2(1.0 / (scale * scale)) {}
    double loss(double r2) const {
        switch (type) {
            case TR_LOSS_TRIVIAL: return r2;
            case TR_LOSS_TRUNCATED:
            case TR_LOSS_TRUNCATED_LE_ZACH: return std::min(r2, s2);
            case TR_LOSS_HUBER: {
                const double r = std::sqrt(r2);
                return (r <= s) ? r2 : s * (2.0 * r - s);
            }
            default: return std::log1p(r2 * inv_s2);
        }
    }
    double weight(double r2) const {
        switch (type) {
            case TR_LOSS_TRIVIAL: return 1.0;
            case TR_LOSS_TRUNCATED: return (r2 < s2) ? 1.0 : 0.0;
            case TR_LOSS_TRUNCATED_LE_ZACH: return le_zach_weight(r2 * inv_s2);
            case TR_LOSS_HUBER: {
                const double r = std::sqrt(r2);
                return (r <= s) ? 1.0 : s / r;
            }
            default: return std::max(DBL_MIN, inv_s2 / (1.0 + r2 * inv_s2));
        }
    }
};

// Cost of jaco.cuh:31-104: points behind the camera (z < 0) contribute 0;
// term i is scaled by wts[i] (RealWeightVector_cuda, jaco_utils.cuh:29-41),
// or by 1 when wts is null (UniformWeightVector_cuda).
double lm_cost(const double* x, const double* X, int64_t n, const double* K, const Loss& L,
               const double q[4], const double t[3], const double* wts) {
    double cost = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double Z[3];
        quat_rot(q, &X[3 * i], Z);
        Z[0] += t[0]; Z[1] += t[1]; Z[2] += t[2];
        if (!(Z[2] >= 0.0)) continue;
        const double iz = 1.0 / Z[2];
        const double r0 = K[0] * (Z[0] * iz) + K[2] - x[2 * i];
        const double r1 = K[1] * (Z[1] * iz) + K[3] - x[2 * i + 1];
        cost += (wts ? wts[i] : 1.0) * L.loss(r0 * r0 + r1 * r1);
    }
    return cost;
}

// Normal equations of jaco.cuh:106-286 for the update R <- R exp([w]x),
// t <- t + R dt (jaco.cuh:288-299).  The 2x6 Jacobian of one point is
// [D (e_j x X) | D] with D = dproj/dZ * R; JtJ lower triangle, row-major 6x6.
void lm_normal(const double* x, const double* X, int64_t n, const double* K, const Loss& L,
               const double q[4], const double t[3], double JtJ[36], double Jtr[6],
               const double* wts) {
    double R[9];
    quat_to_rotmat(q, R);
    for (int i = 0; i < 36; ++i) JtJ[i] = 0.0;
    for (int i = 0; i < 6; ++i) Jtr[i] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double X0 = X[3 * i], X1 = X[3 * i + 1], X2 = X[3 * i + 2];
        const double Z0 = R[0] * X0 + R[1] * X1 + R[2] * X2 + t[0];
        const double Z1 = R[3] * X0 + R[4] * X1 + R[5] * X2 + t[1];
        const double Z2 = R[6] * X0 + R[7] * X1 + R[8] * X2 + t[2];
        const double z0 = Z0 / Z2, z1 = Z1 / Z2;
        const double r0 = K[0] * z0 + K[2] - x[2 * i];
        const double r1 = K[1] * z1 + K[3] - x[2 * i + 1];
        const double w = (wts ? wts[i] : 1.0) * L.weight(r0 * r0 + r1 * r1);
        if (!(Z2 >= 0.0) || w == 0.0) continue;
        const double iz = 1.0 / Z2;
        // dproj/dZ = (1/Z2) [[fx, 0, -fx z0], [0, fy, -fy z1]], then * R.
        const double P0[3] = {K[0] * iz, 0.0, -K[0] * z0 * iz};
        const double P1[3] = {0.0, K[1] * iz, -K[1] * z1 * iz};
        double D0[3], D1[3];
        for (int c = 0; c < 3; ++c) {
            D0[c] = P0[0] * R[c] + P0[1] * R[3 + c] + P0[2] * R[6 + c];
            D1[c] = P1[0] * R[c] + P1[1] * R[3 + c] + P1[2] * R[6 + c];
        }
        const double J0[6] = {-X2 * D0[1] + X1 * D0[2], X2 * D0[0] - X0 * D0[2],
                              -X1 * D0[0] + X0 * D0[1], D0[0], D0[1], D0[2]};
        const double J1[6] = {-X2 * D1[1] + X1 * D1[2], X2 * D1[0] - X0 * D1[2],
                              -X1 * D1[0] + X0 * D1[1], D1[0], D1[1], D1[2]};
        for (int a = 0; a < 6; ++a) {
            for (int b = 0; b <= a; ++b) JtJ[6 * a + b] += w * (J0[a] * J0[b] + J1[a] * J1[b]);
            Jtr[a] += w * (J0[a] * r0 + J1[a] * r1);
        }
    }
}

// LL^T solve reading the lower triangle (lm_utils.cuh:14-100).
void chol6(const double A[36], const double b[6], double x[6]) {
    double Lm[6][6] = {};
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k < j; ++k) s += Lm[i][k] * Lm[j][k];
            if (i == j) {
                double v = A[6 * i + i] - s;
                if (v <= 0.0) v = 1e-15;
                Lm[i][i] = std::sqrt(v);
            } else {
                Lm[i][j] = (A[6 * i + j] - s) / Lm[j][j];
            }
        }
    }
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int j = 0; j < i; ++j) s += Lm[i][j] * y[j];
        y[i] = (b[i] - s) / Lm[i][i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = 0.0;
        for (int j = i + 1; j < 6; ++j) s += Lm[j][i] * x[j];
        x[i] = (y[i] - s) / Lm[i][i];
    }
}

// Levenberg-Marquardt of lm.cuh:15-171.
uint64_t lm(const double* x, const double* X, int64_t n, const double* K,
            const tr_bundle_options& opt, double q[4], double t[3], orc_lm_stats* st,
            const double* wts) {
    const Loss L(opt.loss_type, opt.loss_scale);
    double cost = lm_cost(x, X, n, K, L, q, t, wts);
    double lambda = opt.initial_lambda;
    double JtJ[36], Jtr[6];
    bool recompute = true;
    uint64_t iters = 0, invalid = 0;
    double grad_norm = -1.0, step_norm = -1.0;
    if (st) st->initial_cost = cost;
    for (uint64_t it = 0; it < opt.max_iterations; ++it) {
        bool stop = false;
        if (recompute) {
            lm_normal(x, X, n, K, L, q, t, JtJ, Jtr, wts);
            double g = 0.0;
            for (int i = 0; i < 6; ++i) g += Jtr[i] * Jtr[i];
            grad_norm = std::sqrt(g);
            if (grad_norm < opt.gradient_tol) stop = true;
        }
        for (int i = 0; i < 6; ++i) JtJ[7 * i] += lambda;
        double sol[6];
        chol6(JtJ, Jtr, sol);
        double s2 = 0.0;
        for (int i = 0; i < 6; ++i) {
            sol[i] = -sol[i];
            s2 += sol[i] * sol[i];
        }
        step_norm = std::sqrt(s2);
        if (step_norm < opt.step_tol) stop = true;
        double dq[4], qn[4], tn[3], dt[3];
        quat_exp(sol, dq);
        quat_mul(q, dq, qn);
        quat_rot(q, sol + 3, dt);
        tn[0] = t[0] + dt[0]; tn[1] = t[1] + dt[1]; tn[2] = t[2] + dt[2];
        if (stop) break;
        const double cost_new = lm_cost(x, X, n, K, L, qn, tn, wts);
        if (cost_new < cost) {
            std::memcpy(q, qn, sizeof(qn));
            std::memcpy(t, tn, sizeof(tn));
            lambda = std::max(opt.min_lambda, lambda / 10.0);
            cost = cost_new;
            recompute = true;
        } else {
            ++invalid;
            for (int i = 0; i < 6; ++i) JtJ[7 * i] -= lambda;
            lambda = std::min(opt.max_lambda, lambda * 10.0);
            recompute = false;
        }
        ++iters;
    }
    if (st) {
        st->iterations = iters;
        st->invalid_steps = invalid;
        st->cost = cost;
        st->lambda = lambda;
        st->grad_norm = grad_norm;
        st->step_norm = step_norm;
    }
    return iters;
}

bool all_finite(const double* v, int64_t count) {
    for (int64_t i = 0; i < count; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

// Samples the stop rule asks for (turbo_ransac.cu:223-231), with the
// options honoured: ceil(log(1 - p) / log(1 - w^3) * mult).  No inlier at
// all asks for the cap; w = 1 asks for none.  A w whose cube is below the
// FP64 resolution of 1 - w^3 also asks for the cap (log(1 - w^3) = 0).
uint64_t needed_samples(double w, const tr_ransac_options& o) {
    if (w <= 0.0) return o.max_total_iterations;
    if (w >= 1.0) return 0;
    const double miss = std::log(1.0 - w * w * w);
    if (!(miss < 0.0)) return o.max_total_iterations;
    const double v = std::ceil(std::log(1.0 - o.success_prob) / miss * o.dyn_num_trials_mult);
    if (!(v < 1.8e19)) return o.max_total_iterations;
    return (uint64_t)std::max(0.0, v);
}

}  // namespace

// ==========================================================================
// C API
// ==========================================================================
extern "C" {

double orc_le_zach_weight(double x) { return le_zach_weight(x); }

uint64_t orc_needed_samples(double w, const tr_ransac_options* o) {
    return o ? needed_samples(w, *o) : 0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}

int64_t orc_bounded_from_words(const uint32_t* words, int32_t nwords, int32_t* pos,
                               uint32_t bound) {
    bool ran_out = false;
    int drawn = 0;
    auto next = [&]() -> uint32_t {
        ++drawn;
        if (*pos >= nwords) {
            ran_out = true;
            drawn = kMaxWordsPerSample;  // ends the rejection loop
            return 0u;
        }
        return words[(*pos)++];
    };
    const uint32_t v = bounded(next, bound, &drawn);
    return ran_out ? -1 : (int64_t)v;
}

void orc_sample_triples(uint64_t seed, int64_t n, int64_t first, int64_t count, int32_t* out) {
    for (int64_t s = 0; s < count; ++s) sample_triple(seed, n, (uint64_t)(first + s), out + 3 * s);
}

void orc_sample_prosac(uint64_t seed, int64_t n, uint64_t max_prosac, uint64_t t_max,
                       int64_t first, int64_t count, int32_t* out) {
    const std::vector<uint64_t> E = prosac_events(n, max_prosac, std::min(t_max, max_prosac));
    for (int64_t s = 0; s < count; ++s)
        sample_prosac(seed, n, (uint64_t)(first + s), max_prosac, E, out + 3 * s);
}

int32_t orc_p3p(const double* x, const double* X, double* poses) {
    V3 xs[3], Xs[3];
    for (int i = 0; i < 3; ++i) {
        xs[i] = unit({x[3 * i], x[3 * i + 1], x[3 * i + 2]});
        Xs[i] = {X[3 * i], X[3 * i + 1], X[3 * i + 2]};
    }
    Pose34 out[4];
    const int n = p3p(xs, Xs, out);
    for (int k = 0; k < n; ++k)
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) poses[12 * k + 4 * r + c] = out[k].R[r][c];
            poses[12 * k + 4 * r + 3] = out[k].t[r];
        }
    return n;
}

double orc_msac_score(const double* x, const double* X, int64_t n, const double* pose,
                      double tau2, int64_t* num_inliers) {
    return msac(x, X, n, pose, tau2, num_inliers);
}

void orc_rotmat_to_quat(const double* R, double* q) { rotmat_to_quat(R, q); }
void orc_quat_to_rotmat(const double* q, double* R) { quat_to_rotmat(q, R); }
void orc_quat_exp(const double* w, double* q) { quat_exp(w, q); }
void orc_quat_multiply(const double* a, const double* b, double* out) { quat_mul(a, b, out); }
void orc_quat_rotate(const double* q, const double* p, double* out) { quat_rot(q, p, out); }
void orc_solve_cholesky6(const double* A, const double* b, double* x) { chol6(A, b, x); }

int32_t orc_lm_refine(const double* x, const double* X, int64_t n, const double* K,
                      const tr_bundle_options* opt, double* q, double* t,
                      orc_lm_stats* stats) {
    return (int32_t)lm(x, X, n, K, *opt, q, t, stats, nullptr);
}

double orc_lm_cost(const double* x, const double* X, int64_t n, const double* K,
                   const tr_bundle_options* opt, const double* q, const double* t) {
    return lm_cost(x, X, n, K, Loss(opt->loss_type, opt->loss_scale), q, t, nullptr);
}

void orc_lm_normal_equations(const double* x, const double* X, int64_t n, const double* K,
                             const tr_bundle_options* opt, const double* q,
                             const double* t, double* JtJ, double* Jtr) {
    lm_normal(x, X, n, K, Loss(opt->loss_type, opt->loss_scale), q, t, JtJ, Jtr, nullptr);
}

int32_t orc_lm_refine_weighted(const double* x, const double* X, const double* wts, int64_t n,
                               const double* K, const tr_bundle_options* opt, double* q,
                               double* t, orc_lm_stats* stats) {
    return (int32_t)lm(x, X, n, K, *opt, q, t, stats, wts);
}

double orc_lm_cost_weighted(const double* x, const double* X, const double* wts, int64_t n,
                            const double* K, const tr_bundle_options* opt, const double* q,
                            const double* t) {
    return lm_cost(x, X, n, K, Loss(opt->loss_type, opt->loss_scale), q, t, wts);
}

}  // extern "C"

namespace {

// The whole per-frame path.  wts: the refinement's per-correspondence
// weights or null; init7: the initial pose (q, t) scored first when
// ropt->score_initial_model, else null.
int32_t estimate(const double* points2D, const double* points3D, int64_t n,
                 const tr_camera* camera, const tr_ransac_options* ropt,
                 const tr_bundle_options* bopt, const double* wts, const double* init7,
                 tr_pose_result* res, uint8_t* inliers, orc_trace* trace) {
    if (!points2D || !points3D || !camera || !ropt || !bopt || !res)
        return TR_ERR_INVALID_ARGUMENT;
    std::memset(res, 0, sizeof(*res));
    res->q[0] = 1.0;
    res->best_hypothesis = -1;
    if (n < 3) return res->status = TR_ERR_TOO_FEW_POINTS;
    if (n > 0x7fffffffLL || camera->model_id != 1 || ropt->max_iterations == 0 ||
        !(ropt->max_reproj_error > 0.0) || bopt->loss_type < 0 || bopt->loss_type > TR_LOSS_TRUNCATED_LE_ZACH ||
        !(bopt->loss_scale > 0.0))
        return res->status = TR_ERR_INVALID_ARGUMENT;
    if (!all_finite(camera->params, 4) || !all_finite(points2D, 2 * n) ||
        !all_finite(points3D, 3 * n) || (wts && !all_finite(wts, n)) ||
        (ropt->score_initial_model && (!init7 || !all_finite(init7, 7))))
        return res->status = TR_ERR_NOT_FINITE;
    const double fx = camera->params[0], fy = camera->params[1];
    const double cx = camera->params[2], cy = camera->params[3];
    if (!(fx > 0.0) || !(fy > 0.0)) return res->status = TR_ERR_INVALID_ARGUMENT;

    // Normalised image coordinates and threshold (entry.cu:42-49).
    const double f = 0.5 * (fx + fy);
    std::vector<double> xn(2 * n);
    for (int64_t i = 0; i < n; ++i) {
        xn[2 * i] = (points2D[2 * i] - cx) / fx;
        xn[2 * i + 1] = (points2D[2 * i + 1] - cy) / fy;
    }
    const double tau = ropt->max_reproj_error / f;
    const double tau2 = tau * tau;

    // Rounds of H minimal samples; best = lowest score, earliest sample and
    // solution on ties (ransac.cuh:259-356, turbo_ransac.cu:151-257).
    const uint64_t H = ropt->max_iterations;
    uint64_t drawn = 0, candidates = 0, best_in = 0;
    double best_score = DBL_MAX;
    int64_t best_id = -1;
    Pose34 best{};
    int64_t rounds = 0;
    // The initial model, scored first: the earliest hypothesis (id -2), so a
    // sample's pose replaces it only with a strictly lower score (utils.h:28-29).
    if (ropt->score_initial_model) {
        double Ri[9];
        quat_to_rotmat(init7, Ri);
        double P[12];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) P[4 * r + c] = Ri[3 * r + c];
            P[4 * r + 3] = init7[4 + r];
        }
        int64_t nin = 0;
        best_score = msac(xn.data(), points3D, n, P, tau2, &nin);
        best_id = -2;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) best.R[r][c] = Ri[3 * r + c];
            best.t[r] = init7[4 + r];
        }
        best_in = (uint64_t)nin;
    }
    // PROSAC schedule over the samples this frame may draw.
    std::vector<uint64_t> prosac;
    if (ropt->progressive_sampling) {
        const uint64_t cap = ropt->adaptive_stopping ? std::max(ropt->max_total_iterations, H) : H;
        prosac = prosac_events(n, ropt->max_prosac_iterations,
                               std::min(cap, ropt->max_prosac_iterations));
    }
    for (;;) {
        for (uint64_t s = 0; s < H; ++s) {
            const uint64_t g = drawn + s;
            int32_t idx[3];
            if (ropt->progressive_sampling)
                sample_prosac(ropt->seed, n, g, ropt->max_prosac_iterations, prosac, idx);
            else
                sample_triple(ropt->seed, n, g, idx);
            V3 xb[3], Xb[3];
            for (int k = 0; k < 3; ++k) {
                xb[k] = unit({xn[2 * idx[k]], xn[2 * idx[k] + 1], 1.0});
                Xb[k] = {points3D[3 * idx[k]], points3D[3 * idx[k] + 1], points3D[3 * idx[k] + 2]};
            }
            Pose34 sols[4];
            const int ns = p3p(xb, Xb, sols);
            candidates += (uint64_t)ns;
            const bool rec = trace && (int64_t)g < trace->max_samples;
            if (rec) {
                if (trace->samples) std::memcpy(trace->samples + 3 * g, idx, sizeof(idx));
                if (trace->num_solutions) trace->num_solutions[g] = ns;
                if (trace->scores)
                    for (int k = 0; k < 4; ++k) trace->scores[4 * g + k] = NAN;
            }
            for (int k = 0; k < ns; ++k) {
                double P[12];
                for (int r = 0; r < 3; ++r) {
                    for (int c = 0; c < 3; ++c) P[4 * r + c] = sols[k].R[r][c];
                    P[4 * r + 3] = sols[k].t[r];
                }
                int64_t nin = 0;
                const double sc = msac(xn.data(), points3D, n, P, tau2, &nin);
                if (rec) {
                    if (trace->poses) std::memcpy(trace->poses + 48 * g + 12 * k, P, sizeof(P));
                    if (trace->scores) trace->scores[4 * g + k] = sc;
                }
                if (sc < best_score) {
                    best_score = sc;
                    best_id = (int64_t)((g << 2) | (uint64_t)k);
                    best = sols[k];
                    best_in = (uint64_t)nin;
                }
            }
        }
        drawn += H;
        ++rounds;
        if (!ropt->adaptive_stopping) break;
        if (drawn >= ropt->max_total_iterations) break;
        const double w = (best_id != -1) ? (double)best_in / (double)n : 0.0;
        if (drawn >= ropt->min_iterations && drawn > needed_samples(w, *ropt)) break;
    }

    // Pose of the best hypothesis through its quaternion (entry.cu:77-79,
    // pose.cuh:26-28); identity when P3P produced nothing.
    double q[4] = {1.0, 0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
    if (best_id != -1) {
        const double Rb[9] = {best.R[0][0], best.R[0][1], best.R[0][2], best.R[1][0], best.R[1][1],
                              best.R[1][2], best.R[2][0], best.R[2][1], best.R[2][2]};
        rotmat_to_quat(Rb, q);
        t[0] = best.t[0]; t[1] = best.t[1]; t[2] = best.t[2];
    }
    if (trace) {
        trace->rounds = rounds;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) trace->best_pose[4 * r + c] = best.R[r][c];
            trace->best_pose[4 * r + 3] = best.t[r];
        }
        if (best_id == -1)
            for (int i = 0; i < 12; ++i) trace->best_pose[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }

    // Inlier mask of the RANSAC pose (entry.cu:12-26): strict r^2 < tau^2, z > 0.
    double R[9];
    quat_to_rotmat(q, R);
    std::vector<double> xin, Xin, win;
    xin.reserve(2 * n);
    Xin.reserve(3 * n);
    const double sc = 1.0 / f;
    uint64_t nin = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double* P = &points3D[3 * i];
        const double Z0 = R[0] * P[0] + R[1] * P[1] + R[2] * P[2] + t[0];
        const double Z1 = R[3] * P[0] + R[4] * P[1] + R[5] * P[2] + t[1];
        const double Z2 = R[6] * P[0] + R[7] * P[1] + R[8] * P[2] + t[2];
        const double r0 = Z0 / Z2 - xn[2 * i];
        const double r1 = Z1 / Z2 - xn[2 * i + 1];
        const bool in = (r0 * r0 + r1 * r1 < tau2) && (Z2 > 0.0);
        if (inliers) inliers[i] = in ? 1 : 0;
        if (in) {
            ++nin;
            // Inliers for the refinement: pixels times 1/f, not centred
            // (entry.cu:87-103).
            xin.push_back(points2D[2 * i] * sc);
            xin.push_back(points2D[2 * i + 1] * sc);
            Xin.push_back(P[0]);
            Xin.push_back(P[1]);
            Xin.push_back(P[2]);
            if (wts) win.push_back(wts[i]);
        }
    }

    // Refinement on the focal-normalised camera K / f (entry.cu:105-132),
    // loss scale in the same units.
    const double Kn[4] = {fx * sc, fy * sc, cx * sc, cy * sc};
    tr_bundle_options lo = *bopt;
    lo.loss_scale = bopt->loss_scale * sc;
    orc_lm_stats st{};
    lm(xin.data(), Xin.data(), (int64_t)nin, Kn, lo, q, t, &st, wts ? win.data() : nullptr);

    std::memcpy(res->q, q, sizeof(q));
    std::memcpy(res->t, t, sizeof(t));
    res->iterations = drawn;
    res->refinements = st.iterations;
    res->num_inliers = nin;
    res->inlier_ratio = (double)nin / (double)n;
    res->model_score = (best_id != -1) ? best_score : DBL_MAX;
    res->num_candidates = candidates;
    res->best_hypothesis = best_id;
    res->best_inliers = best_in;
    res->lm_initial_cost = st.initial_cost;
    res->lm_cost = st.cost;
    res->lm_invalid_steps = st.invalid_steps;
    res->lm_lambda = st.lambda;
    res->lm_step_norm = st.step_norm;
    res->lm_grad_norm = st.grad_norm;
    res->status = TR_OK;
    return TR_OK;
}

}  // namespace

extern "C" {

int32_t orc_estimate_absolute_pose(const double* points2D, const double* points3D, int64_t n,
                                   const tr_camera* camera, const tr_ransac_options* ropt,
                                   const tr_bundle_options* bopt, tr_pose_result* res,
                                   uint8_t* inliers, orc_trace* trace) {
    // As tr_estimate_absolute_pose: the initial pose comes in through result.
    double init7[7] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (res && ropt && ropt->score_initial_model) {
        for (int i = 0; i < 4; ++i) init7[i] = res->q[i];
        for (int i = 0; i < 3; ++i) init7[4 + i] = res->t[i];
    }
    return estimate(points2D, points3D, n, camera, ropt, bopt, nullptr, init7, res, inliers,
                    trace);
}

int32_t orc_estimate_frame(const tr_frame* fr, tr_pose_result* res, uint8_t* inliers,
                           orc_trace* trace) {
    if (!fr) return TR_ERR_INVALID_ARGUMENT;
    double init7[7];
    for (int i = 0; i < 4; ++i) init7[i] = fr->initial_q[i];
    for (int i = 0; i < 3; ++i) init7[4 + i] = fr->initial_t[i];
    return estimate(fr->points2D, fr->points3D, fr->num_points, &fr->camera, &fr->ransac,
                    &fr->bundle, fr->weights, init7, res, inliers, trace);
}

int32_t orc_estimate_many(const tr_frame* frames, int32_t num_frames, int32_t threads,
                          tr_pose_result* results) {
    return orc_estimate_many_masks(frames, num_frames, threads, results, nullptr);
}

int32_t orc_estimate_many_masks(const tr_frame* frames, int32_t num_frames, int32_t threads,
                                tr_pose_result* results, uint8_t* const* masks) {
    if (!frames || !results || num_frames < 0) return TR_ERR_INVALID_ARGUMENT;
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    auto work = [&](int tid) {
        for (int32_t f = tid; f < num_frames; f += threads) {
            orc_estimate_frame(&frames[f], &results[f], masks ? masks[f] : nullptr, nullptr);
        }
    };
    if (threads == 1) {
        work(0);
    } else {
        for (int i = 0; i < threads; ++i) pool.emplace_back(work, i);
        for (auto& th : pool) th.join();
    }
    return TR_OK;
}

void orc_pose_error(const double* R_pred, const double* t_pred, const double* R_gt,
                    const double* t_gt, double* rot_deg, double* center_cm) {
    // Camera centres C = -R^T t, rotation error R_gt R_pred^T (main.cu:141-158).
    double Cp[3], Cg[3], E[9];
    for (int i = 0; i < 3; ++i) {
        Cp[i] = -(R_pred[i] * t_pred[0] + R_pred[3 + i] * t_pred[1] + R_pred[6 + i] * t_pred[2]);
        Cg[i] = -(R_gt[i] * t_gt[0] + R_gt[3 + i] * t_gt[1] + R_gt[6 + i] * t_gt[2]);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            E[3 * i + j] = R_gt[3 * i] * R_pred[3 * j] + R_gt[3 * i + 1] * R_pred[3 * j + 1] +
                           R_gt[3 * i + 2] * R_pred[3 * j + 2];
    double qe[4];
    rotmat_to_quat(E, qe);
    const double sh = std::sqrt(qe[1] * qe[1] + qe[2] * qe[2] + qe[3] * qe[3]);
    *rot_deg = 2.0 * std::atan2(sh, std::fabs(qe[0])) * 180.0 / M_PI;
    const double d0 = Cp[0] - Cg[0], d1 = Cp[1] - Cg[1], d2 = Cp[2] - Cg[2];
    *center_cm = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) * 100.0;
}

}  // extern "C"

// ==========================================================================
// Fundamental matrix (the reference paper's Table I estimator; not in the
// reference checkout, README.md:13).  RANSAC over 7-point minimal samples
// of 2D-2D correspondences in pixels, MSAC on the Sampson error with the
// threshold RansacOptions_cuda::max_epipolar_error (utils.h:20), the mask
// of the RANSAC model, then iteratively reweighted 8-point refinement on its
// inliers.  Published methods restated:
//   * Hartley normalisation ("In defense of the eight-point algorithm",
//     PAMI 1997): per image, centroid to the origin, mean distance sqrt(2);
//   * the 7-point algorithm (Hartley and Zisserman, Multiple View Geometry,
//     2nd ed., section 11.1.2): the 2-D null space {F1, F2} of the 7 x 9
//     epipolar system (here by Householder QR of its transpose) and the
//     real roots alpha of det(F2 + alpha (F1 - F2)) = 0;
//   * the Sampson distance (H&Z eq. 11.9) as the MSAC error;
//   * Sampson-weighted iteratively reweighted least squares (Torr and
//     Murray, IJCV 1997; H&Z 11.4): the smallest eigenvector of
//     sum_i a_i a_i^T / den_i(F), then the closest rank-2 matrix.
// ==========================================================================
namespace {

constexpr int kF7 = 7;

// Floyd's algorithm for a uniformly random k-subset of [0, n): for j from
// n - k to n - 1, t uniform in [0, j]; take j when t is already in the set.
void sample_subset(uint64_t seed, int64_t n, uint64_t g, int k, int32_t* out) {
    WordStream ws;
    ws.key[0] = (uint32_t)seed;
    ws.key[1] = (uint32_t)(seed >> 32);
    ws.sample = g;
    auto next = [&]() { return ws.next(); };
    for (int m = 0; m < k; ++m) {
        const uint32_t j = (uint32_t)(n - k + m);
        uint32_t t = bounded(next, j + 1u, &ws.drawn);
        for (int q = 0; q < m; ++q)
            if (out[q] == (int32_t)t) {
                t = j;
                break;
            }
        out[m] = (int32_t)t;
    }
}

// Normalisation of one image's points: x_n = s (x - c), returns s, cx, cy.
void hartley(const double* p, int64_t n, double T[3]) {
    double cx = 0.0, cy = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        cx += p[2 * i];
        cy += p[2 * i + 1];
    }
    cx /= (double)n;
    cy /= (double)n;
    double d = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double dx = p[2 * i] - cx, dy = p[2 * i + 1] - cy;
        d += std::sqrt(dx * dx + dy * dy);
    }
    d /= (double)n;
    T[0] = d > 0.0 ? std::sqrt(2.0) / d : 1.0;
    T[1] = cx;
    T[2] = cy;
}

// Row of the epipolar system x2^T F x1 = 0 for row-major F.
void epipolar_row(double u1, double v1, double u2, double v2, double a[9]) {
    a[0] = u2 * u1; a[1] = u2 * v1; a[2] = u2;
    a[3] = v2 * u1; a[4] = v2 * v1; a[5] = v2;
    a[6] = u1;      a[7] = v1;      a[8] = 1.0;
}

double det3(const double M[9]) {
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
           M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// Cofactor matrix of a 3 x 3 matrix (row-major).
void cofactor3(const double M[9], double C[9]) {
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

// Real roots of c3 x^3 + c2 x^2 + c1 x + c0.  The depressed form of the
// monic cubic only gives the root of largest magnitude accurately (its
// other roots can lose every digit when |c3| is small), so: one root r1 by
// Cardano (one real root; the cancellation-free form of Numerical Recipes,
// 3rd ed., 5.6) or the largest-magnitude root of Viete's trigonometric form
// (three real roots), two Newton steps on it, then the quadratic factor
// x^2 + e x + f, whose discriminant decides the other two roots, of the
// deflation (x - r1)(x^2 + e x + f) from whichever of
// e = a + r1 and e = (f - b) / r1 (with f = -c / r1) rounds less, its roots
// by the stable quadratic formula (Kahan, "To solve a real cubic
// equation", 1986, deflates the same way), and two Newton steps on every
// root against the original coefficients.
int solve_cubic(double c3, double c2, double c1, double c0, double x[3]) {
    const double scale = std::max(std::max(std::fabs(c3), std::fabs(c2)),
                                  std::max(std::fabs(c1), std::fabs(c0)));
    if (!(scale > 0.0)) return 0;
    int nr = 0;
    if (std::fabs(c3) <= 1e-14 * scale) {
        // quadratic (or linear) c2 x^2 + c1 x + c0
        if (std::fabs(c2) <= 1e-14 * scale) {
            if (std::fabs(c1) <= 1e-14 * scale) return 0;
            x[0] = -c0 / c1;
            return 1;
        }
        const double disc = c1 * c1 - 4.0 * c2 * c0;
        if (disc < 0.0) return 0;
        const double sq = std::sqrt(disc);
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
        const double sd = std::sqrt(disc);
        const double A = -std::cbrt(0.5 * std::fabs(q) + sd) * (q >= 0.0 ? 1.0 : -1.0);
        const double B = A != 0.0 ? -p / (3.0 * A) : 0.0;
        r1 = A + B - a / 3.0;
    } else {
        const double r = std::sqrt(std::max(0.0, -p / 3.0));
        const double cosv = (r > 0.0) ? std::min(1.0, std::max(-1.0, -0.5 * q / (r * r * r))) : 1.0;
        const double phi = std::acos(cosv);
        r1 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double xk = 2.0 * r * std::cos((phi + 2.0 * M_PI * (double)k) / 3.0) - a / 3.0;
            if (std::fabs(xk) > std::fabs(r1)) r1 = xk;
        }
    }
    for (int it = 0; it < 2; ++it) {
        const double f = ((r1 + a) * r1 + b) * r1 + c;
        const double df = (3.0 * r1 + 2.0 * a) * r1 + b;
        if (df != 0.0) r1 -= f / df;
    }
    x[nr++] = r1;
    // The quadratic factor decides whether the other two roots are real
    // (the depressed discriminant's sign is not reliable when it cancels).
    double e, f;
    if (r1 != 0.0) {
        f = -c / r1;
        const double err_a = std::max(std::fabs(a), std::fabs(r1));
        const double err_b = (std::fabs(f) + std::fabs(b)) / std::fabs(r1);
        e = err_a <= err_b ? a + r1 : (f - b) / r1;
    } else {
        f = b;
        e = a;
    }
    const double dq = e * e - 4.0 * f;
    if (dq >= 0.0) {
        const double sq = std::sqrt(dq);
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
            const double f = ((c3 * x[i] + c2) * x[i] + c1) * x[i] + c0;
            const double df = (3.0 * c3 * x[i] + 2.0 * c2) * x[i] + c1;
            if (df != 0.0) x[i] -= f / df;
        }
    return nr;
}

// Unit Frobenius norm; returns false for a (near) zero matrix.
bool unit_frobenius(double F[9]) {
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += F[i] * F[i];
    if (!(s > 1e-300)) return false;
    const double inv = 1.0 / std::sqrt(s);
    for (int i = 0; i < 9; ++i) F[i] *= inv;
    return true;
}

// 7-point solver on normalised points u1[7], v1[7], u2[7], v2[7]: up to 3
// fundamental matrices (row-major, unit norm) in F[3][9]; returns the count.
int f7pt(const double* u1, const double* v1, const double* u2, const double* v2,
         double F[3][9]) {
    // M = A^T: column i (i < 7) is the epipolar row of correspondence i.
    double M[9][kF7];
    for (int i = 0; i < kF7; ++i) {
        double a[9];
        epipolar_row(u1[i], v1[i], u2[i], v2[i], a);
        for (int r = 0; r < 9; ++r) M[r][i] = a[r];
    }
    // Householder QR of M: reflection k zeroes M[k+1..8][k]; its vector is
    // kept in M[k..8][k] and 2 / v^T v in beta[k].
    double beta[kF7];
    for (int k = 0; k < kF7; ++k) {
        double nrm = 0.0;
        for (int r = k; r < 9; ++r) nrm += M[r][k] * M[r][k];
        nrm = std::sqrt(nrm);
        if (!(nrm > 0.0)) return 0;
        const double alpha = M[k][k] > 0.0 ? -nrm : nrm;
        M[k][k] -= alpha;
        double vtv = 0.0;
        for (int r = k; r < 9; ++r) vtv += M[r][k] * M[r][k];
        if (!(vtv > 0.0)) return 0;
        beta[k] = 2.0 / vtv;
        for (int c = k + 1; c < kF7; ++c) {
            double d = 0.0;
            for (int r = k; r < 9; ++r) d += M[r][k] * M[r][c];
            d *= beta[k];
            for (int r = k; r < 9; ++r) M[r][c] -= d * M[r][k];
        }
    }
    // Null space of A: the last two columns of Q = H_0 ... H_6.
    double N[2][9];
    for (int j = 0; j < 2; ++j) {
        for (int r = 0; r < 9; ++r) N[j][r] = (r == kF7 + j) ? 1.0 : 0.0;
        for (int k = kF7 - 1; k >= 0; --k) {
            double d = 0.0;
            for (int r = k; r < 9; ++r) d += M[r][k] * N[j][r];
            d *= beta[k];
            for (int r = k; r < 9; ++r) N[j][r] -= d * M[r][k];
        }
    }
    // det(F2 + alpha D) = det F2 + alpha <cof F2, D> + alpha^2 <F2, cof D> + alpha^3 det D
    const double* F1 = N[0];
    const double* F2 = N[1];
    double D[9];
    for (int i = 0; i < 9; ++i) D[i] = F1[i] - F2[i];
    double C2[9], CD[9];
    cofactor3(F2, C2);
    cofactor3(D, CD);
    double k1 = 0.0, k2 = 0.0;
    for (int i = 0; i < 9; ++i) {
        k1 += C2[i] * D[i];
        k2 += F2[i] * CD[i];
    }
    double roots[3];
    const int nr = solve_cubic(det3(D), k2, k1, det3(F2), roots);
    int ns = 0;
    for (int r = 0; r < nr; ++r) {
        for (int i = 0; i < 9; ++i) F[ns][i] = F2[i] + roots[r] * D[i];
        if (unit_frobenius(F[ns])) ++ns;
    }
    return ns;
}

// Sampson distance squared in pixels of the pair (x1n, x2n) (normalised
// coordinates) under F_n: num^2 / (s2^2 (a1^2 + a2^2) + s1^2 (b1^2 + b2^2))
// with a = F_n x1n and b = F_n^T x2n.
double sampson2(const double F[9], double u1, double v1, double u2, double v2, double s1sq,
                double s2sq) {
    const double a1 = F[0] * u1 + F[1] * v1 + F[2];
    const double a2 = F[3] * u1 + F[4] * v1 + F[5];
    const double a3 = F[6] * u1 + F[7] * v1 + F[8];
    const double b1 = F[0] * u2 + F[3] * v2 + F[6];
    const double b2 = F[1] * u2 + F[4] * v2 + F[7];
    const double num = u2 * a1 + v2 * a2 + a3;
    const double den = s2sq * (a1 * a1 + a2 * a2) + s1sq * (b1 * b1 + b2 * b2);
    return num * num / den;
}

// MSAC score of F_n: sum of e^2 over inliers (e^2 < tau2) plus tau2 per
// other point (a point with a zero denominator is not an inlier).
double sampson_msac(const double F[9], const double* x1n, const double* x2n, int64_t n,
                    double s1sq, double s2sq, double tau2, int64_t* nin) {
    double s = 0.0;
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double e2 = sampson2(F, x1n[2 * i], x1n[2 * i + 1], x2n[2 * i], x2n[2 * i + 1],
                                   s1sq, s2sq);
        if (e2 < tau2) {
            s += e2;
            ++k;
        } else {
            s += tau2;
        }
    }
    if (nin) *nin = k;
    return s;
}

// Smallest-eigenvalue eigenvector of a symmetric 3 x 3 matrix (row-major):
// the eigenvalue from the characteristic cubic (Viete), the vector as the
// longest cross product of two rows of S - lambda I.
void smallest_eigvec3(const double S[9], double v[3]) {
    const double p1 = S[1] * S[1] + S[2] * S[2] + S[5] * S[5];
    const double tr = S[0] + S[4] + S[8];
    const double q = tr / 3.0;
    const double p2 = (S[0] - q) * (S[0] - q) + (S[4] - q) * (S[4] - q) + (S[8] - q) * (S[8] - q) +
                      2.0 * p1;
    const double p = std::sqrt(p2 / 6.0);
    double lambda = q;
    if (p > 0.0) {
        double B[9];
        for (int i = 0; i < 9; ++i) B[i] = (S[i] - ((i % 4 == 0) ? q : 0.0)) / p;
        const double r = std::min(1.0, std::max(-1.0, det3(B) / 2.0));
        const double phi = std::acos(r) / 3.0;
        lambda = q + 2.0 * p * std::cos(phi + 2.0 * M_PI / 3.0);  // the smallest
    }
    const double R0[3] = {S[0] - lambda, S[1], S[2]};
    const double R1[3] = {S[3], S[4] - lambda, S[5]};
    const double R2[3] = {S[6], S[7], S[8] - lambda};
    const double* rows[3] = {R0, R1, R2};
    double best[3] = {0.0, 0.0, 1.0}, bn = -1.0;
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) {
            const double* x = rows[a];
            const double* y = rows[b];
            const double c[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2],
                                 x[0] * y[1] - x[1] * y[0]};
            const double nn = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
            if (nn > bn) {
                bn = nn;
                best[0] = c[0]; best[1] = c[1]; best[2] = c[2];
            }
        }
    const double inv = bn > 0.0 ? 1.0 / std::sqrt(bn) : 1.0;
    v[0] = best[0] * inv;
    v[1] = best[1] * inv;
    v[2] = best[2] * inv;
}

// The closest rank-2 matrix (Frobenius): F (I - v v^T) with v the right
// singular vector of the smallest singular value; then unit norm.
void rank2(double F[9]) {
    double S[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            S[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
    double v[3];
    smallest_eigvec3(S, v);
    for (int r = 0; r < 3; ++r) {
        const double fv = F[3 * r] * v[0] + F[3 * r + 1] * v[1] + F[3 * r + 2] * v[2];
        for (int c = 0; c < 3; ++c) F[3 * r + c] -= fv * v[c];
    }
    unit_frobenius(F);
}

// Solves (A + mu I) y = x by Cholesky for a symmetric positive semidefinite
// 9 x 9 A (row-major, lower triangle read); non-positive pivots floored.
void chol9_solve(const double A[81], double mu, const double x[9], double y[9]) {
    double L[9][9] = {};
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = A[9 * i + j] + (i == j ? mu : 0.0);
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            if (i == j) {
                L[i][i] = std::sqrt(s > 1e-300 ? s : 1e-300);
            } else {
                L[i][j] = s / L[j][j];
            }
        }
    double z[9];
    for (int i = 0; i < 9; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s -= L[i][k] * z[k];
        z[i] = s / L[i][i];
    }
    for (int i = 8; i >= 0; --i) {
        double s = z[i];
        for (int k = i + 1; k < 9; ++k) s -= L[k][i] * y[k];
        y[i] = s / L[i][i];
    }
}

constexpr int kFInverseIterations = 6;

// IRLS refinement of F_n (unit norm, rank 2) on inliers x1n, x2n: per
// iteration, M = sum_i a_i a_i^T / den_i(F) (den_i the Sampson denominator
// in pixels), f = the smallest eigenvector of M by inverse iteration from
// the current F (shift mu = 1e-12 trace M), then rank 2, unit norm and the
// sign of the current F; stops after max_it iterations or when the change
// |F' - F| is below step_tol.  Returns the iterations taken.
uint64_t f_irls(const double* x1n, const double* x2n, int64_t n, double s1sq, double s2sq,
                uint64_t max_it, double step_tol, double F[9]) {
    if (n < 8) return 0;
    uint64_t it = 0;
    for (; it < max_it; ++it) {
        double A[81] = {};
        for (int64_t i = 0; i < n; ++i) {
            const double u1 = x1n[2 * i], v1 = x1n[2 * i + 1];
            const double u2 = x2n[2 * i], v2 = x2n[2 * i + 1];
            const double a1 = F[0] * u1 + F[1] * v1 + F[2];
            const double a2 = F[3] * u1 + F[4] * v1 + F[5];
            const double b1 = F[0] * u2 + F[3] * v2 + F[6];
            const double b2 = F[1] * u2 + F[4] * v2 + F[7];
            const double den = s2sq * (a1 * a1 + a2 * a2) + s1sq * (b1 * b1 + b2 * b2);
            if (!(den > 0.0)) continue;
            const double w = 1.0 / den;
            double a[9];
            epipolar_row(u1, v1, u2, v2, a);
            for (int r = 0; r < 9; ++r)
                for (int c = 0; c <= r; ++c) A[9 * r + c] += w * a[r] * a[c];
        }
        double trace = 0.0;
        for (int r = 0; r < 9; ++r) trace += A[10 * r];
        const double mu = 1e-12 * trace;
        double f[9];
        for (int i = 0; i < 9; ++i) f[i] = F[i];
        for (int k = 0; k < kFInverseIterations; ++k) {
            double y[9];
            chol9_solve(A, mu, f, y);
            for (int i = 0; i < 9; ++i) f[i] = y[i];
            if (!unit_frobenius(f)) return it;
        }
        rank2(f);
        double dot = 0.0;
        for (int i = 0; i < 9; ++i) dot += f[i] * F[i];
        if (dot < 0.0)
            for (int i = 0; i < 9; ++i) f[i] = -f[i];
        double step = 0.0;
        for (int i = 0; i < 9; ++i) step += (f[i] - F[i]) * (f[i] - F[i]);
        for (int i = 0; i < 9; ++i) F[i] = f[i];
        if (std::sqrt(step) < step_tol) {
            ++it;
            break;
        }
    }
    return it;
}

// F in pixels from F_n: F = T2^T F_n T1 with T = [[s, 0, -s cx], [0, s, -s cy],
// [0, 0, 1]]; unit norm, sign so that the largest-magnitude entry of the
// last row is positive (the last row always has one: F x1 = 0 only for the
// epipole).
void f_denormalise(const double Fn[9], const double T1[3], const double T2[3], double F[9]) {
    const double s1 = T1[0], s2 = T2[0];
    double G[9];  // F_n T1
    for (int r = 0; r < 3; ++r) {
        G[3 * r] = Fn[3 * r] * s1;
        G[3 * r + 1] = Fn[3 * r + 1] * s1;
        G[3 * r + 2] = Fn[3 * r + 2] - Fn[3 * r] * s1 * T1[1] - Fn[3 * r + 1] * s1 * T1[2];
    }
    for (int c = 0; c < 3; ++c) {  // T2^T G
        F[c] = s2 * G[c];
        F[3 + c] = s2 * G[3 + c];
        F[6 + c] = G[6 + c] - s2 * T2[1] * G[c] - s2 * T2[2] * G[3 + c];
    }
    unit_frobenius(F);
    int big = 6;
    for (int c = 7; c < 9; ++c)
        if (std::fabs(F[c]) > std::fabs(F[big])) big = c;
    if (F[big] < 0.0)
        for (int i = 0; i < 9; ++i) F[i] = -F[i];
}

int32_t estimate_fundamental(const double* p1, const double* p2, int64_t n,
                             const tr_ransac_options* ro, const tr_bundle_options* bo,
                             tr_fundamental_result* res, uint8_t* inliers, orc_ftrace* trace) {
    std::memset(res, 0, sizeof(*res));
    res->best_hypothesis = -1;
    if (n < kF7) return res->status = TR_ERR_TOO_FEW_POINTS;
    if (ro->max_iterations == 0 || !(ro->max_epipolar_error > 0.0))
        return res->status = TR_ERR_INVALID_ARGUMENT;
    double T1[3], T2[3];
    hartley(p1, n, T1);
    hartley(p2, n, T2);
    std::vector<double> x1n(2 * n), x2n(2 * n);
    for (int64_t i = 0; i < n; ++i) {
        x1n[2 * i] = T1[0] * (p1[2 * i] - T1[1]);
        x1n[2 * i + 1] = T1[0] * (p1[2 * i + 1] - T1[2]);
        x2n[2 * i] = T2[0] * (p2[2 * i] - T2[1]);
        x2n[2 * i + 1] = T2[0] * (p2[2 * i + 1] - T2[2]);
    }
    const double s1sq = T1[0] * T1[0], s2sq = T2[0] * T2[0];
    const double tau2 = ro->max_epipolar_error * ro->max_epipolar_error;
    const uint64_t H = ro->max_iterations;
    double best_score = DBL_MAX, best[9] = {};
    int64_t best_id = -1, best_in = 0;
    uint64_t cands = 0;
    for (uint64_t g = 0; g < H; ++g) {
        int32_t idx[kF7];
        sample_subset(ro->seed, n, g, kF7, idx);
        double u1[kF7], v1[kF7], u2[kF7], v2[kF7];
        for (int k = 0; k < kF7; ++k) {
            u1[k] = x1n[2 * idx[k]];
            v1[k] = x1n[2 * idx[k] + 1];
            u2[k] = x2n[2 * idx[k]];
            v2[k] = x2n[2 * idx[k] + 1];
        }
        double F[3][9];
        const int ns = f7pt(u1, v1, u2, v2, F);
        cands += (uint64_t)ns;
        const bool rec = trace && (int64_t)g < trace->max_samples;
        if (rec) {
            if (trace->samples) std::memcpy(trace->samples + kF7 * g, idx, sizeof(idx));
            if (trace->num_solutions) trace->num_solutions[g] = ns;
            if (trace->scores)
                for (int k = 0; k < 3; ++k) trace->scores[3 * g + k] = NAN;
        }
        for (int k = 0; k < ns; ++k) {
            int64_t nin = 0;
            const double sc = sampson_msac(F[k], x1n.data(), x2n.data(), n, s1sq, s2sq, tau2, &nin);
            if (rec) {
                if (trace->models) std::memcpy(trace->models + 27 * g + 9 * k, F[k], sizeof(F[k]));
                if (trace->scores) trace->scores[3 * g + k] = sc;
            }
            if (sc < best_score) {
                best_score = sc;
                best_id = (int64_t)((g << 2) | (uint64_t)k);
                std::memcpy(best, F[k], sizeof(best));
                best_in = nin;
            }
        }
    }
    // Mask of the RANSAC model (strict e^2 < tau^2), inliers for the refinement.
    std::vector<double> i1, i2;
    int64_t nin = 0;
    for (int64_t i = 0; i < n; ++i) {
        const bool in = best_id != -1 && sampson2(best, x1n[2 * i], x1n[2 * i + 1], x2n[2 * i],
                                                  x2n[2 * i + 1], s1sq, s2sq) < tau2;
        if (inliers) inliers[i] = in ? 1 : 0;
        if (in) {
            ++nin;
            i1.push_back(x1n[2 * i]);
            i1.push_back(x1n[2 * i + 1]);
            i2.push_back(x2n[2 * i]);
            i2.push_back(x2n[2 * i + 1]);
        }
    }
    if (trace) std::memcpy(trace->best_model, best, sizeof(best));
    double Fn[9];
    std::memcpy(Fn, best, sizeof(Fn));
    uint64_t refinements = 0;
    if (best_id != -1)
        refinements = f_irls(i1.data(), i2.data(), nin, s1sq, s2sq, bo->max_iterations,
                             bo->step_tol, Fn);
    if (best_id != -1) {
        f_denormalise(Fn, T1, T2, res->F);
    }
    res->iterations = H;
    res->refinements = refinements;
    res->num_inliers = (uint64_t)nin;
    res->inlier_ratio = (double)nin / (double)n;
    res->model_score = best_id != -1 ? best_score : DBL_MAX;
    res->num_candidates = cands;
    res->best_hypothesis = best_id;
    res->best_inliers = (uint64_t)best_in;
    res->status = TR_OK;
    return TR_OK;
}

}  // namespace

extern "C" {

void orc_sample_subset(uint64_t seed, int64_t n, int32_t k, int64_t first, int64_t count,
                       int32_t* out) {
    for (int64_t s = 0; s < count; ++s) sample_subset(seed, n, (uint64_t)(first + s), k, out + k * s);
}

void orc_hartley(const double* points, int64_t n, double* T) { hartley(points, n, T); }

int32_t orc_fundamental_7pt(const double* x1n, const double* x2n, double* F) {
    double u1[kF7], v1[kF7], u2[kF7], v2[kF7];
    for (int i = 0; i < kF7; ++i) {
        u1[i] = x1n[2 * i];
        v1[i] = x1n[2 * i + 1];
        u2[i] = x2n[2 * i];
        v2[i] = x2n[2 * i + 1];
    }
    double Fs[3][9];
    const int ns = f7pt(u1, v1, u2, v2, Fs);
    std::memcpy(F, Fs, sizeof(double) * 9 * ns);
    return ns;
}

int32_t orc_solve_cubic(double c3, double c2, double c1, double c0, double* roots) {
    return solve_cubic(c3, c2, c1, c0, roots);
}

double orc_sampson_msac(const double* F, const double* x1n, const double* x2n, int64_t n,
                        double s1sq, double s2sq, double tau2, int64_t* num_inliers) {
    return sampson_msac(F, x1n, x2n, n, s1sq, s2sq, tau2, num_inliers);
}

uint64_t orc_fundamental_irls(const double* x1n, const double* x2n, int64_t n, double s1sq,
                              double s2sq, uint64_t max_iterations, double step_tol, double* F) {
    return f_irls(x1n, x2n, n, s1sq, s2sq, max_iterations, step_tol, F);
}

void orc_fundamental_rank2(double* F) { rank2(F); }

void orc_fundamental_denormalise(const double* Fn, const double* T1, const double* T2,
                                 double* F) {
    f_denormalise(Fn, T1, T2, F);
}

int32_t orc_estimate_fundamental(const double* points1, const double* points2, int64_t n,
                                 const tr_ransac_options* ransac_opt,
                                 const tr_bundle_options* bundle_opt,
                                 tr_fundamental_result* result, uint8_t* inliers,
                                 orc_ftrace* trace) {
    if (!points1 || !points2 || !ransac_opt || !bundle_opt || !result)
        return TR_ERR_INVALID_ARGUMENT;
    return estimate_fundamental(points1, points2, n, ransac_opt, bundle_opt, result, inliers,
                                trace);
}

}  // extern "C"
