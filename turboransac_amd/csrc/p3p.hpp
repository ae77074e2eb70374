// p3p.hpp - Lambda-Twist P3P in FP64, one minimal sample per lane.
//
// Persson & Nordberg, "Lambda Twist: An Accurate Fast Robust Perspective
// Three Point (P3P) Solver", ECCV 2018, with the branch structure and
// thresholds of the reference implementation (p3p.cuh:9-222,
// p3p_utils.cuh:193-422): one real root of the cubic, the degenerate conic
// split into two lines, two quadratics, five Gauss-Newton depth steps.
// Here every lane of a wavefront solves its own sample; the reference runs
// the solver on one thread of a 384-thread block.
#pragma once

#include <hip/hip_runtime.h>

namespace tr {

struct D3 {
    double x, y, z;
};
__device__ inline D3 d3sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 d3neg(D3 a) { return {-a.x, -a.y, -a.z}; }
__device__ inline double d3dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 d3cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline D3 d3axpy(double s, D3 a, double r, D3 b) {  // s a - r b
    return {s * a.x - r * b.x, s * a.y - r * b.y, s * a.z - r * b.z};
}

// Thresholds: float literals in the reference's single-precision build
// (config.h:20-22), applied to double arithmetic.
#define TR_EPS_ROOT2REAL ((double)1e-6f)
#define TR_EPS_REFINE ((double)1e-6f)

__device__ inline bool p3p_cubic_root(double c2, double c1, double c0, double* root) {
    const double a = c1 - c2 * c2 / 3.0;
    double b = (2.0 * c2 * c2 * c2 - 9.0 * c2 * c1) / 27.0 + c0;
    double c = b * b / 4.0 + a * a * a / 27.0;
    if (c != 0.0) {
        if (c > 0.0) {
            c = sqrt(c);
            b *= -0.5;
            *root = cbrt(b + c) + cbrt(b - c) - c2 / 3.0;
            return true;
        }
        c = 3.0 * b / (2.0 * a) * sqrt(-3.0 / a);
        *root = 2.0 * sqrt(-a / 3.0) * cos(acos(c) / 3.0) - c2 / 3.0;
        return false;
    }
    *root = -c2 / 3.0 + ((a != 0.0) ? (3.0 * b / a) : 0.0);
    return false;
}

__device__ inline bool p3p_quadratic(double b, double c, double* r1, double* r2) {
    const double v = b * b - 4.0 * c;
    if (v < TR_EPS_ROOT2REAL) {
        *r1 = *r2 = -0.5 * b;
        return false;
    }
    const double y = sqrt(v);
    if (b < 0.0) {
        *r1 = 0.5 * (-b + y);
        *r2 = 0.5 * (-b - y);
    } else {
        *r1 = 2.0 * c / (-b + y);
        *r2 = 2.0 * c / (-b - y);
    }
    return true;
}

__device__ inline void p3p_refine(double* l1, double* l2, double* l3, double a12, double a13,
                                  double a23, double b12, double b13, double b23) {
    for (int it = 0; it < 5; ++it) {
        const double u = *l1, v = *l2, w = *l3;
        const double r1 = u * u - 2.0 * u * v * b12 + v * v - a12;
        const double r2 = u * u - 2.0 * u * w * b13 + w * w - a13;
        const double r3 = v * v - 2.0 * v * w * b23 + w * w - a23;
        if (fabs(r1) + fabs(r2) + fabs(r3) < TR_EPS_REFINE) return;
        const double x11 = u - v * b12, x12 = v - u * b12;
        const double x21 = u - w * b13, x23 = w - u * b13;
        const double x32 = v - w * b23, x33 = w - v * b23;
        const double det = 0.5 / (x11 * x23 * x32 + x12 * x21 * x33);
        *l1 += (-x23 * x32 * r1 - x12 * x33 * r2 + x12 * x23 * r3) * det;
        *l2 += (-x21 * x33 * r1 + x11 * x33 * r2 - x11 * x23 * r3) * det;
        *l3 += (x21 * x32 * r1 - x11 * x32 * r2 - x12 * x21 * r3) * det;
    }
}

struct P3PSetup {
    D3 x0, x1, x2, X0;
    double inv[9];  // inverse of [X01 X02 X01xX02], row-major
    double a01, a02, a12, m01, m02, m12, a, b;
};

// Writes R and t as a row-major 3x4 [R | t] to out[0..11], and in FP32 to
// out32[0..11] with rows 0 and 1 scaled by s01 (the scoring kernel's units).
__device__ inline void p3p_emit(const P3PSetup& S, double d0, double d1, double d2, double* out,
                                float* out32, double s01) {
    const D3 v1 = d3axpy(d0, S.x0, d1, S.x1);
    const D3 v2 = d3axpy(d0, S.x0, d2, S.x2);
    const D3 v3 = d3cross(v1, v2);
    const double Y[9] = {v1.x, v2.x, v3.x, v1.y, v2.y, v3.y, v1.z, v2.z, v3.z};
    const double xc[3] = {S.x0.x, S.x0.y, S.x0.z};
    const double Xc[3] = {S.X0.x, S.X0.y, S.X0.z};
    double R[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
            R[3 * i + j] =
                Y[3 * i] * S.inv[j] + Y[3 * i + 1] * S.inv[3 + j] + Y[3 * i + 2] * S.inv[6 + j];
        const double row[4] = {
            R[3 * i], R[3 * i + 1], R[3 * i + 2],
            d0 * xc[i] - (R[3 * i] * Xc[0] + R[3 * i + 1] * Xc[1] + R[3 * i + 2] * Xc[2])};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            out[4 * i + k] = row[k];
            out32[4 * i + k] = (float)(i < 2 ? row[k] * s01 : row[k]);
        }
    }
}

// The solutions of one sample before they are written: the depths of the
// three points along their bearings at fixed positions 2 i + k (line i of
// the degenerate conic, root k of its quadratic), so that no register array
// is indexed by the dynamic solution count; valid[] marks the ones that
// exist.  Their order is the solution order.
struct P3PSolutions {
    P3PSetup S;
    double d[4][3];
    bool valid[4];
};

// Writes the valid solutions in order: the j-th as a row-major [R | t] to
// out + 12 j and its FP32 scoring copy to out32 + 12 j (see p3p_emit).
__device__ inline void p3p_emit_all(const P3PSolutions& P, double* out, float* out32,
                                    double s01) {
    int j = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        if (P.valid[p]) {
            p3p_emit(P.S, P.d[p][0], P.d[p][1], P.d[p][2], out + 12 * j, out32 + 12 * j, s01);
            ++j;
        }
    }
}

// Solves P3P for unit bearings x0..x2 and world points X0..X2 into P;
// returns the count of solutions (written later by p3p_emit_all, once the
// caller knows where their FP32 copies go).
__device__ inline int p3p_solve(D3 x0, D3 x1, D3 x2, D3 X0, D3 X1, D3 X2, P3PSolutions& P) {
    D3 d01 = d3sub(X0, X1), d02 = d3sub(X0, X2);
    const D3 d12 = d3sub(X1, X2);
    double a01 = d3dot(d01, d01), a02 = d3dot(d02, d02), a12 = d3dot(d12, d12);
    // Make the side between points 1 and 2 the longest.
    if (a01 > a02) {
        if (a01 > a12) {
            D3 t = x0; x0 = x2; x2 = t;
            t = X0; X0 = X2; X2 = t;
            double s = a01; a01 = a12; a12 = s;
            d01 = d3neg(d12);
            d02 = d3neg(d02);
        }
    } else if (a02 > a12) {
        D3 t = x0; x0 = x1; x1 = t;
        t = X0; X0 = X1; X1 = t;
        double s = a02; a02 = a12; a12 = s;
        d01 = d3neg(d01);
        d02 = d12;
    }
    P3PSetup& S = P.S;
    S.x0 = x0; S.x1 = x1; S.x2 = x2; S.X0 = X0;
    S.a01 = a01; S.a02 = a02; S.a12 = a12;
    const double inv12 = 1.0 / a12;
    const double a = a01 * inv12, b = a02 * inv12;
    const double m01 = d3dot(x0, x1), m02 = d3dot(x0, x2), m12 = d3dot(x1, x2);
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
    const bool single_root = p3p_cubic_root(k2, k1, k0, &s);

    // Degenerate conic C = C(s) and its split into two lines.
    const double C00 = -a + s * (1.0 - b), C01 = -m02 * s, C02 = a * m12 + b * m12 * s;
    const double C11 = s + 1.0, C12 = -m01, C22 = -a - b * s + 1.0;
    const double A00 = C12 * C12 - C11 * C22;
    const double A11 = C02 * C02 - C00 * C22;
    const double A22 = C01 * C01 - C00 * C11;
    const double A01 = C01 * C22 - C02 * C12;
    const double A02 = C02 * C11 - C01 * C12;
    const double A12 = C00 * C12 - C02 * C01;
    // v = the column of the largest diagonal entry over its square root; only
    // its y and z components enter the two lines.
    double vy, vz, vd;
    if (A00 > A11 && A00 > A22) {
        vd = sqrt(A00); vy = A01 / vd; vz = A02 / vd;
    } else if (!(A00 > A11) && A11 > A22) {
        vd = sqrt(A11); vy = A11 / vd; vz = A12 / vd;
    } else {
        vd = sqrt(A22); vy = A12 / vd; vz = A22 / vd;
    }
    // p = first column, q = first row of C + [v]_x.
    const double lp[2][3] = {{C00, C01 + vz, C02 - vy}, {C00, C01 - vz, C02 + vy}};

    // Inverse of [X01 X02 X01 x X02] (identity when singular).
    {
        const D3 c2 = d3cross(d01, d02);
        const double m00 = d01.x, m01_ = d02.x, m02_ = c2.x;
        const double m10 = d01.y, m11 = d02.y, m12_ = c2.y;
        const double m20 = d01.z, m21 = d02.z, m22 = c2.z;
        const double det = m00 * (m11 * m22 - m12_ * m21) - m01_ * (m10 * m22 - m12_ * m20) +
                           m02_ * (m10 * m21 - m11 * m20);
        if (det == 0.0) {
#pragma unroll
            for (int i = 0; i < 9; ++i) S.inv[i] = (i % 4 == 0) ? 1.0 : 0.0;
        } else {
            S.inv[0] = (m11 * m22 - m12_ * m21) / det;
            S.inv[1] = -(m01_ * m22 - m02_ * m21) / det;
            S.inv[2] = (m01_ * m12_ - m02_ * m11) / det;
            S.inv[3] = -(m10 * m22 - m12_ * m20) / det;
            S.inv[4] = (m00 * m22 - m02_ * m20) / det;
            S.inv[5] = -(m00 * m12_ - m02_ * m10) / det;
            S.inv[6] = (m10 * m21 - m11 * m20) / det;
            S.inv[7] = -(m00 * m21 - m01_ * m20) / det;
            S.inv[8] = (m00 * m11 - m01_ * m10) / det;
        }
    }

    int n = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p) P.valid[p] = false;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double p0 = lp[i][0], p1 = lp[i][1], p2 = lp[i][2];
        // Line p0 d2 + p1 d0 + p2 d1 = 0.
        const bool elim_d0 = fabs(p0) <= fabs(p1);
        double w0, w1, ca, cb, cc;
        if (elim_d0) {
            w0 = -p0 / p1;
            w1 = -p2 / p1;
            ca = 1.0 / (w1 * w1 - b);
            cb = 2.0 * (b * m12 - m02 * w1 + w0 * w1) * ca;
            cc = (w0 * w0 - 2.0 * m02 * w0 - b + 1.0) * ca;
        } else {
            w0 = -p1 / p0;
            w1 = -p2 / p0;
            ca = 1.0 / (-a * w1 * w1 + 2.0 * a * m12 * w1 - a + 1.0);
            cb = 2.0 * (a * m12 * w0 - m01 - a * w0 * w1) * ca;
            cc = (1.0 - a * w0 * w0) * ca;
        }
        double taus[2];
        if (p3p_quadratic(cb, cc, &taus[0], &taus[1])) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const double tau = taus[k];
                if (tau <= 0.0) continue;
                double d0, d1, d2;
                if (elim_d0) {
                    d2 = sqrt(a12 / (tau * (tau - 2.0 * m12) + 1.0));
                    d1 = tau * d2;
                    d0 = w0 * d2 + w1 * d1;
                    if (d0 < 0.0) continue;
                } else {
                    d0 = sqrt(a01 / (tau * (tau - 2.0 * m01) + 1.0));
                    d1 = tau * d0;
                    d2 = w0 * d0 + w1 * d1;
                    if (d2 < 0.0) continue;
                }
                p3p_refine(&d0, &d1, &d2, a01, a02, a12, m01, m02, m12);
                P.d[2 * i + k][0] = d0;
                P.d[2 * i + k][1] = d1;
                P.d[2 * i + k][2] = d2;
                P.valid[2 * i + k] = true;
                ++n;
            }
        }
        if (n > 0 && single_root) break;
    }
    return n;
}

}  // namespace tr
