/*
 * oracle.h - scalar FP64 host oracle of the absolute-pose RANSAC path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU engine is
 * compared against; only the tests, the smoke check and the CPU-baseline
 * leg of the benchmark load it.  The engine never links or calls it.
 *
 * It restates, in plain scalar double-precision C++, the algorithm of the
 * reference library Module/TbrCudaPnP (file:line citations at each
 * function in oracle.cpp), with one deliberate difference: minimal samples
 * come from the counter-based Philox4x32-10 sampler defined below instead
 * of curand XORWOW (ransac_utils.cuh:14-19, ransac.cuh:176-208), so that
 * the GPU engine can reproduce the sample stream bit for bit.
 *
 * Every function is pinned by tests/ against an independent check:
 * Philox known-answer vectors and rocRAND's implementation, brute force,
 * closed forms, numpy/scipy solvers.  See the header of each test module.
 */
#ifndef TR_ORACLE_H
#define TR_ORACLE_H

#include <stdint.h>

#include "../include/turboransac.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 (Salmon et al., SC'11): out = bijection(ctr) under key. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Uniform integer in [0, bound) from a stream of 32-bit words (Lemire's
 * multiply-shift with rejection).  Exposed with an explicit word list for
 * tests: consumes words[*pos...], returns -1 if the list runs out. */
int64_t orc_bounded_from_words(const uint32_t* words, int32_t nwords, int32_t* pos,
                               uint32_t bound);

/* Minimal samples: distinct triples in [0, n) for global sample numbers
 * first .. first+count-1 of the stream keyed by `seed`. */
void orc_sample_triples(uint64_t seed, int64_t n, int64_t first, int64_t count,
                        int32_t* out);

/* Lambda-Twist P3P on three bearings (need not be unit, they are
 * normalised) and three world points, row-major x[3][3], X[3][3].
 * Writes up to 4 poses as row-major 3x4 [R|t]; returns the count. */
int32_t orc_p3p(const double* x, const double* X, double* poses);

/* MSAC score of a 3x4 pose over N correspondences in normalised image
 * coordinates (x: N x 2, X: N x 3): sum of r^2 over inliers plus
 * (N - inliers) * tau2.  Inliers: r^2 < tau2 and z > 0. */
double orc_msac_score(const double* x, const double* X, int64_t n, const double* pose,
                      double tau2, int64_t* num_inliers);

/* Quaternion utilities (w, x, y, z), row-major 3x3 rotation. */
void orc_rotmat_to_quat(const double* R, double* q);
void orc_quat_to_rotmat(const double* q, double* R);
void orc_quat_exp(const double* w, double* q);
void orc_quat_multiply(const double* a, const double* b, double* out);
void orc_quat_rotate(const double* q, const double* p, double* out);

/* Cholesky solve of the 6x6 system A x = b, reading the lower triangle of
 * the row-major A; non-positive pivots are floored to 1e-15. */
void orc_solve_cholesky6(const double* A, const double* b, double* x);

/* Levenberg-Marquardt refinement of (q, t) over correspondences x (N x 2,
 * image units of camera K = fx, fy, cx, cy) and X (N x 3).  Updates q, t in
 * place; returns iterations and writes the initial and final cost. */
typedef struct orc_lm_stats {
    uint64_t iterations;
    uint64_t invalid_steps;
    double initial_cost;
    double cost;
    double lambda;
    double grad_norm;
    double step_norm;
} orc_lm_stats;

int32_t orc_lm_refine(const double* x, const double* X, int64_t n, const double* K,
                      const tr_bundle_options* opt, double* q, double* t,
                      orc_lm_stats* stats);

/* Robust-loss cost and the normal equations (lower triangle of JtJ, Jtr)
 * at (q, t); what one LM iteration consumes. */
double orc_lm_cost(const double* x, const double* X, int64_t n, const double* K,
                   const tr_bundle_options* opt, const double* q, const double* t);
void orc_lm_normal_equations(const double* x, const double* X, int64_t n, const double* K,
                             const tr_bundle_options* opt, const double* q,
                             const double* t, double* JtJ, double* Jtr);

/* Weighted refinement: term i scaled by wts[i] (jaco_utils.cuh:29-41). */
int32_t orc_lm_refine_weighted(const double* x, const double* X, const double* wts, int64_t n,
                               const double* K, const tr_bundle_options* opt, double* q,
                               double* t, orc_lm_stats* stats);
double orc_lm_cost_weighted(const double* x, const double* X, const double* wts, int64_t n,
                            const double* K, const tr_bundle_options* opt, const double* q,
                            const double* t);
/* PROSAC triples (t = first + 1 ...) for N sorted correspondences. */
void orc_sample_prosac(uint64_t seed, int64_t n, uint64_t max_prosac, uint64_t t_max,
                       int64_t first, int64_t count, int32_t* out);

/* Intermediate results of orc_estimate_absolute_pose; every pointer may be
 * NULL.  Sizes with S = samples drawn (<= max_samples): samples S x 3,
 * num_solutions S, poses S x 4 x 12, scores S x 4 (NaN where no solution). */
typedef struct orc_trace {
    int64_t max_samples;
    int32_t* samples;
    int32_t* num_solutions;
    double* poses;
    double* scores;
    int64_t rounds;
    double best_pose[12];
} orc_trace;

/* The whole per-frame path (entry.cu:28-144 with turbo_ransac.cu:25-288 and
 * ba.cu:30-130).  Same arguments and result as tr_estimate_absolute_pose. */
int32_t orc_estimate_absolute_pose(const double* points2D, const double* points3D,
                                   int64_t n, const tr_camera* camera,
                                   const tr_ransac_options* ransac_opt,
                                   const tr_bundle_options* bundle_opt,
                                   tr_pose_result* result, uint8_t* inliers,
                                   orc_trace* trace);

/* One frame with its weights and initial pose (as tr_estimate_frame). */
int32_t orc_estimate_frame(const tr_frame* frame, tr_pose_result* result, uint8_t* inliers,
                           orc_trace* trace);

/* Many frames on `threads` host threads (frame-parallel; 1 = sequential). */
int32_t orc_estimate_many(const tr_frame* frames, int32_t num_frames, int32_t threads,
                          tr_pose_result* results);

/* The same, also writing each frame's inlier mask (masks may be NULL,
 * entries may be NULL). */
int32_t orc_estimate_many_masks(const tr_frame* frames, int32_t num_frames, int32_t threads,
                                tr_pose_result* results, uint8_t* const* masks);

/* IRLS weight of the TRUNCATED_LE_ZACH loss at x = r^2 / s^2. */
double orc_le_zach_weight(double x);

/* Samples the stop rule asks for at best-model inlier ratio w
 * (turbo_ransac.cu:223-231, with the options' success_prob and
 * dyn_num_trials_mult honoured): ceil(log(1 - p) / log(1 - w^3) * mult);
 * max_total_iterations for w <= 0, 0 for w >= 1.  A round loop stops once
 * drawn >= max_total_iterations, or drawn >= min_iterations and drawn
 * exceeds this count. */
uint64_t orc_needed_samples(double w, const tr_ransac_options* ransac_opt);

/* Pose error of main.cu:141-158: rotation angle of R_gt R_pred^T in degrees
 * and camera-centre distance in centimetres.  R row-major, t world-to-camera. */
void orc_pose_error(const double* R_pred, const double* t_pred, const double* R_gt,
                    const double* t_gt, double* rot_deg, double* center_cm);

/* ---- fundamental matrix (tr_estimate_fundamental's algorithm) ---- */

/* k-subset samples (Floyd) of the same counter-based stream, out: count x k. */
void orc_sample_subset(uint64_t seed, int64_t n, int32_t k, int64_t first, int64_t count,
                       int32_t* out);
/* Hartley normalisation of N x 2 points: T = (s, cx, cy), x_n = s (x - c). */
void orc_hartley(const double* points, int64_t n, double* T);
/* 7-point solver on 7 normalised pairs (x1n, x2n: 7 x 2): up to 3 unit-norm
 * row-major F (3 x 9); returns the count. */
int32_t orc_fundamental_7pt(const double* x1n, const double* x2n, double* F);
/* Real roots of c3 x^3 + c2 x^2 + c1 x + c0 (up to 3). */
int32_t orc_solve_cubic(double c3, double c2, double c1, double c0, double* roots);
/* MSAC score (pixels^2) of F_n on normalised pairs with s1^2, s2^2. */
double orc_sampson_msac(const double* F, const double* x1n, const double* x2n, int64_t n,
                        double s1sq, double s2sq, double tau2, int64_t* num_inliers);
/* IRLS refinement of F_n in place; returns the iterations taken. */
uint64_t orc_fundamental_irls(const double* x1n, const double* x2n, int64_t n, double s1sq,
                              double s2sq, uint64_t max_iterations, double step_tol, double* F);
/* Closest rank-2 matrix, unit norm, in place. */
void orc_fundamental_rank2(double* F);
/* F in pixels from F_n and the two normalisations (unit norm, sign rule). */
void orc_fundamental_denormalise(const double* Fn, const double* T1, const double* T2,
                                 double* F);

typedef struct orc_ftrace {
    int64_t max_samples;
    int32_t* samples;        /* S x 7 */
    int32_t* num_solutions;  /* S */
    double* models;          /* S x 3 x 9, normalised coordinates */
    double* scores;          /* S x 3 FP64 MSAC scores, NaN where no solution */
    double best_model[9];    /* the RANSAC model, normalised coordinates */
} orc_ftrace;

int32_t orc_estimate_fundamental(const double* points1, const double* points2, int64_t n,
                                 const tr_ransac_options* ransac_opt,
                                 const tr_bundle_options* bundle_opt,
                                 tr_fundamental_result* result, uint8_t* inliers,
                                 orc_ftrace* trace);

#ifdef __cplusplus
}
#endif

#endif /* TR_ORACLE_H */
