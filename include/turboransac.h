/*
 * turboransac.h - C ABI of the MI355X (gfx950) absolute-pose RANSAC engine.
 *
 * One call estimates a world-to-camera pose from 2D-3D correspondences:
 * seeded minimal sampling -> P3P (FP64) -> MSAC scoring (FP32, LDS-tiled)
 * -> best hypothesis (FP64 re-rank of the FP32 top candidates) -> inlier
 * mask -> Levenberg-Marquardt refinement on the inliers (FP64).
 *
 * Conventions kept from the reference library (Module/TbrCudaPnP):
 *   - points2D are PIXELS, points3D are world coordinates (entry.h:13-19);
 *   - the MSAC threshold is given in pixels and divided by the focal
 *     length f = (fx + fy) / 2 (entry.cu:48-49, camera.cuh:78-81);
 *   - the returned pose maps world to camera, X_cam = R(q) X + t, with
 *     q = [w, x, y, z] (pose.cuh:12-28);
 *   - the inlier mask is computed from the RANSAC pose before refinement,
 *     with the strict tests r^2 < tau^2 and z > 0 (entry.cu:12-26, 79-83);
 *   - refinement runs on the inliers in focal-normalised units with a
 *     Cauchy loss (entry.cu:87-132, jaco.cuh:31-299).
 * Deliberate differences: every call returns a status code instead of
 * exiting the process (cudaCheck.h:4-14), the caller's bundle options and
 * seed are honoured (the reference ignores them, entry.cu:122 and
 * turbo_ransac.cu:159), and RansacStats are filled in.
 *
 * Memory: the caller owns every host buffer.  A context owns one HIP
 * stream and device buffers that grow on demand and are reused across
 * calls.  A context is not thread-safe: use one per host
 * thread, or the server API, whose submit call is thread-safe.
 */
#ifndef TURBORANSAC_H
#define TURBORANSAC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TR_ABI_VERSION 5

typedef enum tr_status {
    TR_OK = 0,
    TR_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad option value, unsupported loss */
    TR_ERR_TOO_FEW_POINTS = 2,   /* fewer than 3 correspondences */
    TR_ERR_NOT_FINITE = 3,       /* NaN or Inf in the correspondences or camera */
    TR_ERR_DEVICE = 4,           /* a HIP runtime call failed; see tr_context_last_error */
    TR_ERR_OUT_OF_MEMORY = 5,    /* device or host allocation failed */
    TR_ERR_CAPACITY = 6,         /* request exceeds a configured limit or HBM budget */
    TR_ERR_NOT_READY = 7,        /* server ticket not finished yet */
    TR_ERR_NO_DEVICE = 8,        /* no HIP device, or the library was built for another arch */
    TR_ERR_INTERNAL = 9
} tr_status;

/* Robust loss of the refinement; BundleOptions_cuda::LossType (utils.h:44-52).
 * The reference declares five and implements Cauchy only (jaco.cuh:37). */
typedef enum tr_loss_type {
    TR_LOSS_TRIVIAL = 0,           /* r^2 */
    TR_LOSS_TRUNCATED = 1,         /* min(r^2, s^2), IRLS weight 1 inside s, 0 outside */
    TR_LOSS_HUBER = 2,             /* r^2 inside s, s (2 r - s) outside */
    TR_LOSS_CAUCHY = 3,            /* s^2 log(1 + r^2 / s^2) up to the factor s^2 */
    TR_LOSS_TRUNCATED_LE_ZACH = 4  /* min(r^2, s^2) with the TR-IRLS weights of Le and
                                      Zach, 3DV 2021 (utils.h:50-51) */
} tr_loss_type;

/* Camera_cuda (camera.cuh:10-85), model PINHOLE only: params = fx, fy, cx, cy.
 * The reference borrows a double*; here the four values are held by value. */
typedef struct tr_camera {
    int32_t model_id; /* 1 = PINHOLE (the only model) */
    int32_t width;    /* informational, as in the reference */
    int32_t height;
    int32_t reserved;
    double params[4];
} tr_camera;

/* RansacOptions_cuda (utils.h:14-31).  Defaults: tr_ransac_options_default. */
typedef struct tr_ransac_options {
    uint64_t max_iterations;       /* minimal samples drawn per round; the reference uses it as
                                      blocks per round (turbo_ransac.cu:159-163). default 100000 */
    uint64_t min_iterations;       /* no stop before this many samples. default 1000 */
    double dyn_num_trials_mult;    /* multiplier of the stop rule. default 3.0 */
    double success_prob;           /* confidence of the stop rule. default 0.9999 */
    double max_reproj_error;       /* MSAC threshold in pixels. default 12.0 */
    uint64_t seed;                 /* key of the counter-based sampler. default 0 */
    uint64_t max_total_iterations; /* hard cap on samples; turbo_ransac.cu:145. default 100000 */
    int32_t adaptive_stopping;     /* 1: rounds until the stop rule of turbo_ransac.cu:223-257
                                      holds; 0: exactly one round. default 1 */
    int32_t progressive_sampling;  /* 1: PROSAC (Chum and Matas 2005) over correspondences
                                      sorted best first (utils.h:22-23, declared there but
                                      not implemented). default 0 */
    uint64_t max_prosac_iterations; /* samples after which PROSAC draws like uniform
                                       sampling (utils.h:24). default 100000 */
    int32_t score_initial_model;   /* 1: the frame's initial pose is scored first and competes
                                      as the earliest hypothesis (utils.h:28-29). default 0 */
    int32_t reserved;
    double max_epipolar_error;     /* Sampson-error threshold in pixels of 2D-2D estimation
                                      (tr_estimate_fundamental; utils.h:20). default 1.0 */
} tr_ransac_options;

/* BundleOptions_cuda (utils.h:33-50).  Defaults: tr_bundle_options_default. */
typedef struct tr_bundle_options {
    uint64_t max_iterations; /* default 100 */
    int32_t loss_type;       /* tr_loss_type, default TR_LOSS_CAUCHY */
    int32_t reserved;
    double loss_scale;       /* pixels, default 1.0 */
    double gradient_tol;     /* default 1e-10 */
    double step_tol;         /* default 1e-8 */
    double initial_lambda;   /* default 1e-3 */
    double min_lambda;       /* default 1e-10 */
    double max_lambda;       /* default 1e10 */
} tr_bundle_options;

/* CameraPose_cuda (pose.cuh:12-28) plus RansacStats_cuda (utils.h:25-32),
 * plus what the refinement and the hypothesis choice report. */
typedef struct tr_pose_result {
    double q[4];                /* w, x, y, z */
    double t[3];
    uint64_t iterations;        /* minimal samples drawn */
    uint64_t refinements;       /* LM iterations taken */
    uint64_t num_inliers;       /* inliers of the RANSAC pose (the mask's count) */
    double inlier_ratio;        /* num_inliers / N */
    double model_score;         /* MSAC score of the chosen hypothesis, normalised units */
    uint64_t num_candidates;    /* P3P poses scored */
    int64_t best_hypothesis;    /* (sample << 2) | solution index, -1 if no P3P pose */
    uint64_t best_inliers;      /* inliers counted while scoring the chosen hypothesis */
    double lm_initial_cost;
    double lm_cost;
    /* The rest of BundleStats_cuda (utils.h:62-72) for the refinement,
     * which the reference computes (lm.cuh:43-160) but never returns: */
    uint64_t lm_invalid_steps;  /* rejected steps */
    double lm_lambda;           /* damping at the end */
    double lm_step_norm;        /* norm of the last step solved, -1 if none */
    double lm_grad_norm;        /* norm of the last gradient computed, -1 if none */
    int32_t status;             /* tr_status of this frame */
    int32_t reserved;
} tr_pose_result;

/* BundleStats_cuda (utils.h:62-72), filled in (the reference returns it
 * default-constructed, ba.cu:104-105). */
typedef struct tr_bundle_stats {
    uint64_t iterations;     /* LM iterations taken */
    double initial_cost;
    double cost;
    double lambda;
    uint64_t invalid_steps;
    double step_norm;
    double grad_norm;
} tr_bundle_stats;

typedef struct tr_context tr_context;

void tr_ransac_options_default(tr_ransac_options* opt);
void tr_bundle_options_default(tr_bundle_options* opt);
const char* tr_status_string(tr_status s);
int32_t tr_abi_version(void);

/* Creates a context on HIP device `device`: one stream, reusable buffers. */
tr_status tr_context_create(int32_t device, tr_context** out);
/* Destroy every batch created on the context first. */
void tr_context_destroy(tr_context* ctx);
const char* tr_context_last_error(const tr_context* ctx);

/*
 * Replaces TbrCudaPnP::estimate_absolute_pose_cuda (entry.h:13-19,
 * entry.cu:28-144): one frame per call, synchronous.
 *   points2D  N x 2 doubles, pixels (host memory)
 *   points3D  N x 3 doubles, world  (host memory)
 *   inliers   N bytes out, 1 = inlier of the RANSAC pose; may be NULL
 *   result    out; with ransac_opt->score_initial_model its q and t are read
 *             first as the initial pose (the reference's in/out pose)
 */
tr_status tr_estimate_absolute_pose(tr_context* ctx, const double* points2D,
                                    const double* points3D, int64_t num_points,
                                    const tr_camera* camera,
                                    const tr_ransac_options* ransac_opt,
                                    const tr_bundle_options* bundle_opt,
                                    tr_pose_result* result, uint8_t* inliers);

/*
 * Replaces TbrCudaPnP::bundle_adjust_cuda (ba.cuh:17-24, ba.cu:30-130):
 * Levenberg-Marquardt refinement of one pose on the caller's
 * correspondences, synchronous (lm.cuh:15-171).
 *   x          N x 2 image points in the units of `camera` (pixels); the
 *              residual is K proj(R(q) X + t) - x (jaco.cuh:31-104)
 *   X          N x 3 world points
 *   opt        loss_scale in the units of x (estimate_absolute_pose_cuda
 *              passes x / f with loss_scale / f, entry.cu:87-132); every
 *              loss type is honoured (the reference always uses Cauchy,
 *              ba.cu:121-129)
 *   weights    N weights of the terms (RealWeightVector_cuda,
 *              jaco_utils.cuh:29-41), or NULL for all 1
 *   q, t       in: starting pose (q = w, x, y, z); out: refined pose
 *   stats      out, may be NULL
 * Host memory.  N = 0 is allowed and returns the pose unchanged.  The
 * reference's poses_num argument is always 1 there (entry.cu:128-134). */
tr_status tr_bundle_adjust(tr_context* ctx, const double* x, const double* X,
                           int64_t num_points, const tr_camera* camera,
                           const tr_bundle_options* opt, const double* weights, double* q,
                           double* t, tr_bundle_stats* stats);

/*
 * Replaces TbrCudaPnP::get_inliers_mx (entry.h:22-25, entry.cu:12-26):
 * inliers[i] = 1 iff |hnormalized(R(q) X_i + t) - x_i|^2 < sq_threshold and
 * the depth of R(q) X_i + t is > 0, else 0.  x: N x 2 image points in
 * normalised coordinates ((u - cx) / fx, (v - cy) / fy); X: N x 3;
 * sq_threshold in the same units.  Host memory; evaluated on the device.
 * num_inliers (may be NULL) receives the count.
 */
tr_status tr_get_inliers(tr_context* ctx, const double* q, const double* t, const double* x,
                         const double* X, int64_t num_points, double sq_threshold,
                         uint8_t* inliers, int64_t* num_inliers);

/* Same call with points2D/points3D already in device memory (HBM) and the
 * mask written to device memory (may be NULL).  No host<->device copy of
 * the correspondences.  Host memory (or another device's) in any of the
 * arrays is refused with TR_ERR_INVALID_ARGUMENT before any kernel runs. */
tr_status tr_estimate_absolute_pose_device(tr_context* ctx, const double* d_points2D,
                                           const double* d_points3D, int64_t num_points,
                                           const tr_camera* camera,
                                           const tr_ransac_options* ransac_opt,
                                           const tr_bundle_options* bundle_opt,
                                           tr_pose_result* result, uint8_t* d_inliers);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc: tr_estimate_absolute_pose_device (same arguments,
 * same results bit for bit) with no allocation in the call.  The pipeline's
 * working memory sits in a fixed device array of 64 MiB per GPU, part of
 * the library's device code, held from the first context on that GPU until
 * the process ends; the host-side plan and its pinned staging are reserved
 * by tr_context_create (and tr_context_set_tuning).  Limits: N <= 200000,
 * ransac_opt->max_iterations <= 16384, and the frame's working set within
 * the 64 MiB; beyond them the call returns TR_ERR_CAPACITY without touching
 * the device.  Every call on one GPU uses that GPU's one array: calls that
 * overlap in time on one GPU (from several contexts or threads) race and
 * must be serialised by the caller.
 */
tr_status tr_estimate_absolute_pose_device_noalloc(tr_context* ctx, const double* d_points2D,
                                                   const double* d_points3D, int64_t num_points,
                                                   const tr_camera* camera,
                                                   const tr_ransac_options* ransac_opt,
                                                   const tr_bundle_options* bundle_opt,
                                                   tr_pose_result* result, uint8_t* d_inliers);

/* ---------------------------------------------------------------------------
 * Fundamental matrix from 2D-2D correspondences: the reference paper's
 * Table I estimator, which its checkout does not contain (README.md:13);
 * the options are the reference's (RansacOptions_cuda::max_epipolar_error,
 * utils.h:20, is the threshold).  Pipeline: Hartley normalisation of each
 * image's points; max_iterations 7-point minimal samples (counter-based
 * sampler, one round) solved in FP64; MSAC of every solution on the Sampson
 * error in FP32, the FP32 leaders re-scored in FP64; the inlier mask of the
 * best model (Sampson error^2 < threshold^2); Sampson-weighted iteratively
 * reweighted 8-point refinement on its inliers (bundle_opt->max_iterations
 * iterations at most, stopping when the change of the unit-norm matrix
 * falls below bundle_opt->step_tol), rank 2 enforced at every iteration.
 * ------------------------------------------------------------------------- */
typedef struct tr_fundamental_result {
    double F[9];                /* row-major, x2^T F x1 = 0 for pixel coordinates; unit
                                   Frobenius norm, the largest entry of its last row > 0 */
    uint64_t iterations;        /* minimal samples drawn */
    uint64_t refinements;       /* refinement iterations taken */
    uint64_t num_inliers;       /* inliers of the RANSAC model (the mask's count) */
    double inlier_ratio;
    double model_score;         /* MSAC score of the chosen hypothesis, pixels^2 */
    uint64_t num_candidates;    /* 7-point solutions scored */
    int64_t best_hypothesis;    /* (sample << 2) | solution, -1 if none */
    uint64_t best_inliers;
    int32_t status;
    int32_t reserved;
} tr_fundamental_result;

/* points1, points2: N x 2 pixels (host memory), N >= 7; inliers: N bytes or
 * NULL.  Synchronous. */
tr_status tr_estimate_fundamental(tr_context* ctx, const double* points1, const double* points2,
                                  int64_t num_points, const tr_ransac_options* ransac_opt,
                                  const tr_bundle_options* bundle_opt,
                                  tr_fundamental_result* result, uint8_t* inliers);

/* Intermediate results of tr_estimate_fundamental for parity tests (every
 * pointer may be NULL): samples S x 7, num_solutions S, models S x 3 x 9 (the
 * FP64 7-point solutions in normalised coordinates, NaN where none),
 * scores32 S x 3 (FP32 MSAC scores), with S = max_samples. */
typedef struct tr_fundamental_debug {
    int64_t max_samples;
    int32_t* samples;
    int32_t* num_solutions;
    double* models;
    float* scores32;
} tr_fundamental_debug;

tr_status tr_debug_estimate_fundamental(tr_context* ctx, const double* points1,
                                        const double* points2, int64_t num_points,
                                        const tr_ransac_options* ransac_opt,
                                        const tr_bundle_options* bundle_opt,
                                        tr_fundamental_result* result, uint8_t* inliers,
                                        tr_fundamental_debug* debug);

/* Many two-view frames of uneven N in one launch sequence: every kernel of
 * the call above takes all the frames at once (a frame's work runs the same
 * partition as alone, so each result and mask is bit-identical to the
 * frame's tr_estimate_fundamental call).  frames: num_frames (1 to 65535)
 * pairs of host arrays with their options; results: num_frames; inliers:
 * num_frames pointers (NULL, or entries NULL, for no mask).  A frame the
 * one-frame call would refuse makes the whole call return that status, the
 * frame's index in tr_context_last_error.  Synchronous.  No reference
 * counterpart (the paper's estimator, one frame per call there). */
typedef struct tr_pair_frame {
    const double* points1;  /* N x 2 pixels, image 1 */
    const double* points2;  /* N x 2 pixels, image 2 */
    int64_t num_points;
    tr_ransac_options ransac;
    tr_bundle_options bundle;
} tr_pair_frame;

tr_status tr_estimate_fundamental_batch(tr_context* ctx, const tr_pair_frame* frames,
                                        int32_t num_frames, tr_fundamental_result* results,
                                        uint8_t* const* inliers);

/* ---------------------------------------------------------------------------
 * Batched engine: many frames of uneven N in one graph launch.  No reference
 * counterpart (the reference handles one frame per call); it repeats the
 * per-frame semantics of estimate_absolute_pose_cuda for every frame.
 * ------------------------------------------------------------------------- */
typedef struct tr_frame {
    const double* points2D; /* N x 2 pixels */
    const double* points3D; /* N x 3 world */
    int64_t num_points;
    tr_camera camera;
    tr_ransac_options ransac;
    tr_bundle_options bundle;
    const double* weights;  /* N weights of the refinement's terms (RealWeightVector_cuda,
                               jaco_utils.cuh:29-41; the reference passes all 1), or NULL */
    double initial_q[4];    /* initial pose (w, x, y, z) and t, scored first when */
    double initial_t[3];    /* ransac.score_initial_model is set */
} tr_frame;

/* One frame described by a tr_frame (host pointers), with its weights and
 * initial pose: the per-frame call with every option.  Same results as the
 * frame in a batch. */
tr_status tr_estimate_frame(tr_context* ctx, const tr_frame* frame, tr_pose_result* result,
                            uint8_t* inliers);

typedef struct tr_batch tr_batch;

/* Uploads the frames (host pointers), plans the work and captures the
 * pipeline into a HIP graph.  The batch owns device copies of the inputs.
 * A frame draws rounds of ransac.max_iterations samples; with
 * adaptive_stopping the stop rule runs on the device after every round,
 * exactly as in the per-frame call (without it, one round).  The graph
 * holds up to 32 rounds; when a frame's max_total_iterations allows more,
 * tr_batch_launch drives the rounds from the host instead, waiting on the
 * device every 8 rounds (the results are the same). */
tr_status tr_batch_create(tr_context* ctx, const tr_frame* frames, int32_t num_frames,
                          tr_batch** out);
/* Launches the captured graph on the batch's own stream (asynchronous).
 * Batches launched one after another overlap on the device: one batch's
 * refinement (latency-bound) runs beside the next one's scoring. */
tr_status tr_batch_launch(tr_batch* batch);
/* Runs the pipeline kernel by kernel with HIP events between the stages
 * (asynchronous); times land in tr_batch_stage_times after tr_batch_sync. */
tr_status tr_batch_launch_profiled(tr_batch* batch);
tr_status tr_batch_sync(tr_batch* batch);
/* Stage times of the last profiled launch, milliseconds; returns the number
 * of stages written (at most capacity) and names them if names != NULL.
 * The sampling stages time the first round; "more_rounds" holds the rest. */
int32_t tr_batch_stage_times(const tr_batch* batch, double* ms, const char** names,
                             int32_t capacity);
/* Work counts of the batch: residuals scored and candidates as of the last
 * tr_batch_results() call (zero before it), samples per launch. */
tr_status tr_batch_work(const tr_batch* batch, uint64_t* residuals, uint64_t* candidates,
                        uint64_t* samples);
/* The MSAC kernel shape the batch runs: engine (0 valu, 1 mfma, 2 gated),
 * J, T, W and the points per work item after the plan's chunk halving for
 * batches of few frames (5 values). */
tr_status tr_batch_score_shape(const tr_batch* batch, int32_t* shape5);
/* Copies results (and masks, each N bytes, entries may be NULL) to host. */
tr_status tr_batch_results(tr_batch* batch, tr_pose_result* results,
                           uint8_t* const* inliers);
/* Device addresses of frame `frame`'s uploaded correspondences (N x 2 and
 * N x 3 doubles), valid while the batch lives; lets a caller run the
 * per-frame device path on frames already in HBM. */
tr_status tr_batch_frame_inputs(const tr_batch* batch, int32_t frame,
                                const double** d_points2D, const double** d_points3D,
                                int64_t* num_points);
void tr_batch_destroy(tr_batch* batch);

/* create + launch + sync + results + destroy in one call. */
tr_status tr_estimate_absolute_pose_batch(tr_context* ctx, const tr_frame* frames,
                                          int32_t num_frames, tr_pose_result* results,
                                          uint8_t* const* inliers);

/* ---------------------------------------------------------------------------
 * Server: continuous batching under an HBM budget.  No reference counterpart.
 * ------------------------------------------------------------------------- */
typedef struct tr_server tr_server;

typedef struct tr_server_options {
    uint64_t hbm_budget_bytes; /* device memory the admitted frames may hold */
    int32_t max_batch_frames;  /* frames per launch */
    int32_t max_inflight;      /* launches in flight on the device (1 to 4;
                                  default 2), each slot on a stream of its
                                  own.  In a process whose only streams are
                                  the server's, 3 gave p50 0.43 ms and p99
                                  0.83 ms at 8000 frames/s against 0.50 and
                                  0.97 ms with 2; after other streams were
                                  created, 3 gave 0.58 and 1.08 ms.  HIP gives
                                  a process GPU_MAX_HW_QUEUES hardware queues
                                  (4 by default) and shares them among the
                                  streams beyond; a slot on a shared queue
                                  was measured 1.4-1.8x slower, so create the
                                  server before the process's other streams
                                  or keep those few. */
    int32_t return_masks;      /* keep inlier masks for tr_server_wait */
    int32_t reserved;
} tr_server_options;

void tr_server_options_default(tr_server_options* opt);
tr_status tr_server_create(int32_t device, const tr_server_options* opt, tr_server** out);
/* Copies the frame; returns a ticket.  Thread-safe.  TR_ERR_CAPACITY if the
 * frame alone exceeds the HBM budget. */
tr_status tr_server_submit(tr_server* srv, const tr_frame* frame, uint64_t* ticket);
/* The same for a frame whose points2D, points3D (and weights, if any) are
 * device addresses (HBM) on the server's device, e.g. written there by an
 * earlier GPU stage.  Nothing is copied: the server's launch reads the
 * caller's arrays, which must stay allocated and unchanged until
 * tr_server_wait has returned for the ticket (or the server is destroyed).
 * The footprint charged to the HBM budget leaves out the inputs.  The
 * values are not checked on the host (as in the _device per-frame call);
 * the addresses are: host memory or another device's memory is refused
 * with TR_ERR_INVALID_ARGUMENT.
 * No reference counterpart. */
tr_status tr_server_submit_device(tr_server* srv, const tr_frame* frame, uint64_t* ticket);
/* Blocks until `ticket` is done (timeout_ms < 0: forever).  Writes the
 * result, the mask (N bytes, if kept and inliers != NULL) and the latency
 * from submit to completion in microseconds. */
tr_status tr_server_wait(tr_server* srv, uint64_t ticket, int64_t timeout_ms,
                         tr_pose_result* result, uint8_t* inliers, double* latency_us);
/* Submits frames[i] at arrival_s[i] seconds after the call (open-loop
 * replay of an arrival trace), waits for all and writes latencies (us) and
 * tickets.  Returns after the last frame finished.  results (num_frames,
 * may be NULL) receives every frame's result; inliers (num_frames
 * pointers, may be NULL, entries may be NULL) its mask of N bytes when the
 * server keeps masks. */
tr_status tr_server_replay(tr_server* srv, const tr_frame* frames, const double* arrival_s,
                           int32_t num_frames, uint64_t* tickets, double* latency_us,
                           double* wall_s, tr_pose_result* results, uint8_t* const* inliers);
/* tr_server_replay for frames submitted with tr_server_submit_device. */
tr_status tr_server_replay_device(tr_server* srv, const tr_frame* frames,
                                  const double* arrival_s, int32_t num_frames, uint64_t* tickets,
                                  double* latency_us, double* wall_s, tr_pose_result* results,
                                  uint8_t* const* inliers);
/* Counters: frames done, launches, peak admitted bytes. */
tr_status tr_server_stats(const tr_server* srv, uint64_t* frames_done, uint64_t* launches,
                          uint64_t* peak_bytes);
void tr_server_destroy(tr_server* srv);
const char* tr_server_last_error(const tr_server* srv);

/* ---------------------------------------------------------------------------
 * Tuning of the MSAC kernel's shape.  No reference counterpart (the
 * reference clamps its block to 384 threads, turbo_ransac.cu:80-95).
 * A table line is "single|batch MAX_N ENGINE J T W CHUNK": frames with
 * N <= MAX_N (-1: any N) score with ENGINE "valu" (J hypotheses per lane,
 * LDS tiles of T points), "mfma" (the projection on the matrix cores, J
 * tiles of 16 hypotheses per wavefront, LDS tiles of T points) or "gated"
 * (J points per lane, T hypotheses per work item, the division only where
 * a wavefront holds an inlier), W wavefronts per workgroup and work items
 * of CHUNK points.  Contexts load tuning/gfx950.txt beside the
 * library at creation (TR_TUNING_FILE overrides the path).
 * ------------------------------------------------------------------------- */

/* Compiled (engine, J, T, W) variants, engine 0 = valu, 1 = mfma, 2 = gated; returns
 * their count, writes up to capacity quadruples to ejtw. */
int32_t tr_score_variants(int32_t* ejtw, int32_t capacity);
/* Replaces the context's table (batches created later use it). */
tr_status tr_context_set_tuning(tr_context* ctx, const char* table_text);
/* Writes the table as text (NUL-terminated, truncated to capacity);
 * returns the full length. */
int32_t tr_context_get_tuning(const tr_context* ctx, char* buf, int32_t capacity);

/* ---------------------------------------------------------------------------
 * Stage access for parity tests.  No reference counterpart.
 * ------------------------------------------------------------------------- */

/* Sampler alone: triples of distinct indices in [0, n) for global sample
 * numbers first .. first+count-1 (device kernel, copied to host). */
tr_status tr_sample_triples(tr_context* ctx, uint64_t seed, int64_t n, int64_t first,
                            int64_t count, int32_t* out_triples);

/* Per-frame path with the intermediate results copied out.  Every pointer
 * may be NULL.  Sizes with S = samples drawn (<= max_samples):
 *   samples  S x 3, num_solutions S, poses S x 4 x 12 (row-major [R|t]),
 *   scores32 S x 4 (FP32 MSAC score, NaN where no solution),
 *   topk_ids/topk_scores64 up to 16 (FP64 re-ranked candidates). */
typedef struct tr_debug_frame {
    int64_t max_samples;
    int32_t* samples;
    int32_t* num_solutions;
    double* poses;
    float* scores32;
    int64_t* topk_ids;
    double* topk_scores64;
    int32_t topk_count;
    int32_t reserved;
} tr_debug_frame;

tr_status tr_debug_estimate(tr_context* ctx, const tr_frame* frame, tr_pose_result* result,
                            uint8_t* inliers, tr_debug_frame* debug);

#ifdef __cplusplus
}
#endif

#endif /* TURBORANSAC_H */
