// kernels.hpp - device data layout of a batch and the kernel launchers.
//
// A batch is a set of frames processed by one pass of the pipeline
//   reset -> prepare -> sample+P3P -> MSAC score -> top-K select
//   -> FP64 re-rank -> choose -> inlier mask -> LM refinement.
// Every kernel is a grid-stride loop over an item table whose length it
// reads from device memory (BatchHeader), so one launch shape (and one
// captured HIP graph) serves batches of any composition up to capacity.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "turboransac.h"

namespace tr {

struct LmParams {
    int32_t max_iterations;
    int32_t loss_type;
    double loss_scale;  // normalised units (pixels / f)
    double gradient_tol, step_tol, initial_lambda, min_lambda, max_lambda;
};

// Levenberg-Marquardt state of one frame, in device memory between the
// launches of the LM kernel.  The step works on a register copy of the
// LmCore part; the statistics, which it only writes, stay in LDS.
struct LmCore {
    double cur[7];     // accepted pose (q, t)
    double prop[7];    // pose of the next pass
    double JtJ[36];    // lower triangle; the diagonal carries the damping between passes
    double Jtr[6];
    double cost, initial_cost, lambda;
    uint64_t iters;    // passes after the initial one
    int32_t nin;       // compacted inliers
    int32_t chunks;    // inlier chunks of kLmChunk (at least 1)
    uint32_t arrived;  // chunks finished in the current launch
    int32_t recompute; // last step accepted: test the gradient before the next solve
    int32_t done;
    int32_t pad;
};
struct LmState : LmCore {
    double grad_norm;  // last gradient norm computed, -1 before the first
    double step_norm;  // last step norm solved, -1 before the first
    uint64_t invalid;  // rejected steps
};

// Per-frame state written by the kernels.
struct FrameState {
    uint32_t n_cand;       // P3P poses appended this round
    uint32_t topk_count;
    uint32_t topk_idx[kTopK];   // slot of the leader
    uint32_t topk_id[kTopK];    // (global sample << 2) | solution
    float topk_score32[kTopK];
    double topk_score64[kTopK];
    uint32_t topk_inliers[kTopK];
    uint32_t n_inliers;    // mask count
    // Best hypothesis over all rounds so far.
    double best_score;
    int64_t best_id;
    uint64_t best_inliers;
    uint64_t total_candidates;
    uint64_t samples_drawn;
    double best_pose[12];
    double mask_R[9];      // rotation of the RANSAC pose through its quaternion
    double mask_q[4];
    double mask_t[3];
    int32_t status;
    int32_t active;        // 1 while the frame draws more rounds
    tr_pose_result result;
};

// Frame descriptor: inputs, sizes and the frame's slices of the arenas.
struct FrameDesc {
    const double* uv;   // N x 2 pixels
    const double* xyz;  // N x 3 world
    int32_t n;
    int32_t n_pad;      // multiple of kPad
    double fx, fy, cx, cy;
    double inv_f;       // 1 / ((fx + fy) / 2)
    double tau2;        // (threshold / f)^2
    double inv_tau;     // f / threshold: the FP32 scoring arrays measure image
                        // coordinates in units of the threshold
    uint64_t seed;
    // Stop rule (turbo_ransac.cu:223-257, options honoured): after a round,
    // go on while drawn < max_samples and not (drawn >= min_samples and
    // drawn > ceil(log(1 - success_prob) / log(1 - w^3) * trials_mult)).
    uint64_t min_samples, max_samples;
    double success_prob, trials_mult;
    int32_t adaptive;
    // PROSAC (Chum and Matas 2005) over correspondences sorted best first:
    // the events E_0 = 1 < E_1 < ... (prosac_len of them, then the T' that
    // follows), samples past max_prosac uniform.
    int32_t progressive;
    const uint64_t* prosac;
    uint64_t max_prosac;
    int32_t prosac_len;
    int32_t score_init;    // the initial pose competes as hypothesis -2
    double init[7];        // initial pose q, t
    int32_t H;             // samples per round
    int32_t cand_cap;      // 4 H
    int32_t n_chunks;      // scoring chunks
    int32_t chunk_pts;     // points per scoring chunk (multiple of kPad)
    int32_t rr_chunks;     // FP64 re-rank chunks
    int32_t rr_chunk_pts;
    int32_t lm_chunks;     // ceil(N / kLmChunk): LM items planned for the frame
    int32_t mask_blocks;   // ceil(N / kPtBlock)
    LmParams lm;

    float* pts32;          // SoA u, v, X, Y, Z; each n_pad floats
    double* xn;            // N x 2 normalised image coordinates
    // P3P output: the FP32 scoring copy by candidate index (the order of
    // cand_slot), the FP64 pose by slot = 4 * sample + solution (sample
    // within the round).
    float* cand32;         // cand_cap x 12, row-major [R | t], rows 0 and 1 / tau
    double* slot64;        // cand_cap x 12
    uint32_t* cand_slot;   // cand_cap: the filled slots, appended by the sampler
    float* partial;        // n_chunks x cand_cap, indexed like cand_slot
    double* rr_acc;        // rr_chunks x kRrSlots (the leaders, then the initial pose)
    uint32_t* rr_cnt;      // rr_chunks x kRrSlots
    uint8_t* mask;         // N
    uint32_t* blk_count;   // mask_blocks: inliers per block of kPtBlock points
    double* lm_x;          // N x 2 inliers, pixels / f
    double* lm_X;          // N x 3
    const double* w;       // N refinement weights (input), or null
    double* lm_w;          // N compacted weights, null without weights
    LmState* lm_state;
    double* lm_part;       // lm_chunks x kLmSums: per-chunk sums of the current pass
    FrameState* state;

    // Optional stage dumps (null unless a debug call asked for them).
    int32_t* dbg_samples;  // S x 3
    int32_t* dbg_nsol;     // S
    double* dbg_poses;     // S x 4 x 12
    float* dbg_scores32;   // S x 4
    int64_t dbg_max_samples;
};

struct BlockItem {  // prepare, sampler, mask: a block of a frame
    int32_t frame;
    int32_t start;
};

struct ScoreItem {  // MSAC: candidates [cbase, cbase + CB) x points [p0, p1)
    int32_t frame;
    int32_t cbase;
    int32_t p0;
    int32_t p1;
    int32_t chunk;
    int32_t pad;
};

struct RangeItem {  // FP64 re-rank: points [p0, p1) of a frame
    int32_t frame;
    int32_t chunk;
    int32_t p0;
    int32_t p1;
};

// The MSAC work items sit in kScoreQueues queues, one per XCD: the items of
// one chunk of points go to one queue, so the candidate blocks sharing the
// chunk run on the XCD whose L2 holds it.
constexpr int kScoreQueues = 8;
constexpr int kQHeadStride = 64;  // uint32 words between queue heads: one 256 B line each

struct BatchHeader {
    int32_t n_frames;
    int32_t n_pt_blocks;
    int32_t n_sample_blocks;
    int32_t n_score_items;
    int32_t n_rr_items;
    int32_t n_cand_blocks;
    int32_t n_lm_items;
    // Frames active in rounds of even / odd index: k_choose of round r
    // counts the frames going on into slot (r + 1) & 1, and the kernels of
    // a round whose count is 0 return at once.
    uint32_t n_active[2];
    // LM items of the launch: n_lm_items is the planned capacity (chunks
    // of every point), n_lm_live the chunks of the actual inliers, which
    // k_lm_compact writes at the front of lm_items.
    uint32_t n_lm_live;
    int32_t q_start[kScoreQueues + 1];  // queue q holds score items [q_start[q], q_start[q+1])
    // Items taken from each queue (reset per round).  Each head has a line
    // of its own, apart from the fields every workgroup reads: atomics on a
    // line that is also read everywhere slow the whole kernel several-fold.
    alignas(256) uint32_t q_head[kScoreQueues * kQHeadStride];
};

struct BatchDev {
    BatchHeader* hdr;
    FrameDesc* frames;
    BlockItem* pt_blocks;
    BlockItem* sample_blocks;
    ScoreItem* score_items;
    RangeItem* rr_items;
    BlockItem* cand_blocks;  // candidate slots [start, start + 256) of a frame
    BlockItem* lm_items;     // LM chunk `start` of a frame (the live ones, from k_lm_compact)
};

// Shape of the MSAC kernel (the autotuner's search space).
struct ScoreShape {
    int32_t engine;         // 0: k_msac (vector ALU), 1: k_msac_mfma (projection on MFMA),
                            // 2: k_msac_gated (points on lanes, gated division),
                            // 3: k_msac_gated2 (the same, two hypotheses per pass)
    int32_t cand_per_lane;  // k_msac: J hypotheses per lane; k_msac_mfma: M tiles of 16 per
                            // wave; k_msac_gated: J points per lane
    int32_t tile_pts;       // T: points per LDS tile; k_msac_gated: hypotheses per work item
    int32_t waves;          // wavefronts per workgroup
    int32_t chunk_pts;      // points per work item (multiple of kPad)
};

inline int score_block_candidates(const ScoreShape& s) {
    if (s.engine >= 2) return s.tile_pts;  // k_msac_gated(2): hypotheses per work item
    return s.engine ? 16 * s.cand_per_lane * s.waves : 64 * s.waves * s.cand_per_lane;
}

// Grid sizes of one launch of the pipeline.
struct LaunchGrid {
    int prepare, sample, score, sumscores, select, rerank, choose, mask, lm;
    int rounds;       // sampling rounds in one launch (adaptive stopping on the device)
    int lm_iter;      // grid of each LM iteration launch
    int lm_launches;  // LM iteration launches (the last one finishes every frame)
};

bool score_shape_supported(const ScoreShape& s);
// Compiled (engine, J, T, W) variants; writes up to cap quadruples, returns
// the count.
int score_variant_list(int32_t* ejtw, int cap);

// Stage launchers (stream-ordered, asynchronous).
// The stages of a round take the round's index: frames that stopped are
// skipped, and a round with no active frame returns at once.
hipError_t launch_prepare(const BatchDev* b, int grid, hipStream_t st);
hipError_t launch_sample(const BatchDev* b, int grid, int round, hipStream_t st);
hipError_t launch_score(const BatchDev* b, const ScoreShape& s, int grid, int round,
                        hipStream_t st);
// Folds each candidate's per-chunk partial scores into chunk 0, then picks
// the frame's leaders.
hipError_t launch_select(const BatchDev* b, int grid_sum, int grid, int round, hipStream_t st);
hipError_t launch_rerank(const BatchDev* b, int grid, int round, hipStream_t st);
// Best hypothesis so far, the stop rule, and the next round's counters.
hipError_t launch_choose(const BatchDev* b, int grid, int round, hipStream_t st);
hipError_t launch_mask(const BatchDev* b, int grid, hipStream_t st);
// Compaction, then `launches` LM iteration launches.
hipError_t launch_lm(const BatchDev* b, int grid, int grid_iter, int launches, hipStream_t st);
// Refinement of frames whose correspondences are already in lm_x / lm_X
// (lm_w) with F.n of them, from the pose F.init: the LM state and items are
// set up by k_lm_start instead of the compaction.
hipError_t launch_lm_given(const BatchDev* b, int grid_iter, int launches, hipStream_t st);
// inliers[i] = |hnormalized(R(q) X_i + t) - x_i|^2 < thr and depth > 0;
// *count (zeroed by the caller) receives the number of inliers.
struct PoseArg {
    double q[4];
    double t[3];
};
hipError_t launch_inliers(PoseArg pose, const double* x, const double* X, int64_t n, double thr,
                          uint8_t* inliers, unsigned long long* count, hipStream_t st);
hipError_t launch_sample_only(uint64_t seed, uint32_t n, uint64_t first, int64_t count,
                              int32_t* out, hipStream_t st);

// Workgroups of the MSAC kernel that fit on one CU at once.
int score_blocks_per_cu(const ScoreShape& s);

}  // namespace tr
