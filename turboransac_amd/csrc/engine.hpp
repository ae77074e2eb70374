// engine.hpp - host side of the engine: plans, device arenas, contexts.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "turboransac.h"

namespace tr {

// MSAC engine names of the tuning table: 0 "valu", 1 "mfma", 2 "gated", 3 "gated2";
// score_engine_id returns -1 for an unknown name.
const char* score_engine_name(int engine);
int score_engine_id(const std::string& name);

// A HIP error turned into a status and a message.
struct Status {
    tr_status code = TR_OK;
    std::string msg;
    bool ok() const { return code == TR_OK; }
    static Status Ok() { return {}; }
    static Status Err(tr_status c, std::string m) { return {c, std::move(m)}; }
};

Status hip_status(hipError_t e, const char* what);

#define TR_HIP(expr)                                            \
    do {                                                        \
        hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) return ::tr::hip_status(e_, #expr); \
    } while (0)

#define TR_TRY(expr)                        \
    do {                                    \
        ::tr::Status s_ = (expr);           \
        if (!s_.ok()) return s_;            \
    } while (0)

// Grow-only device allocation.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    Status reserve(size_t need);
    void release();
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// Grow-only pinned host allocation.
struct HostBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    Status reserve(size_t need);
    void release();
    ~HostBuf() { release(); }
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
};

// What a frame may carry beyond the reference's arguments.
struct FrameExtras {
    const double* weights = nullptr;  // N refinement weights (same memory as the points), or null
    double init[7] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // initial pose q, t
};

// Checks bundle options (loss type, scale, iteration cap).
Status validate_bundle_options(const tr_bundle_options* bo);

// Checks the options of one frame; the boundary's validation.
Status validate_frame(const double* uv, const double* xyz, int64_t n, const tr_camera* cam,
                      const tr_ransac_options* ro, const tr_bundle_options* bo,
                      bool host_pointers, const FrameExtras* extras = nullptr);

// Copies n doubles and reports whether all of them are finite, reading each
// value once: a host input's check and its staging copy in one pass.  A
// double is NaN or infinite exactly when its 11 exponent bits are all ones;
// the test is an integer OR over the elements, with no branch, so the loop
// vectorises.
inline bool copy_finite(double* dst, const double* src, size_t n) {
    constexpr uint64_t kExp = 0x7ff0000000000000ull;
    uint64_t bad = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t b;
        std::memcpy(&b, src + i, sizeof b);
        std::memcpy(dst + i, &b, sizeof b);
        bad |= (uint64_t)((~b & kExp) == 0);
    }
    return bad == 0;
}

// Checks that a frame's device arrays (points, and the weights and the
// mask destination when given) are device memory of `device` (or managed
// memory), so that an entry point given host memory returns an error
// instead of launching kernels that would read it.
Status validate_device_arrays(int device, const double* uv, const double* xyz,
                              const double* weights, const uint8_t* mask);

// PROSAC schedule of a frame (Chum and Matas 2005): the sample numbers
// t = 1, 2, ... at which the subset of top correspondences grows, up to
// t_max, followed by the T' after the last of them, appended to `out`.
void prosac_events(int64_t n, uint64_t max_prosac, uint64_t t_max, std::vector<uint64_t>& out);

// One frame as the planner sees it: inputs already in device memory.
struct FrameInput {
    const double* uv = nullptr;   // device, N x 2 pixels
    const double* xyz = nullptr;  // device, N x 3
    int64_t n = 0;
    tr_camera cam{};
    tr_ransac_options ro{};
    tr_bundle_options bo{};
    uint8_t* mask_out = nullptr;  // device mask destination, or null
    const double* w = nullptr;    // device, N refinement weights, or null
    double init[7] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // initial pose (score_initial_model)
};

// Tuned shape of the MSAC kernel for a frame of n points and a batch.
struct Tuning {
    struct Entry {
        int64_t max_n;  // frames with n <= max_n use this shape
        ScoreShape shape;
    };
    std::vector<Entry> single;  // one frame per call
    std::vector<Entry> batch;   // batches
    std::string source = "built-in";
    ScoreShape pick(int64_t n, bool batched) const;
    static Tuning defaults();
    bool load(const std::string& path, std::string* err);
    // Text format, one entry per line: "single|batch MAX_N ENGINE J T W CHUNK"
    // with ENGINE valu or mfma (MAX_N = -1 for no bound); '#' starts a comment.
    bool parse(const std::string& text, const std::string& source, std::string* err);
    std::string to_text() const;
};

// Limits of the allocation-free entry point, whose arena and metadata sit
// in a fixed device array of kNoAllocScratchBytes per GPU.
constexpr int64_t kNoAllocMaxPoints = 200000;
constexpr uint64_t kNoAllocMaxSamples = 16384;
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

// Work layout of a batch: item tables, arena offsets, launch grids.
class Plan {
  public:
    // Item counts of a one-frame plan (the per-frame path) for n points,
    // H samples per round and the one-frame table's shape for n.
    struct Counts {
        size_t pt_blocks, sample_blocks, score_items, rr_items, cand_blocks, lm_items, groups,
            prosac;
    };
    static Counts single_counts(int64_t n, uint64_t H, bool progressive, const Tuning& tune);
    // Grows the containers to hold a plan of these counts.
    void reserve(const Counts& c);
    // True when the containers hold such a plan without growing: build()
    // then allocates nothing.
    bool fits(const Counts& c) const;

    Status build(const FrameInput* frames, int nf, const Tuning& tune, bool batched,
                 int num_cus, int score_blocks_per_cu);
    // Points the descriptors at the arena and writes the packed metadata
    // (header, descriptors, tables) to host staging `meta_host`.
    void bind(char* arena, char* meta_dev, char* meta_host);

    int num_frames() const { return (int)frames_.size(); }
    size_t arena_bytes() const { return arena_bytes_; }
    size_t meta_bytes() const { return meta_bytes_; }
    const ScoreShape& shape() const { return shape_; }
    const LaunchGrid& grid() const { return grid_; }
    const BatchDev* dev() const { return dev_; }  // device address (metadata base)
    FrameDesc& frame(int f) { return frames_[f]; }
    size_t state_offset() const { return state_off_; }       // FrameState array
    size_t mask_offset(int f) const { return mask_off_[f]; } // internal mask of frame f
    size_t frame_desc_offset(int f) const;                   // within meta
    uint64_t residual_upper() const { return residual_upper_; }
    // Rounds the frames' stop rule may ask for (grid().rounds is what one
    // captured launch holds).
    int rounds_needed() const { return rounds_needed_; }
    size_t header_offset() const { return meta_hdr_off_; }
    uint64_t samples() const { return samples_; }
    std::vector<int64_t> n_points;

  private:
    void assign_score_queues();
    struct Group {
        size_t first, count;
        double work;
        int queue;
    };
    std::vector<Group> groups_;       // assign_score_queues' working sets, kept so that
    std::vector<size_t> order_;       // a plan rebuilt within its capacity allocates
    std::vector<ScoreItem> sorted_;   // nothing
    std::vector<FrameDesc> frames_;
    int32_t q_start_[kScoreQueues + 1] = {};
    std::vector<BlockItem> pt_blocks_, sample_blocks_;
    std::vector<ScoreItem> score_items_;
    std::vector<RangeItem> rr_items_;
    std::vector<BlockItem> cand_blocks_;
    std::vector<BlockItem> lm_items_;
    std::vector<size_t> off_pts_, off_xn_, off_c32_, off_s64_, off_cslot_, off_part_, off_rracc_,
        off_rrcnt_, mask_off_, off_blk_, off_lmx_, off_lmX_, off_lmp_, off_lmw_;
    std::vector<uint64_t> prosac_;        // all frames' PROSAC schedules, concatenated
    std::vector<size_t> prosac_first_;    // frame f's schedule starts at prosac_[prosac_first_[f]]
    size_t state_off_ = 0, lm_state_off_ = 0, arena_bytes_ = 0, meta_bytes_ = 0;
    size_t meta_hdr_off_ = 0, meta_frames_off_ = 0, meta_pt_off_ = 0, meta_smp_off_ = 0,
           meta_score_off_ = 0, meta_rr_off_ = 0, meta_cb_off_ = 0, meta_lm_off_ = 0,
           meta_prosac_off_ = 0;
    ScoreShape shape_{};
    LaunchGrid grid_{};
    const BatchDev* dev_ = nullptr;
    BatchHeader hdr_{};
    uint64_t residual_upper_ = 0, samples_ = 0;
    int rounds_needed_ = 1;
};

// Device bytes one frame occupies in a plan's arena (inputs excluded).
size_t frame_arena_bytes(int64_t n, uint64_t H, int chunk_pts);

// Stage names of a profiled launch.
constexpr int kNumStages = 9;
extern const char* const kStageNames[kNumStages];

Status launch_pipeline(const Plan& p, hipStream_t st, hipEvent_t* ev /* kNumStages+1 or null */);
// Same pipeline with fixed grids (grid-stride kernels read their item counts
// from device memory), so one captured graph serves any batch whose
// metadata sits at the same device address.
Status launch_pipeline_fixed(const BatchDev* d, const ScoreShape& shape, const LaunchGrid& g,
                             hipStream_t st);

class Context {
  public:
    static Status create(int device, Context** out);
    ~Context();

    Status estimate(const double* uv, const double* xyz, int64_t n, const tr_camera* cam,
                    const tr_ransac_options* ro, const tr_bundle_options* bo,
                    const FrameExtras& extras,
                    tr_pose_result* res, uint8_t* mask, bool device_io,
                    tr_debug_frame* dbg);
    Status sample_triples(uint64_t seed, int64_t n, int64_t first, int64_t count, int32_t* out);
    // The per-frame device path without any allocation: arena and metadata
    // in this device's scratch array, the plan in containers and pinned
    // staging reserved by setup_noalloc.
    Status estimate_noalloc(const double* uv, const double* xyz, int64_t n, const tr_camera* cam,
                            const tr_ransac_options* ro, const tr_bundle_options* bo,
                            const FrameExtras& extras, tr_pose_result* res, uint8_t* mask);
    // Reserves what estimate_noalloc needs for the current tuning table
    // (at creation and after every table change).
    Status setup_noalloc();
    // bundle_adjust_cuda (ba.cu:30-130) on host inputs: q, t in and out.
    Status bundle_adjust(const double* x, const double* X, int64_t n, const tr_camera* cam,
                         const tr_bundle_options* bo, const double* w, double* q, double* t,
                         tr_bundle_stats* stats);
    // get_inliers_mx (entry.cu:12-26) on host inputs.
    Status get_inliers(const double* q, const double* t, const double* x, const double* X,
                       int64_t n, double sq_threshold, uint8_t* mask, int64_t* count);
    // Fundamental matrix from 2D-2D correspondences (host inputs, one round
    // of ro->max_iterations 7-point samples, IRLS refinement); dbg may be null.
    Status estimate_fundamental(const double* p1, const double* p2, int64_t n,
                                const tr_ransac_options* ro, const tr_bundle_options* bo,
                                tr_fundamental_result* res, uint8_t* mask,
                                tr_fundamental_debug* dbg);
    // Many two-view frames through the same kernels in one launch sequence.
    Status estimate_fundamental_batch(const tr_pair_frame* frames, int nf,
                                      tr_fundamental_result* res, uint8_t* const* masks);

    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    std::string arch;
    std::string last_error;
    Tuning tuning;
    int score_occ(const ScoreShape& s);
    // The cached occupancy, or the query's answer without caching it.
    int score_occ_peek(const ScoreShape& s) const;

  private:
    char* scratch_ = nullptr;  // this device's copy of the scratch array
    Plan na_plan_;             // the allocation-free path's plan
    HostBuf na_host_;          // pinned: its metadata staging, then the frame state
    size_t na_meta_cap_ = 0;   // metadata bytes na_host_ holds
    DevBuf in_;     // staged host inputs
    HostBuf in_host_;  // their pinned staging (frames of up to kPinnedStageBytes of inputs)
    DevBuf arena_;
    DevBuf meta_;
    DevBuf dbg_;
    HostBuf meta_host_;
    HostBuf io_host_;
    DevBuf f_dev_;     // the fundamental-matrix path's arena
    HostBuf f_host_;   // its pinned staging: frame table and inputs up, headers and masks down
    // The fundamental-matrix launch of nf frames (one for the one-frame
    // call, with its stage dumps when dbg is given).
    Status run_fundamental(const tr_pair_frame* frames, int nf, tr_fundamental_result* res,
                           uint8_t* const* masks, tr_fundamental_debug* dbg);
    std::vector<std::pair<ScoreShape, int>> occ_cache_;
};

class Batch {
  public:
    static Status create(Context* ctx, const tr_frame* frames, int nf, Batch** out);
    ~Batch();
    // Asynchronous when the captured rounds cover every frame's stop rule;
    // otherwise the rounds run from the host, which waits on the device
    // every few rounds.
    Status launch();
    Status launch_profiled();
    Status sync();
    Status results(tr_pose_result* res, uint8_t* const* masks);
    int stage_times(double* ms, const char** names, int cap) const;
    Context* ctx = nullptr;
    Plan plan;
    uint64_t candidates_seen = 0;  // P3P poses scored, from the last results()
    uint64_t residuals_seen = 0;   // sum over frames of poses x N

  private:
    Status launch_eager(hipEvent_t* ev);
    hipStream_t stream_ = nullptr;  // every launch and copy of the batch
    DevBuf in_, arena_, meta_;
    HostBuf meta_host_, out_host_, active_host_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipEvent_t ev_[kNumStages + 1] = {};
    bool profiled_ = false;
    double stage_ms_[kNumStages] = {};
    std::vector<int64_t> n_;
};


}  // namespace tr
