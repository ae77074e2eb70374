// engine.hip - planner, device arenas, the per-frame path and batches.
#include "engine.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace tr {

const char* const kStageNames[kNumStages] = {"prepare", "sample_p3p",  "msac_score", "select",
                                             "rerank",  "choose",      "more_rounds", "mask",
                                             "lm"};

Status hip_status(hipError_t e, const char* what) {
    const tr_status c = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                            ? TR_ERR_OUT_OF_MEMORY
                            : TR_ERR_DEVICE;
    return Status::Err(c, std::string(what) + ": " + hipGetErrorString(e));
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

Status DevBuf::reserve(size_t need) {
    if (need <= bytes && ptr) return Status::Ok();
    release();
    const size_t sz = align_up(std::max<size_t>(need, 1), 1 << 20);
    const hipError_t e = hipMalloc(&ptr, sz);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_status(e, "hipMalloc");
    }
    bytes = sz;
    return Status::Ok();
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

Status HostBuf::reserve(size_t need) {
    if (need <= bytes && ptr) return Status::Ok();
    release();
    const size_t sz = align_up(std::max<size_t>(need, 1), 1 << 16);
    const hipError_t e = hipHostMalloc(&ptr, sz, hipHostMallocDefault);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_status(e, "hipHostMalloc");
    }
    bytes = sz;
    return Status::Ok();
}

void HostBuf::release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// ---------------------------------------------------------------------------
// Validation at the boundary (the reference asserts on a size mismatch and
// exits on any CUDA error; here every problem is a status).
// ---------------------------------------------------------------------------
// Correspondences per frame: below 2^28, so that every 32-bit index product
// the kernels form (3 i into N x 3 arrays, 4 n_pad + i into the FP32 SoA)
// stays below 2^31.
static constexpr int64_t kMaxPoints = int64_t(1) << 28;
// MSAC work items of one launch: the device indexes them with 32-bit
// integers and the host plan holds them; beyond this a launch is refused
// (2^28 points with 2^24 samples per round would ask for 2^36 of them).
static constexpr int64_t kMaxScoreItems = int64_t(1) << 26;
static constexpr uint64_t kMaxSamplesPerRound = uint64_t(1) << 24;
static constexpr uint64_t kMaxTotalSamples = uint64_t(1) << 29;

Status validate_bundle_options(const tr_bundle_options* bo) {
    if (!bo) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null bundle options");
    if (bo->loss_type < TR_LOSS_TRIVIAL || bo->loss_type > TR_LOSS_TRUNCATED_LE_ZACH)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "unsupported loss_type");
    if (!(bo->loss_scale > 0.0) || !std::isfinite(bo->loss_scale))
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "loss_scale must be positive");
    if (bo->max_iterations > 100000)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "bundle max_iterations above 100000");
    return Status::Ok();
}

static Status validate_camera(const tr_camera* cam) {
    if (!cam) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null camera");
    if (cam->model_id != 1) return Status::Err(TR_ERR_INVALID_ARGUMENT, "camera model must be PINHOLE (1)");
    for (int i = 0; i < 4; ++i)
        if (!std::isfinite(cam->params[i]))
            return Status::Err(TR_ERR_NOT_FINITE, "camera parameter is not finite");
    if (!(cam->params[0] > 0.0) || !(cam->params[1] > 0.0))
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "focal lengths must be positive");
    return Status::Ok();
}

static Status check_finite(const double* v, int64_t count, const char* what) {
    for (int64_t i = 0; i < count; ++i)
        if (!std::isfinite(v[i])) return Status::Err(TR_ERR_NOT_FINITE, std::string(what) + " not finite");
    return Status::Ok();
}

Status validate_frame(const double* uv, const double* xyz, int64_t n, const tr_camera* cam,
                      const tr_ransac_options* ro, const tr_bundle_options* bo,
                      bool host_pointers, const FrameExtras* extras) {
    if (!uv || !xyz || !cam || !ro || !bo)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 3) return Status::Err(TR_ERR_TOO_FEW_POINTS, "fewer than 3 correspondences");
    if (n >= kMaxPoints) return Status::Err(TR_ERR_CAPACITY, "2^28 or more correspondences");
    TR_TRY(validate_camera(cam));
    if (ro->max_iterations == 0 || ro->max_iterations > kMaxSamplesPerRound)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "max_iterations must be in [1, 2^24]");
    if (ro->adaptive_stopping && ro->max_total_iterations > kMaxTotalSamples)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "max_total_iterations above 2^29");
    if (!(ro->max_reproj_error > 0.0) || !std::isfinite(ro->max_reproj_error))
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "max_reproj_error must be positive");
    if (ro->adaptive_stopping &&
        (!(ro->success_prob > 0.0) || !(ro->success_prob <= 1.0) || !(ro->dyn_num_trials_mult > 0.0)))
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "success_prob in (0,1], dyn_num_trials_mult > 0");
    TR_TRY(validate_bundle_options(bo));
    if (ro->progressive_sampling && ro->max_prosac_iterations > kMaxTotalSamples)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "max_prosac_iterations above 2^29");
    if (ro->score_initial_model) {
        if (!extras) return Status::Err(TR_ERR_INVALID_ARGUMENT, "score_initial_model needs a pose");
        double qq = 0.0;
        for (int i = 0; i < 7; ++i)
            if (!std::isfinite(extras->init[i]))
                return Status::Err(TR_ERR_NOT_FINITE, "initial pose is not finite");
        for (int i = 0; i < 4; ++i) qq += extras->init[i] * extras->init[i];
        if (!(qq > 0.0)) return Status::Err(TR_ERR_INVALID_ARGUMENT, "initial quaternion is zero");
    }
    if (host_pointers && extras && extras->weights) {
        for (int64_t i = 0; i < n; ++i)
            if (!std::isfinite(extras->weights[i]))
                return Status::Err(TR_ERR_NOT_FINITE, "weights not finite");
    }
    if (host_pointers) {
        for (int64_t i = 0; i < 2 * n; ++i)
            if (!std::isfinite(uv[i])) return Status::Err(TR_ERR_NOT_FINITE, "points2D not finite");
        for (int64_t i = 0; i < 3 * n; ++i)
            if (!std::isfinite(xyz[i])) return Status::Err(TR_ERR_NOT_FINITE, "points3D not finite");
    }
    return Status::Ok();
}

// One device array: device memory of `device` or managed memory.  Plain
// host memory reads as hipMemoryTypeUnregistered (ROCm 7), pinned host
// memory as hipMemoryTypeHost; both are refused.
static Status check_device_array(int device, const void* p, const char* what) {
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return Status::Err(TR_ERR_INVALID_ARGUMENT,
                           std::string(what) + ": not an address the runtime knows");
    }
    if (a.type == hipMemoryTypeManaged || a.isManaged) return Status::Ok();
    if (a.type != hipMemoryTypeDevice)
        return Status::Err(TR_ERR_INVALID_ARGUMENT,
                           std::string(what) + ": host memory where device memory is expected");
    if (a.device != device)
        return Status::Err(TR_ERR_INVALID_ARGUMENT,
                           std::string(what) + ": memory of another device");
    return Status::Ok();
}

Status validate_device_arrays(int device, const double* uv, const double* xyz,
                              const double* weights, const uint8_t* mask) {
    TR_TRY(check_device_array(device, uv, "points2D"));
    TR_TRY(check_device_array(device, xyz, "points3D"));
    if (weights) TR_TRY(check_device_array(device, weights, "weights"));
    if (mask) TR_TRY(check_device_array(device, mask, "inliers"));
    return Status::Ok();
}

// PROSAC schedule (Chum and Matas, CVPR 2005, sections 2-3) for samples of
// m = 3: T_m = T_N prod_{i<m} (m-i)/(N-i) with T_N = max_prosac, T_{n+1} =
// T_n (n+1)/(n+1-m), T'_m = 1, T'_{n+1} = T'_n + ceil(T_{n+1} - T_n).  The
// subset grows at t = T'_n, so the events are the successive T' values,
// appended to `out` (at most n_pts - 2 of them).
void prosac_events(int64_t n_pts, uint64_t max_prosac, uint64_t t_max,
                   std::vector<uint64_t>& out) {
    const int m = 3;
    double Tn = (double)max_prosac;
    for (int i = 0; i < m; ++i) Tn *= (double)(m - i) / (double)(n_pts - i);
    uint64_t last = 1;
    out.push_back(last);
    int64_t n = m;
    while (n < n_pts && last <= t_max) {
        const double Tn1 = Tn * (double)(n + 1) / (double)(n + 1 - m);
        last += (uint64_t)std::ceil(Tn1 - Tn);
        out.push_back(last);
        Tn = Tn1;
        ++n;
    }
}

// ---------------------------------------------------------------------------
// Tuning table: MSAC kernel shape by frame size.
// ---------------------------------------------------------------------------
static const char* const kEngineNames[] = {"valu", "mfma", "gated", "gated2"};

const char* score_engine_name(int engine) {
    return (engine >= 0 && engine < 4) ? kEngineNames[engine] : "?";
}

int score_engine_id(const std::string& name) {
    for (int e = 0; e < 4; ++e)
        if (name == kEngineNames[e]) return e;
    return -1;
}

Tuning Tuning::defaults() {
    Tuning t;
    t.single = {{4096, {0, 1, 512, 4, 512}},
                {32768, {0, 1, 1024, 4, 1024}},
                {INT64_MAX, {0, 2, 1024, 4, 2048}}};
    t.batch = {{INT64_MAX, {0, 2, 1024, 4, 512}}};
    return t;
}

ScoreShape Tuning::pick(int64_t n, bool batched) const {
    const auto& v = batched ? batch : single;
    for (const auto& e : v)
        if (n <= e.max_n) return e.shape;
    return v.empty() ? ScoreShape{0, 2, 1024, 4, 2048} : v.back().shape;
}

bool Tuning::load(const std::string& path, std::string* err) {
    std::ifstream in(path);
    if (!in) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::stringstream text;
    text << in.rdbuf();
    return parse(text.str(), path, err);
}

std::string Tuning::to_text() const {
    std::ostringstream o;
    auto put = [&](const char* kind, const std::vector<Entry>& v) {
        for (const auto& e : v)
            o << kind << ' ' << (e.max_n == INT64_MAX ? -1LL : (long long)e.max_n) << ' '
              << score_engine_name(e.shape.engine) << ' ' << e.shape.cand_per_lane << ' '
              << e.shape.tile_pts << ' ' << e.shape.waves << ' ' << e.shape.chunk_pts << '\n';
    };
    put("single", single);
    put("batch", batch);
    return o.str();
}

bool Tuning::parse(const std::string& text, const std::string& source, std::string* err) {
    std::istringstream in(text);
    Tuning t;
    std::string line;
    while (std::getline(in, line)) {
        const auto h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ss(line);
        std::string kind, engine;
        long long max_n;
        ScoreShape s;
        if (!(ss >> kind)) continue;
        const bool ok = (bool)(ss >> max_n >> engine >> s.cand_per_lane >> s.tile_pts >> s.waves >>
                               s.chunk_pts);
        s.engine = score_engine_id(engine);
        if (!ok || s.engine < 0 ||
            (kind != "single" && kind != "batch") || !score_shape_supported(s) ||
            s.chunk_pts < kPad || s.chunk_pts % kPad) {
            if (err) *err = "bad tuning line: " + line;
            return false;
        }
        Entry e{max_n < 0 ? INT64_MAX : (int64_t)max_n, s};
        (kind == "single" ? t.single : t.batch).push_back(e);
    }
    auto by_n = [](const Entry& a, const Entry& b) { return a.max_n < b.max_n; };
    std::sort(t.single.begin(), t.single.end(), by_n);
    std::sort(t.batch.begin(), t.batch.end(), by_n);
    const Tuning d = defaults();
    if (t.single.empty()) t.single = d.single;
    if (t.batch.empty()) t.batch = d.batch;
    t.source = source;
    *this = t;
    return true;
}

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------
static constexpr int kRerankChunk = 1024;

static size_t cdiv(size_t a, size_t b) { return (a + b - 1) / b; }

Plan::Counts Plan::single_counts(int64_t n, uint64_t H, bool progressive, const Tuning& tune) {
    const ScoreShape shape = tune.pick(n, false);
    const size_t np = align_up((size_t)n, kPad);
    const size_t cap = 4 * (size_t)H;
    const size_t nch = cdiv(np, (size_t)shape.chunk_pts);
    Counts c;
    c.pt_blocks = np / kPtBlock;
    c.sample_blocks = cdiv((size_t)H, kSampleBlock);
    c.score_items = nch * cdiv(cap, (size_t)score_block_candidates(shape));
    c.rr_items = cdiv((size_t)n, kRerankChunk);
    c.cand_blocks = cdiv(cap, kPtBlock);
    c.lm_items = std::max<size_t>(1, cdiv((size_t)n, kLmChunk));
    c.groups = nch;
    c.prosac = progressive ? (size_t)n : 0;  // prosac_events writes at most n - 2
    return c;
}

void Plan::reserve(const Counts& c) {
    frames_.reserve(1);
    n_points.reserve(1);
    prosac_first_.reserve(1);
    for (auto* v : {&off_pts_, &off_xn_, &off_c32_, &off_s64_, &off_cslot_, &off_part_,
                    &off_rracc_, &off_rrcnt_, &mask_off_, &off_blk_, &off_lmx_, &off_lmX_, &off_lmp_,
                    &off_lmw_})
        v->reserve(1);
    pt_blocks_.reserve(c.pt_blocks);
    sample_blocks_.reserve(c.sample_blocks);
    score_items_.reserve(c.score_items);
    sorted_.reserve(c.score_items);
    rr_items_.reserve(c.rr_items);
    cand_blocks_.reserve(c.cand_blocks);
    lm_items_.reserve(c.lm_items);
    groups_.reserve(c.groups);
    order_.reserve(c.groups);
    prosac_.reserve(c.prosac);
}

bool Plan::fits(const Counts& c) const {
    return frames_.capacity() >= 1 && off_lmw_.capacity() >= 1 &&
           pt_blocks_.capacity() >= c.pt_blocks && sample_blocks_.capacity() >= c.sample_blocks &&
           score_items_.capacity() >= c.score_items && sorted_.capacity() >= c.score_items &&
           rr_items_.capacity() >= c.rr_items && cand_blocks_.capacity() >= c.cand_blocks &&
           lm_items_.capacity() >= c.lm_items && groups_.capacity() >= c.groups &&
           order_.capacity() >= c.groups && prosac_.capacity() >= c.prosac;
}

size_t frame_arena_bytes(int64_t n, uint64_t H, int chunk_pts) {
    const size_t np = align_up((size_t)n, kPad);
    const size_t cap = 4 * (size_t)H;
    const size_t nch = (np + chunk_pts - 1) / chunk_pts;
    const size_t rrch = ((size_t)n + kRerankChunk - 1) / kRerankChunk;
    const size_t lmch = std::max<size_t>(1, ((size_t)n + kLmChunk - 1) / kLmChunk);
    const size_t mb = ((size_t)n + kPtBlock - 1) / kPtBlock;
    // (weights compacted for the refinement counted whether given or not)
    return 20 * np + 16 * n + 48 * cap + 96 * cap + 4 * cap + 4 * nch * cap +
           12 * kRrSlots * rrch + 4 * mb + align_up((size_t)n, 256) + 48 * n +
           sizeof(FrameState) + sizeof(LmState) + 8 * kLmSums * lmch + 17 * 256;
}

Status Plan::build(const FrameInput* in, int nf, const Tuning& tune, bool batched, int num_cus,
                   int score_bpc) {
    if (nf <= 0) return Status::Err(TR_ERR_INVALID_ARGUMENT, "no frames");
    frames_.assign(nf, FrameDesc{});
    pt_blocks_.clear();
    sample_blocks_.clear();
    score_items_.clear();
    rr_items_.clear();
    cand_blocks_.clear();
    lm_items_.clear();
    n_points.assign(nf, 0);
    for (auto* v : {&off_pts_, &off_xn_, &off_c32_, &off_s64_, &off_cslot_, &off_part_,
                    &off_rracc_, &off_rrcnt_, &mask_off_, &off_blk_, &off_lmx_, &off_lmX_, &off_lmp_,
                    &off_lmw_})
        v->assign(nf, 0);
    prosac_.clear();
    prosac_first_.assign(nf, 0);
    int64_t max_n = 0;
    for (int f = 0; f < nf; ++f) max_n = std::max(max_n, in[f].n);
    shape_ = tune.pick(max_n, batched);
    const int CB = score_block_candidates(shape_);
    // A batch of few frames halves the tuned chunk (down to kMinChunk)
    // until there are two (candidate block, point chunk) items per resident
    // scoring workgroup, so that a launch of one or two frames still fills
    // the device.
    int chunk = shape_.chunk_pts;
    if (batched) {
        // k_msac_gated(2) take their items per wavefront, the others per workgroup.
        const int64_t per_wg = shape_.engine >= 2 ? shape_.waves : 1;
        const int64_t target = 2 * (int64_t)num_cus * std::max(1, score_bpc) * per_wg;
        auto items = [&](int c) {
            int64_t k = 0;
            for (int f = 0; f < nf; ++f) {
                const int64_t np = (int64_t)align_up((size_t)in[f].n, kPad);
                const int64_t cap = 4 * (int64_t)in[f].ro.max_iterations;
                k += ((cap + CB - 1) / CB) * ((np + c - 1) / c);
            }
            return k;
        };
        while (chunk > kMinChunk && items(chunk) < target)
            chunk = std::max(kMinChunk, (int)align_up((size_t)(chunk / 2), kPad));
        shape_.chunk_pts = chunk;
    }
    {
        int64_t items = 0;
        for (int f = 0; f < nf; ++f) {
            const int64_t np = (int64_t)align_up((size_t)in[f].n, kPad);
            const int64_t cap = 4 * (int64_t)in[f].ro.max_iterations;
            items += ((cap + CB - 1) / CB) * ((np + chunk - 1) / chunk);
        }
        if (items > kMaxScoreItems)
            return Status::Err(TR_ERR_CAPACITY,
                               "the launch's scoring work (points x samples) exceeds 2^26 items");
    }

    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = align_up(off, 256);
        const size_t o = off;
        off += bytes;
        return o;
    };
    state_off_ = take(sizeof(FrameState) * nf);
    lm_state_off_ = take(sizeof(LmState) * nf);
    size_t mask_total = 0;
    for (int f = 0; f < nf; ++f) mask_total += align_up((size_t)in[f].n, 256);
    size_t mask_region = take(mask_total);
    residual_upper_ = 0;
    samples_ = 0;
    int lm_launches = 1;
    int rounds = 1;
    for (int f = 0; f < nf; ++f) {
        const FrameInput& I = in[f];
        const int n = (int)I.n;
        const int np = (int)align_up((size_t)n, kPad);
        const int H = (int)I.ro.max_iterations;
        const int cap = 4 * H;
        const int nch = (np + chunk - 1) / chunk;
        const int rrch = (n + kRerankChunk - 1) / kRerankChunk;
        const int lmch = std::max(1, (n + kLmChunk - 1) / kLmChunk);
        const int mb = (n + kPtBlock - 1) / kPtBlock;
        n_points[f] = n;
        off_pts_[f] = take(sizeof(float) * 5 * (size_t)np);
        off_xn_[f] = take(sizeof(double) * 2 * (size_t)n);
        off_c32_[f] = take(sizeof(float) * 12 * (size_t)cap);
        off_s64_[f] = take(sizeof(double) * 12 * (size_t)cap);
        off_cslot_[f] = take(sizeof(uint32_t) * (size_t)cap);
        off_part_[f] = take(sizeof(float) * (size_t)nch * cap);
        off_rracc_[f] = take(sizeof(double) * kRrSlots * (size_t)rrch);
        off_rrcnt_[f] = take(sizeof(uint32_t) * kRrSlots * (size_t)rrch);
        off_blk_[f] = take(sizeof(uint32_t) * (size_t)mb);
        mask_off_[f] = mask_region;
        mask_region += align_up((size_t)n, 256);
        off_lmx_[f] = take(sizeof(double) * 2 * (size_t)n);
        off_lmX_[f] = take(sizeof(double) * 3 * (size_t)n);
        off_lmp_[f] = take(sizeof(double) * kLmSums * (size_t)lmch);
        off_lmw_[f] = I.w ? take(sizeof(double) * (size_t)n) : 0;

        FrameDesc& F = frames_[f];
        F.uv = I.uv;
        F.xyz = I.xyz;
        F.n = n;
        F.n_pad = np;
        F.fx = I.cam.params[0];
        F.fy = I.cam.params[1];
        F.cx = I.cam.params[2];
        F.cy = I.cam.params[3];
        const double focal = 0.5 * (F.fx + F.fy);
        F.inv_f = 1.0 / focal;
        const double tau = I.ro.max_reproj_error / focal;
        F.tau2 = tau * tau;
        F.inv_tau = 1.0 / tau;
        F.seed = I.ro.seed;
        F.adaptive = I.ro.adaptive_stopping ? 1 : 0;
        F.w = I.w;
        F.score_init = I.ro.score_initial_model ? 1 : 0;
        for (int i = 0; i < 7; ++i) F.init[i] = I.init[i];
        F.progressive = I.ro.progressive_sampling ? 1 : 0;
        F.max_prosac = I.ro.max_prosac_iterations;
        if (F.progressive) {
            // the samples this frame may draw, or up to max_prosac
            const uint64_t cap = I.ro.adaptive_stopping
                                     ? std::max<uint64_t>(I.ro.max_total_iterations, (uint64_t)H)
                                     : (uint64_t)H;
            prosac_first_[f] = prosac_.size();
            prosac_events(n, F.max_prosac, std::min<uint64_t>(cap, F.max_prosac), prosac_);
            F.prosac_len = (int32_t)(prosac_.size() - prosac_first_[f]) - 1;
        }
        F.min_samples = I.ro.min_iterations;
        F.max_samples = I.ro.max_total_iterations;
        F.success_prob = I.ro.success_prob;
        F.trials_mult = I.ro.dyn_num_trials_mult;
        F.H = H;
        F.cand_cap = cap;
        F.n_chunks = nch;
        F.chunk_pts = chunk;
        F.rr_chunks = rrch;
        F.rr_chunk_pts = kRerankChunk;
        F.lm_chunks = lmch;
        F.mask_blocks = mb;
        F.lm.max_iterations = (int32_t)I.bo.max_iterations;
        F.lm.loss_type = I.bo.loss_type;
        F.lm.loss_scale = I.bo.loss_scale * F.inv_f;
        F.lm.gradient_tol = I.bo.gradient_tol;
        F.lm.step_tol = I.bo.step_tol;
        F.lm.initial_lambda = I.bo.initial_lambda;
        F.lm.min_lambda = I.bo.min_lambda;
        F.lm.max_lambda = I.bo.max_lambda;

        for (int s = 0; s < np; s += kPtBlock) pt_blocks_.push_back({f, s});
        for (int s = 0; s < H; s += kSampleBlock) sample_blocks_.push_back({f, s});
        for (int ch = 0; ch < nch; ++ch)
            for (int cb = 0; cb < cap; cb += CB)
                score_items_.push_back({f, cb, ch * chunk, std::min(np, (ch + 1) * chunk), ch, 0});
        for (int ch = 0; ch < rrch; ++ch)
            rr_items_.push_back({f, ch, ch * kRerankChunk, std::min(n, (ch + 1) * kRerankChunk)});
        for (int s = 0; s < cap; s += kPtBlock) cand_blocks_.push_back({f, s});
        for (int c = 0; c < lmch; ++c) lm_items_.push_back({f, c});
        // Rounds the stop rule may ask for: until max_total_iterations.
        if (I.ro.adaptive_stopping) {
            const uint64_t r = std::max<uint64_t>(
                1, (I.ro.max_total_iterations + (uint64_t)H - 1) / (uint64_t)H);
            rounds = std::max<int>(rounds, (int)std::min<uint64_t>(r, 1u << 30));
        }
        residual_upper_ += (uint64_t)cap * (uint64_t)np;
        samples_ += (uint64_t)H;
        // One LM iteration per launch while the launches last; the last
        // launch finishes the frames still running.  In a launch of a few
        // frames, a frame of a few chunks is refined by one workgroup in one
        // launch instead: same arithmetic, fewer launches.
        if (n > kLmChunk && ((batched && nf > kSmallLaunchFrames) || n > kLmSoloMaxPoints))
            lm_launches = std::max<int64_t>(
                lm_launches, std::min<int64_t>(kLmMaxLaunches, (int64_t)I.bo.max_iterations + 1));
    }
    arena_bytes_ = align_up(off, 256);
    assign_score_queues();

    hdr_ = BatchHeader{};
    hdr_.n_frames = nf;
    hdr_.n_pt_blocks = (int32_t)pt_blocks_.size();
    hdr_.n_sample_blocks = (int32_t)sample_blocks_.size();
    hdr_.n_score_items = (int32_t)score_items_.size();
    hdr_.n_rr_items = (int32_t)rr_items_.size();
    hdr_.n_cand_blocks = (int32_t)cand_blocks_.size();
    hdr_.n_lm_items = (int32_t)lm_items_.size();
    for (int q = 0; q <= kScoreQueues; ++q) hdr_.q_start[q] = q_start_[q];

    // Metadata: the table of pointers (BatchDev) at offset 0, then the
    // header, descriptors and item tables.  Kernels receive only the base
    // address, so a captured graph stays valid for any later layout written
    // into the same buffer.
    meta_hdr_off_ = align_up(sizeof(BatchDev), 256);
    size_t m = meta_hdr_off_ + align_up(sizeof(BatchHeader), 256);
    meta_frames_off_ = m;
    m = align_up(m + sizeof(FrameDesc) * nf, 256);
    meta_pt_off_ = m;
    m = align_up(m + sizeof(BlockItem) * pt_blocks_.size(), 256);
    meta_smp_off_ = m;
    m = align_up(m + sizeof(BlockItem) * sample_blocks_.size(), 256);
    meta_score_off_ = m;
    m = align_up(m + sizeof(ScoreItem) * score_items_.size(), 256);
    meta_rr_off_ = m;
    m = align_up(m + sizeof(RangeItem) * rr_items_.size(), 256);
    meta_cb_off_ = m;
    m = align_up(m + sizeof(BlockItem) * cand_blocks_.size(), 256);
    meta_lm_off_ = m;
    m = align_up(m + sizeof(BlockItem) * lm_items_.size(), 256);
    meta_prosac_off_ = m;
    m = align_up(m + sizeof(uint64_t) * prosac_.size(), 256);
    meta_bytes_ = m;

    auto cap_grid = [](size_t v) { return (int)std::max<size_t>(1, std::min<size_t>(v, 65535)); };
    grid_.prepare = cap_grid(pt_blocks_.size());
    grid_.sample = cap_grid(sample_blocks_.size());
    grid_.score = std::max(1, num_cus * std::max(1, score_bpc));
    grid_.score = std::min<int>(grid_.score, std::max<int>(1, (int)score_items_.size()));
    grid_.select = cap_grid(nf);
    grid_.sumscores = cap_grid(cand_blocks_.size());
    grid_.rerank = cap_grid(rr_items_.size());
    grid_.choose = cap_grid(nf);  // one wave per frame
    grid_.mask = cap_grid(pt_blocks_.size());
    grid_.lm = cap_grid(pt_blocks_.size());
    grid_.lm_iter = cap_grid(lm_items_.size());  // one item per workgroup
    grid_.lm_launches = lm_launches;
    rounds_needed_ = rounds;
    // One frame per call: the host runs the rounds.  A batch captures up to
    // kMaxDeviceRounds; a launch whose frames may need more runs them from
    // the host (Batch::launch_eager).
    grid_.rounds = batched ? std::min(rounds, kMaxDeviceRounds) : 1;
    return Status::Ok();
}

// Orders the score items into the per-XCD queues: the items of one chunk of
// points (all its candidate blocks) form a group, and the groups go,
// largest first (ties in item order), to the queue with the least work so
// far.  Works in the plan's own containers, so a plan whose containers
// already hold the capacity allocates nothing here.
void Plan::assign_score_queues() {
    groups_.clear();
    for (size_t i = 0; i < score_items_.size();) {
        size_t j = i;
        while (j < score_items_.size() && score_items_[j].frame == score_items_[i].frame &&
               score_items_[j].chunk == score_items_[i].chunk)
            ++j;
        const double pts = (double)(score_items_[i].p1 - score_items_[i].p0);
        groups_.push_back({i, j - i, pts * (double)(j - i), 0});
        i = j;
    }
    order_.clear();
    for (size_t g = 0; g < groups_.size(); ++g) order_.push_back(g);
    std::sort(order_.begin(), order_.end(), [&](size_t a, size_t b) {
        if (groups_[a].work != groups_[b].work) return groups_[a].work > groups_[b].work;
        return a < b;
    });
    double load[kScoreQueues] = {};
    for (size_t g : order_) {
        int best = 0;
        for (int q = 1; q < kScoreQueues; ++q)
            if (load[q] < load[best]) best = q;
        groups_[g].queue = best;
        load[best] += groups_[g].work;
    }
    sorted_.clear();
    for (int q = 0; q < kScoreQueues; ++q) {
        q_start_[q] = (int32_t)sorted_.size();
        for (size_t g : order_)
            if (groups_[g].queue == q)
                for (size_t k = 0; k < groups_[g].count; ++k)
                    sorted_.push_back(score_items_[groups_[g].first + k]);
    }
    q_start_[kScoreQueues] = (int32_t)sorted_.size();
    score_items_.swap(sorted_);
}

size_t Plan::frame_desc_offset(int f) const { return meta_frames_off_ + sizeof(FrameDesc) * f; }

void Plan::bind(char* arena, char* meta_dev, char* meta_host) {
    FrameState* states = reinterpret_cast<FrameState*>(arena + state_off_);
    for (int f = 0; f < (int)frames_.size(); ++f) {
        FrameDesc& F = frames_[f];
        F.pts32 = reinterpret_cast<float*>(arena + off_pts_[f]);
        F.xn = reinterpret_cast<double*>(arena + off_xn_[f]);
        F.cand32 = reinterpret_cast<float*>(arena + off_c32_[f]);
        F.slot64 = reinterpret_cast<double*>(arena + off_s64_[f]);
        F.cand_slot = reinterpret_cast<uint32_t*>(arena + off_cslot_[f]);
        F.partial = reinterpret_cast<float*>(arena + off_part_[f]);
        F.rr_acc = reinterpret_cast<double*>(arena + off_rracc_[f]);
        F.rr_cnt = reinterpret_cast<uint32_t*>(arena + off_rrcnt_[f]);
        if (!F.mask) F.mask = reinterpret_cast<uint8_t*>(arena + mask_off_[f]);
        F.blk_count = reinterpret_cast<uint32_t*>(arena + off_blk_[f]);
        F.lm_x = reinterpret_cast<double*>(arena + off_lmx_[f]);
        F.lm_X = reinterpret_cast<double*>(arena + off_lmX_[f]);
        F.lm_state = reinterpret_cast<LmState*>(arena + lm_state_off_) + f;
        F.lm_w = F.w ? reinterpret_cast<double*>(arena + off_lmw_[f]) : nullptr;
        F.prosac = F.progressive ? reinterpret_cast<const uint64_t*>(meta_dev + meta_prosac_off_) +
                                       prosac_first_[f]
                                 : nullptr;
        F.lm_part = reinterpret_cast<double*>(arena + off_lmp_[f]);
        F.state = states + f;
    }
    std::memcpy(meta_host + meta_hdr_off_, &hdr_, sizeof(hdr_));
    std::memcpy(meta_host + meta_frames_off_, frames_.data(), sizeof(FrameDesc) * frames_.size());
    std::memcpy(meta_host + meta_pt_off_, pt_blocks_.data(), sizeof(BlockItem) * pt_blocks_.size());
    std::memcpy(meta_host + meta_smp_off_, sample_blocks_.data(),
                sizeof(BlockItem) * sample_blocks_.size());
    std::memcpy(meta_host + meta_score_off_, score_items_.data(),
                sizeof(ScoreItem) * score_items_.size());
    std::memcpy(meta_host + meta_rr_off_, rr_items_.data(), sizeof(RangeItem) * rr_items_.size());
    std::memcpy(meta_host + meta_cb_off_, cand_blocks_.data(),
                sizeof(BlockItem) * cand_blocks_.size());
    std::memcpy(meta_host + meta_lm_off_, lm_items_.data(), sizeof(BlockItem) * lm_items_.size());
    if (!prosac_.empty())
        std::memcpy(meta_host + meta_prosac_off_, prosac_.data(), sizeof(uint64_t) * prosac_.size());
    BatchDev d;
    d.hdr = reinterpret_cast<BatchHeader*>(meta_dev + meta_hdr_off_);
    d.frames = reinterpret_cast<FrameDesc*>(meta_dev + meta_frames_off_);
    d.pt_blocks = reinterpret_cast<BlockItem*>(meta_dev + meta_pt_off_);
    d.sample_blocks = reinterpret_cast<BlockItem*>(meta_dev + meta_smp_off_);
    d.score_items = reinterpret_cast<ScoreItem*>(meta_dev + meta_score_off_);
    d.rr_items = reinterpret_cast<RangeItem*>(meta_dev + meta_rr_off_);
    d.cand_blocks = reinterpret_cast<BlockItem*>(meta_dev + meta_cb_off_);
    d.lm_items = reinterpret_cast<BlockItem*>(meta_dev + meta_lm_off_);
    std::memcpy(meta_host, &d, sizeof(d));
    dev_ = reinterpret_cast<const BatchDev*>(meta_dev);
}

static Status launch_rounds_head(const Plan& p, hipStream_t st, hipEvent_t* ev) {
    const BatchDev* d = p.dev();
    const LaunchGrid& g = p.grid();
    if (ev) TR_HIP(hipEventRecord(ev[0], st));
    TR_HIP(launch_prepare(d, g.prepare, st));
    if (ev) TR_HIP(hipEventRecord(ev[1], st));
    return Status::Ok();
}

// Round r of sampling and scoring; stage events (ev[3..7]) for round 0.
static Status launch_round(const BatchDev* d, const ScoreShape& shape, const LaunchGrid& g, int r,
                           hipStream_t st, hipEvent_t* ev) {
    TR_HIP(launch_sample(d, g.sample, r, st));
    if (ev) TR_HIP(hipEventRecord(ev[2], st));
    TR_HIP(launch_score(d, shape, g.score, r, st));
    if (ev) TR_HIP(hipEventRecord(ev[3], st));
    TR_HIP(launch_select(d, g.sumscores, g.select, r, st));
    if (ev) TR_HIP(hipEventRecord(ev[4], st));
    TR_HIP(launch_rerank(d, g.rerank, r, st));
    if (ev) TR_HIP(hipEventRecord(ev[5], st));
    TR_HIP(launch_choose(d, g.choose, r, st));
    if (ev) TR_HIP(hipEventRecord(ev[6], st));
    return Status::Ok();
}

static Status launch_tail(const BatchDev* d, const LaunchGrid& g, hipStream_t st, hipEvent_t* ev) {
    TR_HIP(launch_mask(d, g.mask, st));
    if (ev) TR_HIP(hipEventRecord(ev[8], st));
    TR_HIP(launch_lm(d, g.lm, g.lm_iter, g.lm_launches, st));
    if (ev) TR_HIP(hipEventRecord(ev[9], st));
    return Status::Ok();
}

// Rounds 1.. of a launch: the frames the stop rule keeps going (the stage
// "more_rounds"; empty rounds return at once on the device).
static Status launch_more_rounds(const BatchDev* d, const ScoreShape& shape, const LaunchGrid& g,
                                 hipStream_t st, hipEvent_t* ev) {
    for (int r = 1; r < g.rounds; ++r) TR_TRY(launch_round(d, shape, g, r, st, nullptr));
    if (ev) TR_HIP(hipEventRecord(ev[7], st));
    return Status::Ok();
}

Status launch_pipeline(const Plan& p, hipStream_t st, hipEvent_t* ev) {
    TR_TRY(launch_rounds_head(p, st, ev));
    TR_TRY(launch_round(p.dev(), p.shape(), p.grid(), 0, st, ev));
    TR_TRY(launch_more_rounds(p.dev(), p.shape(), p.grid(), st, ev));
    TR_TRY(launch_tail(p.dev(), p.grid(), st, ev));
    return Status::Ok();
}

Status launch_pipeline_fixed(const BatchDev* d, const ScoreShape& shape, const LaunchGrid& g,
                             hipStream_t st) {
    TR_HIP(launch_prepare(d, g.prepare, st));
    TR_TRY(launch_round(d, shape, g, 0, st, nullptr));
    TR_TRY(launch_more_rounds(d, shape, g, st, nullptr));
    TR_TRY(launch_tail(d, g, st, nullptr));
    return Status::Ok();
}

// ---------------------------------------------------------------------------
// Scratch memory of the allocation-free entry point: one array per GPU, in
// the library's device code object, so it exists from the first use of the
// library on a device until the process ends.  Host code never passes the
// array's name to a kernel or a copy (on the host it is a stand-in, not a
// device address): it takes each device's address once with
// hipGetSymbolAddress, with that device current, and uses that.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on `device` (the current device).
static Status scratch_address(int device, char** out) {
    if (device < 0 || device >= kMaxDevices)
        return Status::Err(TR_ERR_NO_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        TR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return Status::Err(TR_ERR_DEVICE, "scratch array has no device address");
        g_scratch_ptr[device] = static_cast<char*>(p);
    }
    *out = g_scratch_ptr[device];
    return Status::Ok();
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
static std::string library_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&library_dir), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const auto s = p.rfind('/');
        if (s != std::string::npos) return p.substr(0, s);
    }
    return ".";
}

Status Context::create(int device, Context** out) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return Status::Err(TR_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count)
        return Status::Err(TR_ERR_NO_DEVICE, "device index out of range");
    TR_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    TR_HIP(hipGetDeviceProperties(&prop, device));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0)
        return Status::Err(TR_ERR_NO_DEVICE, "library is built for gfx950, device is " + arch);
    Context* c = new Context();
    c->device = device;
    c->arch = arch;
    c->num_cus = prop.multiProcessorCount;
    const hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_status(e, "hipStreamCreateWithFlags");
    }
    c->tuning = Tuning::defaults();
    const char* env = std::getenv("TR_TUNING_FILE");
    std::string path = env ? std::string(env) : library_dir() + "/../tuning/gfx950.txt";
    std::string err;
    if (!c->tuning.load(path, &err) && env) {
        delete c;
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "TR_TUNING_FILE: " + err);
    }
    Status s = scratch_address(device, &c->scratch_);
    if (s.ok()) s = c->setup_noalloc();
    if (!s.ok()) {
        delete c;
        return s;
    }
    *out = c;
    return Status::Ok();
}

Context::~Context() {
    (void)hipSetDevice(device);
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

int Context::score_occ(const ScoreShape& s) {
    for (const auto& e : occ_cache_)
        if (e.first.engine == s.engine && e.first.cand_per_lane == s.cand_per_lane &&
            e.first.tile_pts == s.tile_pts && e.first.waves == s.waves)
            return e.second;
    const int v = std::max(1, score_blocks_per_cu(s));
    occ_cache_.push_back({s, v});
    return v;
}

int Context::score_occ_peek(const ScoreShape& s) const {
    for (const auto& e : occ_cache_)
        if (e.first.engine == s.engine && e.first.cand_per_lane == s.cand_per_lane &&
            e.first.tile_pts == s.tile_pts && e.first.waves == s.waves)
            return e.second;
    return std::max(1, score_blocks_per_cu(s));
}

// For every row of the one-frame table, the largest frame the limits allow
// (the row's N bound, at most kNoAllocMaxPoints, with kNoAllocMaxSamples
// samples and PROSAC): the plan's containers, the pinned staging of its
// metadata and the row's occupancy are set up here, so that
// estimate_noalloc allocates nothing.
Status Context::setup_noalloc() {
    size_t meta = 0;
    for (const auto& e : tuning.single) {
        const int64_t n = std::min<int64_t>(e.max_n, kNoAllocMaxPoints);
        const Plan::Counts c = Plan::single_counts(n, kNoAllocMaxSamples, true, tuning);
        na_plan_.reserve(c);
        (void)score_occ(e.shape);
        size_t m = align_up(sizeof(BatchDev), 256) + align_up(sizeof(BatchHeader), 256) +
                   align_up(sizeof(FrameDesc), 256);
        m += align_up(sizeof(BlockItem) * c.pt_blocks, 256) +
             align_up(sizeof(BlockItem) * c.sample_blocks, 256) +
             align_up(sizeof(ScoreItem) * c.score_items, 256) +
             align_up(sizeof(RangeItem) * c.rr_items, 256) +
             align_up(sizeof(BlockItem) * c.cand_blocks, 256) +
             align_up(sizeof(BlockItem) * c.lm_items, 256) +
             align_up(sizeof(uint64_t) * c.prosac, 256);
        meta = std::max(meta, m);
    }
    TR_TRY(na_host_.reserve(align_up(meta, 256) + sizeof(FrameState) + 256));
    na_meta_cap_ = meta;
    return Status::Ok();
}

Status Context::estimate_noalloc(const double* uv, const double* xyz, int64_t n,
                                 const tr_camera* cam, const tr_ransac_options* ro,
                                 const tr_bundle_options* bo, const FrameExtras& extras,
                                 tr_pose_result* res, uint8_t* mask) {
    if (!res) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null result");
    TR_TRY(validate_frame(uv, xyz, n, cam, ro, bo, false, &extras));
    TR_TRY(validate_device_arrays(device, uv, xyz, extras.weights, mask));
    if (n > kNoAllocMaxPoints)
        return Status::Err(TR_ERR_CAPACITY,
                           "the allocation-free entry point takes at most 200000 correspondences");
    if (ro->max_iterations > kNoAllocMaxSamples)
        return Status::Err(TR_ERR_CAPACITY,
                           "the allocation-free entry point draws at most 16384 samples per round");
    if (!scratch_) return Status::Err(TR_ERR_INTERNAL, "no scratch array on this device");
    const Plan::Counts c =
        Plan::single_counts(n, ro->max_iterations, ro->progressive_sampling != 0, tuning);
    if (!na_plan_.fits(c))
        return Status::Err(TR_ERR_CAPACITY, "the frame's plan is larger than the "
                                            "allocation-free path reserved for the tuning table");
    TR_HIP(hipSetDevice(device));
    FrameInput in;
    in.n = n;
    in.cam = *cam;
    in.ro = *ro;
    in.bo = *bo;
    for (int i = 0; i < 7; ++i) in.init[i] = extras.init[i];
    in.uv = uv;
    in.xyz = xyz;
    const ScoreShape shape = tuning.pick(n, false);
    TR_TRY(na_plan_.build(&in, 1, tuning, false, num_cus, score_occ_peek(shape)));
    const size_t meta_off = align_up(na_plan_.arena_bytes(), 256);
    if (meta_off + na_plan_.meta_bytes() > kNoAllocScratchBytes)
        return Status::Err(TR_ERR_CAPACITY, "the frame needs more than the 64 MiB scratch array "
                                            "of the allocation-free entry point");
    if (na_plan_.meta_bytes() > na_meta_cap_)
        return Status::Err(TR_ERR_INTERNAL, "metadata larger than its reserved staging");
    if (mask) na_plan_.frame(0).mask = mask;
    char* arena = scratch_;
    char* meta = scratch_ + meta_off;
    char* host = static_cast<char*>(na_host_.ptr);
    na_plan_.bind(arena, meta, host);
    TR_HIP(hipMemcpyAsync(meta, host, na_plan_.meta_bytes(), hipMemcpyHostToDevice, stream));
    TR_TRY(launch_rounds_head(na_plan_, stream, nullptr));
    FrameState* st_host = reinterpret_cast<FrameState*>(host + align_up(na_meta_cap_, 256));
    const FrameState* st_dev = reinterpret_cast<const FrameState*>(arena + na_plan_.state_offset());
    for (int r = 0;; ++r) {
        TR_TRY(launch_round(na_plan_.dev(), na_plan_.shape(), na_plan_.grid(), r, stream, nullptr));
        if (!ro->adaptive_stopping) break;
        TR_HIP(hipMemcpyAsync(st_host, st_dev, sizeof(FrameState), hipMemcpyDeviceToHost, stream));
        TR_HIP(hipStreamSynchronize(stream));
        if (!st_host->active) break;
    }
    TR_TRY(launch_tail(na_plan_.dev(), na_plan_.grid(), stream, nullptr));
    TR_HIP(hipMemcpyAsync(st_host, st_dev, sizeof(FrameState), hipMemcpyDeviceToHost, stream));
    TR_HIP(hipStreamSynchronize(stream));
    *res = st_host->result;
    return Status::Ok();
}

Status Context::sample_triples(uint64_t seed, int64_t n, int64_t first, int64_t count,
                               int32_t* out) {
    if (!out || n < 3 || n >= kMaxPoints || first < 0 || count < 0)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "bad sampler arguments");
    if (count == 0) return Status::Ok();
    TR_HIP(hipSetDevice(device));
    TR_TRY(dbg_.reserve(sizeof(int32_t) * 3 * (size_t)count));
    TR_HIP(launch_sample_only(seed, (uint32_t)n, (uint64_t)first, count,
                              static_cast<int32_t*>(dbg_.ptr), stream));
    TR_HIP(hipMemcpyAsync(out, dbg_.ptr, sizeof(int32_t) * 3 * (size_t)count,
                          hipMemcpyDeviceToHost, stream));
    TR_HIP(hipStreamSynchronize(stream));
    return Status::Ok();
}

// Host inputs staged through pinned memory by the one-frame call, at most.
constexpr size_t kPinnedStageBytes = size_t(64) << 20;

Status Context::estimate(const double* uv, const double* xyz, int64_t n, const tr_camera* cam,
                         const tr_ransac_options* ro, const tr_bundle_options* bo,
                         const FrameExtras& extras, tr_pose_result* res, uint8_t* mask,
                         bool device_io, tr_debug_frame* dbg) {
    if (!res) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null result");
    // Host inputs of up to kPinnedStageBytes are checked for finite values
    // while they are copied into pinned staging (one pass over the bytes) and
    // go up in one asynchronous copy; larger ones are checked, then copied
    // from the caller's memory.
    const size_t in_bytes = sizeof(double) * (extras.weights ? 6 : 5) * (size_t)std::max<int64_t>(n, 0);
    const bool stage_pinned = !device_io && in_bytes <= kPinnedStageBytes;
    TR_TRY(validate_frame(uv, xyz, n, cam, ro, bo, !device_io && !stage_pinned, &extras));
    if (device_io) TR_TRY(validate_device_arrays(device, uv, xyz, extras.weights, mask));
    TR_HIP(hipSetDevice(device));
    FrameInput in;
    in.n = n;
    in.cam = *cam;
    in.ro = *ro;
    in.bo = *bo;
    for (int i = 0; i < 7; ++i) in.init[i] = extras.init[i];
    if (device_io) {
        in.uv = uv;
        in.xyz = xyz;
        in.w = extras.weights;
        in.mask_out = mask;
    } else {
        TR_TRY(in_.reserve(sizeof(double) * 6 * (size_t)n + 256));
        double* d = static_cast<double*>(in_.ptr);
        if (stage_pinned) {
            // The previous call's upload from in_host_ has finished (it
            // synchronised), or is queued before this call's on the stream
            // and its data is overwritten before any kernel reads it.
            TR_TRY(in_host_.reserve(in_bytes + 256));
            double* h = static_cast<double*>(in_host_.ptr);
            if (!copy_finite(h, uv, 2 * (size_t)n))
                return Status::Err(TR_ERR_NOT_FINITE, "points2D not finite");
            if (!copy_finite(h + 2 * n, xyz, 3 * (size_t)n))
                return Status::Err(TR_ERR_NOT_FINITE, "points3D not finite");
            if (extras.weights && !copy_finite(h + 5 * n, extras.weights, (size_t)n))
                return Status::Err(TR_ERR_NOT_FINITE, "weights not finite");
            TR_HIP(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, stream));
        } else {
            TR_HIP(hipMemcpyAsync(d, uv, sizeof(double) * 2 * n, hipMemcpyHostToDevice, stream));
            TR_HIP(hipMemcpyAsync(d + 2 * n, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                                  stream));
            if (extras.weights)
                TR_HIP(hipMemcpyAsync(d + 5 * n, extras.weights, sizeof(double) * n,
                                      hipMemcpyHostToDevice, stream));
        }
        in.uv = d;
        in.xyz = d + 2 * n;
        if (extras.weights) in.w = d + 5 * n;
    }
    Plan plan;
    const ScoreShape shape = tuning.pick(n, false);
    TR_TRY(plan.build(&in, 1, tuning, false, num_cus, score_occ(shape)));
    if (device_io && mask) plan.frame(0).mask = mask;

    // Debug dumps.
    int64_t S = 0;
    if (dbg && dbg->max_samples > 0) {
        S = dbg->max_samples;
        const size_t need = (size_t)S * (12 + 4 + 384 + 16) + 1024;
        TR_TRY(dbg_.reserve(need));
        char* base = static_cast<char*>(dbg_.ptr);
        FrameDesc& F = plan.frame(0);
        F.dbg_poses = reinterpret_cast<double*>(base);
        F.dbg_samples = reinterpret_cast<int32_t*>(base + (size_t)S * 384);
        F.dbg_nsol = reinterpret_cast<int32_t*>(base + (size_t)S * 396);
        F.dbg_scores32 = reinterpret_cast<float*>(base + (size_t)S * 400);
        F.dbg_max_samples = S;
        TR_HIP(hipMemsetAsync(base, 0xff, (size_t)S * 416, stream));
    }

    TR_TRY(arena_.reserve(plan.arena_bytes()));
    TR_TRY(meta_.reserve(plan.meta_bytes()));
    TR_TRY(meta_host_.reserve(plan.meta_bytes()));
    TR_TRY(io_host_.reserve(sizeof(FrameState) + (size_t)n + 4096));
    // The previous call's uploads read meta_host_; they finished when that
    // call synchronised.
    plan.bind(static_cast<char*>(arena_.ptr), static_cast<char*>(meta_.ptr),
              static_cast<char*>(meta_host_.ptr));
    TR_HIP(hipMemcpyAsync(meta_.ptr, meta_host_.ptr, plan.meta_bytes(), hipMemcpyHostToDevice,
                          stream));
    TR_TRY(launch_rounds_head(plan, stream, nullptr));
    FrameState* st_host = static_cast<FrameState*>(io_host_.ptr);
    const FrameState* st_dev =
        reinterpret_cast<const FrameState*>(static_cast<char*>(arena_.ptr) + plan.state_offset());
    // Rounds until the device's stop rule ends the frame (k_choose clears
    // `active`): the same decision a batch takes on the device.
    for (int r = 0;; ++r) {
        TR_TRY(launch_round(plan.dev(), plan.shape(), plan.grid(), r, stream, nullptr));
        if (!ro->adaptive_stopping) break;
        TR_HIP(hipMemcpyAsync(st_host, st_dev, sizeof(FrameState), hipMemcpyDeviceToHost, stream));
        TR_HIP(hipStreamSynchronize(stream));
        if (!st_host->active) break;
    }
    TR_TRY(launch_tail(plan.dev(), plan.grid(), stream, nullptr));
    TR_HIP(hipMemcpyAsync(st_host, st_dev, sizeof(FrameState), hipMemcpyDeviceToHost, stream));
    uint8_t* mask_host = static_cast<uint8_t*>(io_host_.ptr) + align_up(sizeof(FrameState), 64) + 64;
    if (mask && !device_io)
        TR_HIP(hipMemcpyAsync(mask_host, static_cast<char*>(arena_.ptr) + plan.mask_offset(0),
                              (size_t)n, hipMemcpyDeviceToHost, stream));
    if (S > 0) {
        char* base = static_cast<char*>(dbg_.ptr);
        if (dbg->poses)
            TR_HIP(hipMemcpyAsync(dbg->poses, base, (size_t)S * 384, hipMemcpyDeviceToHost, stream));
        if (dbg->samples)
            TR_HIP(hipMemcpyAsync(dbg->samples, base + (size_t)S * 384, (size_t)S * 12,
                                  hipMemcpyDeviceToHost, stream));
        if (dbg->num_solutions)
            TR_HIP(hipMemcpyAsync(dbg->num_solutions, base + (size_t)S * 396, (size_t)S * 4,
                                  hipMemcpyDeviceToHost, stream));
        if (dbg->scores32)
            TR_HIP(hipMemcpyAsync(dbg->scores32, base + (size_t)S * 400, (size_t)S * 16,
                                  hipMemcpyDeviceToHost, stream));
    }
    TR_HIP(hipStreamSynchronize(stream));
    *res = st_host->result;
    if (mask && !device_io) std::memcpy(mask, mask_host, (size_t)n);
    if (dbg) {
        dbg->topk_count = (int32_t)st_host->topk_count;
        for (uint32_t k = 0; k < st_host->topk_count && k < (uint32_t)kTopK; ++k) {
            if (dbg->topk_ids) dbg->topk_ids[k] = (int64_t)st_host->topk_id[k];
            if (dbg->topk_scores64) dbg->topk_scores64[k] = st_host->topk_score64[k];
        }
    }
    return Status::Ok();
}

// ---------------------------------------------------------------------------
// Refinement of a given pose on given correspondences (bundle_adjust_cuda)
// and the inlier test of a given pose (get_inliers_mx).
// ---------------------------------------------------------------------------
Status Context::bundle_adjust(const double* x, const double* X, int64_t n, const tr_camera* cam,
                              const tr_bundle_options* bo, const double* w, double* q, double* t,
                              tr_bundle_stats* stats) {
    if (!q || !t) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null pose");
    if (n < 0 || (n > 0 && (!x || !X)))
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "bad correspondences");
    if (n >= kMaxPoints) return Status::Err(TR_ERR_CAPACITY, "2^28 or more correspondences");
    TR_TRY(validate_camera(cam));
    TR_TRY(validate_bundle_options(bo));
    TR_TRY(check_finite(q, 4, "q"));
    TR_TRY(check_finite(t, 3, "t"));
    TR_TRY(check_finite(x, 2 * n, "x"));
    TR_TRY(check_finite(X, 3 * n, "X"));
    if (w) TR_TRY(check_finite(w, n, "weights"));
    TR_HIP(hipSetDevice(device));

    const int chunks = n ? (int)((n + kLmChunk - 1) / kLmChunk) : 1;
    // Metadata: the pointer table, the header, one descriptor, the items.
    const size_t m_hdr = align_up(sizeof(BatchDev), 256);
    const size_t m_fr = m_hdr + align_up(sizeof(BatchHeader), 256);
    const size_t m_it = m_fr + align_up(sizeof(FrameDesc), 256);
    const size_t m_bytes = m_it + align_up(sizeof(BlockItem) * chunks, 256);
    // Arena: the frame's state, the LM state, the chunk sums, the inputs.
    const size_t a_lms = align_up(sizeof(FrameState), 256);
    const size_t a_part = a_lms + align_up(sizeof(LmState), 256);
    const size_t a_x = a_part + align_up(sizeof(double) * kLmSums * chunks, 256);
    const size_t a_X = a_x + align_up(sizeof(double) * 2 * (size_t)n, 256);
    const size_t a_w = a_X + align_up(sizeof(double) * 3 * (size_t)n, 256);
    const size_t a_bytes = a_w + (w ? sizeof(double) * (size_t)n : 0) + 256;
    TR_TRY(arena_.reserve(a_bytes));
    TR_TRY(meta_.reserve(m_bytes));
    TR_TRY(meta_host_.reserve(m_bytes));
    TR_TRY(io_host_.reserve(sizeof(FrameState) + 256));
    char* A = static_cast<char*>(arena_.ptr);
    char* M = static_cast<char*>(meta_.ptr);
    char* H = static_cast<char*>(meta_host_.ptr);

    FrameDesc F{};
    F.n = (int32_t)n;
    F.fx = cam->params[0];
    F.fy = cam->params[1];
    F.cx = cam->params[2];
    F.cy = cam->params[3];
    F.inv_f = 1.0;  // x is in the camera's own units
    for (int i = 0; i < 4; ++i) F.init[i] = q[i];
    for (int i = 0; i < 3; ++i) F.init[4 + i] = t[i];
    F.lm_chunks = chunks;
    F.lm.max_iterations = (int32_t)bo->max_iterations;
    F.lm.loss_type = bo->loss_type;
    F.lm.loss_scale = bo->loss_scale;
    F.lm.gradient_tol = bo->gradient_tol;
    F.lm.step_tol = bo->step_tol;
    F.lm.initial_lambda = bo->initial_lambda;
    F.lm.min_lambda = bo->min_lambda;
    F.lm.max_lambda = bo->max_lambda;
    F.state = reinterpret_cast<FrameState*>(A);
    F.lm_state = reinterpret_cast<LmState*>(A + a_lms);
    F.lm_part = reinterpret_cast<double*>(A + a_part);
    F.lm_x = reinterpret_cast<double*>(A + a_x);
    F.lm_X = reinterpret_cast<double*>(A + a_X);
    F.lm_w = w ? reinterpret_cast<double*>(A + a_w) : nullptr;
    BatchHeader hdr{};
    hdr.n_frames = 1;
    hdr.n_lm_items = chunks;
    BatchDev d{};
    d.hdr = reinterpret_cast<BatchHeader*>(M + m_hdr);
    d.frames = reinterpret_cast<FrameDesc*>(M + m_fr);
    d.lm_items = reinterpret_cast<BlockItem*>(M + m_it);
    std::memcpy(H, &d, sizeof(d));
    std::memcpy(H + m_hdr, &hdr, sizeof(hdr));
    std::memcpy(H + m_fr, &F, sizeof(F));
    TR_HIP(hipMemcpyAsync(M, H, m_bytes, hipMemcpyHostToDevice, stream));
    if (n) {
        TR_HIP(hipMemcpyAsync(F.lm_x, x, sizeof(double) * 2 * n, hipMemcpyHostToDevice, stream));
        TR_HIP(hipMemcpyAsync(F.lm_X, X, sizeof(double) * 3 * n, hipMemcpyHostToDevice, stream));
        if (w) TR_HIP(hipMemcpyAsync(F.lm_w, w, sizeof(double) * n, hipMemcpyHostToDevice, stream));
    }
    // The launch forms of a one-frame call (Plan::build): one launch up to
    // kLmSoloMaxPoints correspondences, one iteration per launch above.
    int launches = 1;
    if (n > kLmChunk && n > kLmSoloMaxPoints)
        launches = (int)std::min<int64_t>(kLmMaxLaunches, (int64_t)bo->max_iterations + 1);
    TR_HIP(launch_lm_given(reinterpret_cast<const BatchDev*>(M), chunks, launches, stream));
    FrameState* st = static_cast<FrameState*>(io_host_.ptr);
    TR_HIP(hipMemcpyAsync(st, A, sizeof(FrameState), hipMemcpyDeviceToHost, stream));
    TR_HIP(hipStreamSynchronize(stream));
    const tr_pose_result& r = st->result;
    for (int i = 0; i < 4; ++i) q[i] = r.q[i];
    for (int i = 0; i < 3; ++i) t[i] = r.t[i];
    if (stats) {
        stats->iterations = r.refinements;
        stats->initial_cost = r.lm_initial_cost;
        stats->cost = r.lm_cost;
        stats->lambda = r.lm_lambda;
        stats->invalid_steps = r.lm_invalid_steps;
        stats->step_norm = r.lm_step_norm;
        stats->grad_norm = r.lm_grad_norm;
    }
    return Status::Ok();
}

Status Context::get_inliers(const double* q, const double* t, const double* x, const double* X,
                            int64_t n, double sq_threshold, uint8_t* mask, int64_t* count) {
    if (!q || !t) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null pose");
    if (n < 0 || (n > 0 && (!x || !X || !mask)))
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "bad correspondences");
    if (n >= kMaxPoints) return Status::Err(TR_ERR_CAPACITY, "2^28 or more correspondences");
    if (std::isnan(sq_threshold)) return Status::Err(TR_ERR_NOT_FINITE, "threshold is NaN");
    TR_TRY(check_finite(q, 4, "q"));
    TR_TRY(check_finite(t, 3, "t"));
    if (count) *count = 0;
    if (n == 0) return Status::Ok();
    TR_HIP(hipSetDevice(device));
    const size_t o_X = align_up(sizeof(double) * 2 * (size_t)n, 256);
    const size_t o_m = o_X + align_up(sizeof(double) * 3 * (size_t)n, 256);
    const size_t o_c = o_m + align_up((size_t)n, 256);
    TR_TRY(in_.reserve(o_c + 256));
    TR_TRY(io_host_.reserve(64));
    char* D = static_cast<char*>(in_.ptr);
    TR_HIP(hipMemcpyAsync(D, x, sizeof(double) * 2 * n, hipMemcpyHostToDevice, stream));
    TR_HIP(hipMemcpyAsync(D + o_X, X, sizeof(double) * 3 * n, hipMemcpyHostToDevice, stream));
    TR_HIP(hipMemsetAsync(D + o_c, 0, sizeof(unsigned long long), stream));
    PoseArg pose;
    for (int i = 0; i < 4; ++i) pose.q[i] = q[i];
    for (int i = 0; i < 3; ++i) pose.t[i] = t[i];
    unsigned long long* dcount = reinterpret_cast<unsigned long long*>(D + o_c);
    TR_HIP(launch_inliers(pose, reinterpret_cast<const double*>(D),
                          reinterpret_cast<const double*>(D + o_X), n, sq_threshold,
                          reinterpret_cast<uint8_t*>(D + o_m), dcount, stream));
    TR_HIP(hipMemcpyAsync(mask, D + o_m, (size_t)n, hipMemcpyDeviceToHost, stream));
    unsigned long long* hcount = static_cast<unsigned long long*>(io_host_.ptr);
    TR_HIP(hipMemcpyAsync(hcount, dcount, sizeof(*hcount), hipMemcpyDeviceToHost, stream));
    TR_HIP(hipStreamSynchronize(stream));
    if (count) *count = (int64_t)*hcount;
    return Status::Ok();
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
Status Batch::create(Context* ctx, const tr_frame* frames, int nf, Batch** out) {
    if (!ctx || !frames || !out || nf <= 0)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "bad batch arguments");
    size_t total = 0;  // doubles staged: points, and weights where given
    for (int f = 0; f < nf; ++f) {
        const tr_frame& fr = frames[f];
        FrameExtras x;
        x.weights = fr.weights;
        for (int i = 0; i < 4; ++i) x.init[i] = fr.initial_q[i];
        for (int i = 0; i < 3; ++i) x.init[4 + i] = fr.initial_t[i];
        // Finite values: checked below, while the frame is staged.
        Status s = validate_frame(fr.points2D, fr.points3D, fr.num_points, &fr.camera,
                                  &fr.ransac, &fr.bundle, false, &x);
        if (!s.ok()) return Status::Err(s.code, "frame " + std::to_string(f) + ": " + s.msg);
        total += (size_t)fr.num_points * (fr.weights ? 6 : 5);
    }
    TR_HIP(hipSetDevice(ctx->device));
    Batch* b = new Batch();
    b->ctx = ctx;
    auto fail = [&](Status s) {
        delete b;
        return s;
    };
    // A stream of its own: batches launched one after another overlap on
    // the device (one batch's latency-bound refinement under the next
    // one's scoring).
    const hipError_t se = hipStreamCreateWithFlags(&b->stream_, hipStreamNonBlocking);
    if (se != hipSuccess) return fail(hip_status(se, "hipStreamCreateWithFlags"));
    Status s = b->in_.reserve(sizeof(double) * total + 256);
    if (!s.ok()) return fail(s);
    // out_host_ stages the inputs here and receives the frames' states in
    // results(): a batch of many small frames needs more for the latter.
    s = b->out_host_.reserve(
        std::max(sizeof(double) * total + 256, sizeof(FrameState) * (size_t)nf));
    if (!s.ok()) return fail(s);
    std::vector<FrameInput> in(nf);
    double* host = static_cast<double*>(b->out_host_.ptr);
    double* dev = static_cast<double*>(b->in_.ptr);
    size_t o = 0;
    for (int f = 0; f < nf; ++f) {
        const tr_frame& fr = frames[f];
        const size_t n = (size_t)fr.num_points;
        const char* bad = !copy_finite(host + o, fr.points2D, 2 * n)          ? "points2D"
                          : !copy_finite(host + o + 2 * n, fr.points3D, 3 * n) ? "points3D"
                          : (fr.weights && !copy_finite(host + o + 5 * n, fr.weights, n)) ? "weights"
                                                                                            : nullptr;
        if (bad)
            return fail(Status::Err(TR_ERR_NOT_FINITE,
                                    "frame " + std::to_string(f) + ": " + bad + " not finite"));
        in[f].uv = dev + o;
        in[f].xyz = dev + o + 2 * n;
        in[f].n = fr.num_points;
        in[f].cam = fr.camera;
        in[f].ro = fr.ransac;
        in[f].bo = fr.bundle;
        for (int i = 0; i < 4; ++i) in[f].init[i] = fr.initial_q[i];
        for (int i = 0; i < 3; ++i) in[f].init[4 + i] = fr.initial_t[i];
        b->n_.push_back(fr.num_points);
        o += 5 * n;
        if (fr.weights) {
            in[f].w = dev + o;  // staged above
            o += n;
        }
    }
    hipError_t e = hipMemcpyAsync(dev, host, sizeof(double) * o, hipMemcpyHostToDevice, b->stream_);
    if (e != hipSuccess) return fail(hip_status(e, "upload"));
    int64_t max_n = 0;
    for (int f = 0; f < nf; ++f) max_n = std::max(max_n, in[f].n);
    const ScoreShape shape = ctx->tuning.pick(max_n, true);
    s = b->plan.build(in.data(), nf, ctx->tuning, true, ctx->num_cus, ctx->score_occ(shape));
    if (!s.ok()) return fail(s);
    if (!(s = b->arena_.reserve(b->plan.arena_bytes())).ok()) return fail(s);
    if (!(s = b->meta_.reserve(b->plan.meta_bytes())).ok()) return fail(s);
    if (!(s = b->active_host_.reserve(64)).ok()) return fail(s);
    if (!(s = b->meta_host_.reserve(b->plan.meta_bytes())).ok()) return fail(s);
    b->plan.bind(static_cast<char*>(b->arena_.ptr), static_cast<char*>(b->meta_.ptr),
                 static_cast<char*>(b->meta_host_.ptr));
    e = hipMemcpyAsync(b->meta_.ptr, b->meta_host_.ptr, b->plan.meta_bytes(),
                       hipMemcpyHostToDevice, b->stream_);
    if (e != hipSuccess) return fail(hip_status(e, "upload meta"));
    e = hipStreamSynchronize(b->stream_);
    if (e != hipSuccess) return fail(hip_status(e, "sync"));
    for (int i = 0; i <= kNumStages; ++i) {
        e = hipEventCreate(&b->ev_[i]);
        if (e != hipSuccess) return fail(hip_status(e, "hipEventCreate"));
    }
    // Capture the pipeline once; every launch replays it.
    e = hipStreamBeginCapture(b->stream_, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail(hip_status(e, "hipStreamBeginCapture"));
    s = launch_pipeline(b->plan, b->stream_, nullptr);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->stream_, &g);
    if (!s.ok()) {
        if (g) (void)hipGraphDestroy(g);
        return fail(s);
    }
    if (e != hipSuccess) return fail(hip_status(e, "hipStreamEndCapture"));
    b->graph_ = g;
    e = hipGraphInstantiate(&b->exec_, g, nullptr, nullptr, 0);
    if (e != hipSuccess) return fail(hip_status(e, "hipGraphInstantiate"));
    *out = b;
    return Status::Ok();
}

Batch::~Batch() {
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        if (stream_) (void)hipStreamSynchronize(stream_);
    }
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

// The pipeline with the rounds driven from the host: after every
// kRoundsPerCheck rounds the count of frames going on is read back, and the
// rounds stop when it is 0.  For batches whose stop rule may ask for more
// rounds than a captured graph holds.
Status Batch::launch_eager(hipEvent_t* ev) {
    hipStream_t st = stream_;
    const BatchDev* d = plan.dev();
    const LaunchGrid& g = plan.grid();
    TR_TRY(launch_rounds_head(plan, st, ev));
    uint32_t* active = static_cast<uint32_t*>(active_host_.ptr);
    const char* hdr = static_cast<const char*>(meta_.ptr) + plan.header_offset();
    for (int r = 0;; ++r) {
        TR_TRY(launch_round(d, plan.shape(), g, r, st, r == 0 ? ev : nullptr));
        if (r + 1 >= plan.rounds_needed()) break;
        if ((r + 1) % kRoundsPerCheck == 0) {
            // frames active in round r + 1
            TR_HIP(hipMemcpyAsync(active, hdr + offsetof(BatchHeader, n_active) +
                                              sizeof(uint32_t) * ((r + 1) & 1),
                                  sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            TR_HIP(hipStreamSynchronize(st));
            if (*active == 0) break;
        }
    }
    if (ev) TR_HIP(hipEventRecord(ev[7], st));
    TR_TRY(launch_tail(d, g, st, ev));
    return Status::Ok();
}

Status Batch::launch() {
    TR_HIP(hipSetDevice(ctx->device));
    profiled_ = false;
    if (plan.rounds_needed() > plan.grid().rounds) return launch_eager(nullptr);
    TR_HIP(hipGraphLaunch(exec_, stream_));
    return Status::Ok();
}

Status Batch::launch_profiled() {
    TR_HIP(hipSetDevice(ctx->device));
    if (plan.rounds_needed() > plan.grid().rounds)
        TR_TRY(launch_eager(ev_));
    else
        TR_TRY(launch_pipeline(plan, stream_, ev_));
    profiled_ = true;
    return Status::Ok();
}

Status Batch::sync() {
    TR_HIP(hipSetDevice(ctx->device));
    TR_HIP(hipStreamSynchronize(stream_));
    if (profiled_) {
        for (int i = 0; i < kNumStages; ++i) {
            float ms = 0.0f;
            TR_HIP(hipEventElapsedTime(&ms, ev_[i], ev_[i + 1]));
            stage_ms_[i] = ms;
        }
    }
    return Status::Ok();
}

int Batch::stage_times(double* ms, const char** names, int cap) const {
    const int k = std::min(cap, kNumStages);
    for (int i = 0; i < k; ++i) {
        if (ms) ms[i] = stage_ms_[i];
        if (names) names[i] = kStageNames[i];
    }
    return k;
}

Status Batch::results(tr_pose_result* res, uint8_t* const* masks) {
    TR_HIP(hipSetDevice(ctx->device));
    const int nf = plan.num_frames();
    if (sizeof(FrameState) * (size_t)nf > out_host_.bytes)
        return Status::Err(TR_ERR_INTERNAL, "result staging smaller than the frames' states");
    FrameState* host = static_cast<FrameState*>(out_host_.ptr);
    TR_HIP(hipMemcpyAsync(host, static_cast<char*>(arena_.ptr) + plan.state_offset(),
                          sizeof(FrameState) * nf, hipMemcpyDeviceToHost, stream_));
    TR_HIP(hipStreamSynchronize(stream_));
    uint64_t cands = 0, resid = 0;
    for (int f = 0; f < nf; ++f) {
        if (res) res[f] = host[f].result;
        cands += host[f].result.num_candidates;
        resid += host[f].result.num_candidates * (uint64_t)n_[f];
    }
    candidates_seen = cands;
    residuals_seen = resid;
    if (masks) {
        for (int f = 0; f < nf; ++f) {
            if (!masks[f]) continue;
            TR_HIP(hipMemcpy(masks[f], static_cast<char*>(arena_.ptr) + plan.mask_offset(f),
                             (size_t)n_[f], hipMemcpyDeviceToHost));
        }
    }
    return Status::Ok();
}

}  // namespace tr
