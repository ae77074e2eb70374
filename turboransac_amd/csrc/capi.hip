// capi.hip - the extern "C" boundary declared in include/turboransac.h.
#include <algorithm>
#include <cstring>
#include <new>

#include "engine.hpp"
#include "server.hpp"
#include "turboransac.h"

struct tr_context {
    tr::Context* impl;
};

struct tr_batch {
    tr::Batch* impl;
    tr_context* ctx;
};

struct tr_server {
    tr::Server* impl;
};

namespace {

tr_status fail(tr_context* ctx, const tr::Status& s) {
    if (ctx && ctx->impl) ctx->impl->last_error = s.msg;
    return s.code;
}

// Runs an entry point's body: no C++ exception crosses the C boundary.  A
// failed host allocation (std::bad_alloc) returns TR_ERR_OUT_OF_MEMORY, any
// other exception TR_ERR_INTERNAL, with the message in the context's last
// error when there is a context.
template <class Body>
tr_status guarded(tr_context* ctx, Body&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(ctx, tr::Status::Err(TR_ERR_OUT_OF_MEMORY, "host allocation failed"));
    } catch (const std::exception& e) {
        return fail(ctx, tr::Status::Err(TR_ERR_INTERNAL, e.what()));
    } catch (...) {
        return fail(ctx, tr::Status::Err(TR_ERR_INTERNAL, "unknown exception"));
    }
}

}  // namespace

extern "C" {

int32_t tr_abi_version(void) { return TR_ABI_VERSION; }

void tr_ransac_options_default(tr_ransac_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_iterations = 100000;
    o->min_iterations = 1000;
    o->dyn_num_trials_mult = 3.0;
    o->success_prob = 0.9999;
    o->max_reproj_error = 12.0;
    o->seed = 0;
    o->max_total_iterations = 100000;
    o->adaptive_stopping = 1;
    o->progressive_sampling = 0;
    o->max_prosac_iterations = 100000;  // utils.h:24
    o->score_initial_model = 0;
    o->max_epipolar_error = 1.0;  // utils.h:20
}

void tr_bundle_options_default(tr_bundle_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_iterations = 100;
    o->loss_type = TR_LOSS_CAUCHY;
    o->loss_scale = 1.0;
    o->gradient_tol = 1e-10;
    o->step_tol = 1e-8;
    o->initial_lambda = 1e-3;
    o->min_lambda = 1e-10;
    o->max_lambda = 1e10;
}

const char* tr_status_string(tr_status s) {
    switch (s) {
        case TR_OK: return "TR_OK";
        case TR_ERR_INVALID_ARGUMENT: return "TR_ERR_INVALID_ARGUMENT";
        case TR_ERR_TOO_FEW_POINTS: return "TR_ERR_TOO_FEW_POINTS";
        case TR_ERR_NOT_FINITE: return "TR_ERR_NOT_FINITE";
        case TR_ERR_DEVICE: return "TR_ERR_DEVICE";
        case TR_ERR_OUT_OF_MEMORY: return "TR_ERR_OUT_OF_MEMORY";
        case TR_ERR_CAPACITY: return "TR_ERR_CAPACITY";
        case TR_ERR_NOT_READY: return "TR_ERR_NOT_READY";
        case TR_ERR_NO_DEVICE: return "TR_ERR_NO_DEVICE";
        case TR_ERR_INTERNAL: return "TR_ERR_INTERNAL";
    }
    return "unknown tr_status";
}

tr_status tr_context_create(int32_t device, tr_context** out) {
    return guarded(nullptr, [&]() -> tr_status {
        if (!out) return TR_ERR_INVALID_ARGUMENT;
        *out = nullptr;
        tr::Context* c = nullptr;
        const tr::Status s = tr::Context::create(device, &c);
        if (!s.ok()) return s.code;
        tr_context* h = new (std::nothrow) tr_context{c};
        if (!h) {
            delete c;
            return TR_ERR_OUT_OF_MEMORY;
        }
        *out = h;
        return TR_OK;
    });
}

void tr_context_destroy(tr_context* ctx) {
    if (!ctx) return;
    delete ctx->impl;
    delete ctx;
}

const char* tr_context_last_error(const tr_context* ctx) {
    if (!ctx || !ctx->impl) return "";
    return ctx->impl->last_error.c_str();
}

// The extras of a frame beyond the reference's arguments.
static tr::FrameExtras extras_of(const tr_frame& f) {
    tr::FrameExtras x;
    x.weights = f.weights;
    for (int i = 0; i < 4; ++i) x.init[i] = f.initial_q[i];
    for (int i = 0; i < 3; ++i) x.init[4 + i] = f.initial_t[i];
    return x;
}

// The initial pose of the reference-shaped calls: result's q and t (the
// reference's in/out pose).
static tr::FrameExtras extras_of_result(const tr_pose_result* r) {
    tr::FrameExtras x;
    if (r) {
        for (int i = 0; i < 4; ++i) x.init[i] = r->q[i];
        for (int i = 0; i < 3; ++i) x.init[4 + i] = r->t[i];
    }
    return x;
}

tr_status tr_estimate_absolute_pose(tr_context* ctx, const double* points2D,
                                    const double* points3D, int64_t num_points,
                                    const tr_camera* camera, const tr_ransac_options* ransac_opt,
                                    const tr_bundle_options* bundle_opt, tr_pose_result* result,
                                    uint8_t* inliers) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s =
            ctx->impl->estimate(points2D, points3D, num_points, camera, ransac_opt, bundle_opt,
                                extras_of_result(result), result, inliers, false, nullptr);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_estimate_absolute_pose_device(tr_context* ctx, const double* d_points2D,
                                           const double* d_points3D, int64_t num_points,
                                           const tr_camera* camera,
                                           const tr_ransac_options* ransac_opt,
                                           const tr_bundle_options* bundle_opt,
                                           tr_pose_result* result, uint8_t* d_inliers) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s =
            ctx->impl->estimate(d_points2D, d_points3D, num_points, camera, ransac_opt, bundle_opt,
                                extras_of_result(result), result, d_inliers, true, nullptr);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_estimate_absolute_pose_device_noalloc(tr_context* ctx, const double* d_points2D,
                                                   const double* d_points3D, int64_t num_points,
                                                   const tr_camera* camera,
                                                   const tr_ransac_options* ransac_opt,
                                                   const tr_bundle_options* bundle_opt,
                                                   tr_pose_result* result, uint8_t* d_inliers) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s =
            ctx->impl->estimate_noalloc(d_points2D, d_points3D, num_points, camera, ransac_opt,
                                        bundle_opt, extras_of_result(result), result, d_inliers);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_bundle_adjust(tr_context* ctx, const double* x, const double* X,
                           int64_t num_points, const tr_camera* camera,
                           const tr_bundle_options* opt, const double* weights, double* q,
                           double* t, tr_bundle_stats* stats) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s =
            ctx->impl->bundle_adjust(x, X, num_points, camera, opt, weights, q, t, stats);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_get_inliers(tr_context* ctx, const double* q, const double* t, const double* x,
                         const double* X, int64_t num_points, double sq_threshold,
                         uint8_t* inliers, int64_t* num_inliers) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s =
            ctx->impl->get_inliers(q, t, x, X, num_points, sq_threshold, inliers, num_inliers);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_estimate_fundamental(tr_context* ctx, const double* points1, const double* points2,
                                  int64_t num_points, const tr_ransac_options* ransac_opt,
                                  const tr_bundle_options* bundle_opt,
                                  tr_fundamental_result* result, uint8_t* inliers) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s = ctx->impl->estimate_fundamental(points1, points2, num_points, ransac_opt,
                                                             bundle_opt, result, inliers, nullptr);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_estimate_fundamental_batch(tr_context* ctx, const tr_pair_frame* frames,
                                        int32_t num_frames, tr_fundamental_result* results,
                                        uint8_t* const* inliers) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s =
            ctx->impl->estimate_fundamental_batch(frames, num_frames, results, inliers);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_debug_estimate_fundamental(tr_context* ctx, const double* points1,
                                        const double* points2, int64_t num_points,
                                        const tr_ransac_options* ransac_opt,
                                        const tr_bundle_options* bundle_opt,
                                        tr_fundamental_result* result, uint8_t* inliers,
                                        tr_fundamental_debug* debug) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl || !debug) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s = ctx->impl->estimate_fundamental(points1, points2, num_points, ransac_opt,
                                                             bundle_opt, result, inliers, debug);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_estimate_frame(tr_context* ctx, const tr_frame* f, tr_pose_result* result,
                            uint8_t* inliers) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl || !f) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s =
            ctx->impl->estimate(f->points2D, f->points3D, f->num_points, &f->camera, &f->ransac,
                                &f->bundle, extras_of(*f), result, inliers, false, nullptr);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

tr_status tr_debug_estimate(tr_context* ctx, const tr_frame* f, tr_pose_result* result,
                            uint8_t* inliers, tr_debug_frame* debug) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl || !debug || !f) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s =
            ctx->impl->estimate(f->points2D, f->points3D, f->num_points, &f->camera, &f->ransac,
                                &f->bundle, extras_of(*f), result, inliers, false, debug);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

int32_t tr_score_variants(int32_t* jtw, int32_t capacity) {
    return tr::score_variant_list(jtw, capacity);
}

tr_status tr_context_set_tuning(tr_context* ctx, const char* table_text) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl || !table_text) return TR_ERR_INVALID_ARGUMENT;
        std::string err;
        if (!ctx->impl->tuning.parse(table_text, "tr_context_set_tuning", &err))
            return fail(ctx, tr::Status::Err(TR_ERR_INVALID_ARGUMENT, err));
        const tr::Status s = ctx->impl->setup_noalloc();
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

int32_t tr_context_get_tuning(const tr_context* ctx, char* buf, int32_t capacity) {
    if (!ctx || !ctx->impl) return 0;
    const std::string t = ctx->impl->tuning.to_text();
    if (buf && capacity > 0) {
        const size_t k = std::min<size_t>(t.size(), (size_t)capacity - 1);
        std::memcpy(buf, t.data(), k);
        buf[k] = '\0';
    }
    return (int32_t)t.size();
}

tr_status tr_sample_triples(tr_context* ctx, uint64_t seed, int64_t n, int64_t first,
                            int64_t count, int32_t* out_triples) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s = ctx->impl->sample_triples(seed, n, first, count, out_triples);
        return s.ok() ? TR_OK : fail(ctx, s);
    });
}

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------
tr_status tr_batch_create(tr_context* ctx, const tr_frame* frames, int32_t num_frames,
                          tr_batch** out) {
    return guarded(ctx, [&]() -> tr_status {
        if (!ctx || !ctx->impl || !out) return TR_ERR_INVALID_ARGUMENT;
        *out = nullptr;
        tr::Batch* b = nullptr;
        const tr::Status s = tr::Batch::create(ctx->impl, frames, num_frames, &b);
        if (!s.ok()) return fail(ctx, s);
        tr_batch* h = new (std::nothrow) tr_batch{b, ctx};
        if (!h) {
            delete b;
            return TR_ERR_OUT_OF_MEMORY;
        }
        *out = h;
        return TR_OK;
    });
}

tr_status tr_batch_launch(tr_batch* b) {
    return guarded((b ? b->ctx : nullptr), [&]() -> tr_status {
        if (!b || !b->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s = b->impl->launch();
        return s.ok() ? TR_OK : fail(b->ctx, s);
    });
}

tr_status tr_batch_launch_profiled(tr_batch* b) {
    return guarded((b ? b->ctx : nullptr), [&]() -> tr_status {
        if (!b || !b->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s = b->impl->launch_profiled();
        return s.ok() ? TR_OK : fail(b->ctx, s);
    });
}

tr_status tr_batch_sync(tr_batch* b) {
    return guarded((b ? b->ctx : nullptr), [&]() -> tr_status {
        if (!b || !b->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s = b->impl->sync();
        return s.ok() ? TR_OK : fail(b->ctx, s);
    });
}

int32_t tr_batch_stage_times(const tr_batch* b, double* ms, const char** names,
                             int32_t capacity) {
    if (!b || !b->impl || capacity <= 0) return 0;
    return b->impl->stage_times(ms, names, capacity);
}

tr_status tr_batch_work(const tr_batch* b, uint64_t* residuals, uint64_t* candidates,
                        uint64_t* samples) {
    return guarded((b ? b->ctx : nullptr), [&]() -> tr_status {
        if (!b || !b->impl) return TR_ERR_INVALID_ARGUMENT;
        // Residuals of the last results() call: candidates x N, frame by frame.
        if (residuals) *residuals = b->impl->residuals_seen;
        if (candidates) *candidates = b->impl->candidates_seen;
        if (samples) *samples = b->impl->plan.samples();
        return TR_OK;
    });
}

tr_status tr_batch_score_shape(const tr_batch* b, int32_t* shape5) {
    return guarded((b ? b->ctx : nullptr), [&]() -> tr_status {
        if (!b || !b->impl || !shape5) return TR_ERR_INVALID_ARGUMENT;
        const tr::ScoreShape& s = b->impl->plan.shape();
        shape5[0] = s.engine;
        shape5[1] = s.cand_per_lane;
        shape5[2] = s.tile_pts;
        shape5[3] = s.waves;
        shape5[4] = s.chunk_pts;
        return TR_OK;
    });
}

tr_status tr_batch_results(tr_batch* b, tr_pose_result* results, uint8_t* const* inliers) {
    return guarded((b ? b->ctx : nullptr), [&]() -> tr_status {
        if (!b || !b->impl) return TR_ERR_INVALID_ARGUMENT;
        const tr::Status s = b->impl->results(results, inliers);
        return s.ok() ? TR_OK : fail(b->ctx, s);
    });
}

tr_status tr_batch_frame_inputs(const tr_batch* b, int32_t frame, const double** d_points2D,
                                const double** d_points3D, int64_t* num_points) {
    return guarded((b ? b->ctx : nullptr), [&]() -> tr_status {
        if (!b || !b->impl || frame < 0 || frame >= b->impl->plan.num_frames())
            return TR_ERR_INVALID_ARGUMENT;
        const tr::FrameDesc& F = b->impl->plan.frame(frame);
        if (d_points2D) *d_points2D = F.uv;
        if (d_points3D) *d_points3D = F.xyz;
        if (num_points) *num_points = F.n;
        return TR_OK;
    });
}

void tr_batch_destroy(tr_batch* b) {
    if (!b) return;
    delete b->impl;
    delete b;
}

tr_status tr_estimate_absolute_pose_batch(tr_context* ctx, const tr_frame* frames,
                                          int32_t num_frames, tr_pose_result* results,
                                          uint8_t* const* inliers) {
    return guarded(ctx, [&]() -> tr_status {
        tr_batch* b = nullptr;
        tr_status s = tr_batch_create(ctx, frames, num_frames, &b);
        if (s != TR_OK) return s;
        if ((s = tr_batch_launch(b)) == TR_OK && (s = tr_batch_sync(b)) == TR_OK)
            s = tr_batch_results(b, results, inliers);
        tr_batch_destroy(b);
        return s;
    });
}

// ---------------------------------------------------------------------------
// Server
// ---------------------------------------------------------------------------
void tr_server_options_default(tr_server_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->hbm_budget_bytes = uint64_t(8) << 30;
    o->max_batch_frames = 256;
    o->max_inflight = 2;
    o->return_masks = 0;
}

tr_status tr_server_create(int32_t device, const tr_server_options* opt, tr_server** out) {
    return guarded(nullptr, [&]() -> tr_status {
        if (!out) return TR_ERR_INVALID_ARGUMENT;
        *out = nullptr;
        tr_server_options o;
        if (opt)
            o = *opt;
        else
            tr_server_options_default(&o);
        tr::Server* s = nullptr;
        const tr::Status st = tr::Server::create(device, o, &s);
        if (!st.ok()) return st.code;
        *out = new tr_server{s};
        return TR_OK;
    });
}

tr_status tr_server_submit(tr_server* srv, const tr_frame* frame, uint64_t* ticket) {
    return guarded(nullptr, [&]() -> tr_status {
        if (!srv || !srv->impl) return TR_ERR_INVALID_ARGUMENT;
        return srv->impl->submit(frame, ticket).code;
    });
}

tr_status tr_server_submit_device(tr_server* srv, const tr_frame* frame, uint64_t* ticket) {
    return guarded(nullptr, [&]() -> tr_status {
        if (!srv || !srv->impl) return TR_ERR_INVALID_ARGUMENT;
        return srv->impl->submit(frame, ticket, true).code;
    });
}

tr_status tr_server_wait(tr_server* srv, uint64_t ticket, int64_t timeout_ms,
                         tr_pose_result* result, uint8_t* inliers, double* latency_us) {
    return guarded(nullptr, [&]() -> tr_status {
        if (!srv || !srv->impl) return TR_ERR_INVALID_ARGUMENT;
        return srv->impl->wait(ticket, timeout_ms, result, inliers, latency_us).code;
    });
}

tr_status tr_server_replay(tr_server* srv, const tr_frame* frames, const double* arrival_s,
                           int32_t num_frames, uint64_t* tickets, double* latency_us,
                           double* wall_s, tr_pose_result* results, uint8_t* const* inliers) {
    return guarded(nullptr, [&]() -> tr_status {
        if (!srv || !srv->impl) return TR_ERR_INVALID_ARGUMENT;
        return srv->impl
            ->replay(frames, arrival_s, num_frames, tickets, latency_us, wall_s, results, inliers)
            .code;
    });
}

tr_status tr_server_replay_device(tr_server* srv, const tr_frame* frames,
                                  const double* arrival_s, int32_t num_frames, uint64_t* tickets,
                                  double* latency_us, double* wall_s, tr_pose_result* results,
                                  uint8_t* const* inliers) {
    return guarded(nullptr, [&]() -> tr_status {
        if (!srv || !srv->impl) return TR_ERR_INVALID_ARGUMENT;
        return srv->impl
            ->replay(frames, arrival_s, num_frames, tickets, latency_us, wall_s, results, inliers,
                     true)
            .code;
    });
}

tr_status tr_server_stats(const tr_server* srv, uint64_t* frames_done, uint64_t* launches,
                          uint64_t* peak_bytes) {
    return guarded(nullptr, [&]() -> tr_status {
        if (!srv || !srv->impl) return TR_ERR_INVALID_ARGUMENT;
        srv->impl->stats(frames_done, launches, peak_bytes);
        return TR_OK;
    });
}

void tr_server_destroy(tr_server* srv) {
    if (!srv) return;
    delete srv->impl;
    delete srv;
}

const char* tr_server_last_error(const tr_server* srv) {
    if (!srv || !srv->impl) return "";
    return srv->impl->last_error();
}

}  // extern "C"
