// server.hip - continuous batching under an HBM budget.
//
// Requests are admitted first-in first-out while the device footprint of
// the admitted frames stays under the budget and a launch slot is free.
// Each slot owns a stream, device buffers and a HIP graph of the pipeline
// captured with fixed grids; the grid-stride kernels read the batch's item
// counts from device memory, so the same graph replays every composition.
// With two slots, frames that arrive while one launch runs are uploaded and
// launched on the other stream and overlap with it on the device.
#include "server.hpp"

#include <sys/prctl.h>

#include <algorithm>
#include <cstring>

namespace tr {

static size_t align_up_s(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Device footprint of a request: its inputs' copy in the slot's input
// buffer (none for a device-resident frame, read in place), its arena and
// its share of the metadata.
static size_t request_bytes(int64_t n, uint64_t H, int chunk, size_t input_doubles_per_point) {
    return sizeof(double) * input_doubles_per_point * (size_t)n + frame_arena_bytes(n, H, chunk) +
           4096;
}

Status Server::create(int device, const tr_server_options& opt, Server** out) {
    if (opt.max_batch_frames <= 0 || opt.max_batch_frames > 4096 || opt.max_inflight <= 0 ||
        opt.max_inflight > 4 || opt.hbm_budget_bytes == 0)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "bad server options");
    Context* probe = nullptr;
    TR_TRY(Context::create(device, &probe));
    std::unique_ptr<Server> s(new Server());
    s->device_ = device;
    s->num_cus_ = probe->num_cus;
    s->opt_ = opt;
    auto add_shape = [&](int64_t max_n, const ScoreShape& sh) {
        ShapeChoice c;
        c.max_n = max_n;
        c.shape = sh;
        c.tuning = probe->tuning;
        c.tuning.batch = {{INT64_MAX, sh}};
        c.occ = std::max(1, probe->score_occ(sh));
        s->shapes_.push_back(c);
    };
    add_shape(INT64_MAX, probe->tuning.pick(INT64_MAX, true));
    for (const auto& e : probe->tuning.single) add_shape(e.max_n, e.shape);
    delete probe;

    const int mf = opt.max_batch_frames;
    LaunchGrid& g = s->grid_;
    g.prepare = s->num_cus_ * 8;
    g.sample = s->num_cus_ * 8;
    g.score = s->num_cus_;  // per launch: num_cus x the shape's occupancy
    g.select = mf;
    g.sumscores = s->num_cus_ * 8;
    g.rerank = s->num_cus_ * 4;
    g.choose = mf;
    g.mask = s->num_cus_ * 8;
    g.lm = s->num_cus_ * 8;
    g.lm_iter = s->num_cus_ * kLmServerBlocksPerCu;
    g.lm_launches = 1;  // per launch, from the plan
    g.rounds = 1;       // per launch, from the plan

    TR_HIP(hipSetDevice(device));
    for (int i = 0; i < opt.max_inflight; ++i) {
        auto slot = std::make_unique<Slot>();
        TR_HIP(hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking));
        TR_HIP(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming));
        s->slots_.push_back(std::move(slot));
    }
    for (auto& slot : s->slots_)
        slot->waiter = std::thread([p = s.get(), sl = slot.get()] { p->wait_slot(sl); });
    s->worker_ = std::thread([p = s.get()] { p->run(); });
    *out = s.release();
    return Status::Ok();
}

int Server::pick_shape(int nf, int64_t max_n) const {
    if (nf > kSmallLaunchFrames || shapes_.size() < 2) return 0;
    for (size_t i = 1; i < shapes_.size(); ++i)
        if (max_n <= shapes_[i].max_n) return (int)i;
    return (int)shapes_.size() - 1;
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
    {
        std::lock_guard<std::mutex> lk(mu_);
        waiters_stop_ = true;
    }
    cv_slot_.notify_all();
    for (auto& s : slots_)
        if (s->waiter.joinable()) s->waiter.join();
    (void)hipSetDevice(device_);
    for (auto& s : slots_) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        for (auto& kv : s->exec)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
        if (s->done) (void)hipEventDestroy(s->done);
        if (s->stream) (void)hipStreamDestroy(s->stream);
    }
}

void Server::set_error(const std::string& m) {
    std::lock_guard<std::mutex> lk(mu_);
    error_ = m;
}

const char* Server::last_error() const {
    std::lock_guard<std::mutex> lk(mu_);
    return error_.c_str();
}

void Server::stats(uint64_t* done, uint64_t* launches, uint64_t* peak) const {
    std::lock_guard<std::mutex> lk(mu_);
    if (done) *done = frames_done_;
    if (launches) *launches = launches_;
    if (peak) *peak = peak_bytes_;
}

Status Server::submit(const tr_frame* fr, uint64_t* ticket, bool device) {
    if (!fr || !ticket) return Status::Err(TR_ERR_INVALID_ARGUMENT, "null argument");
    FrameExtras x;
    x.weights = fr->weights;
    for (int i = 0; i < 4; ++i) x.init[i] = fr->initial_q[i];
    for (int i = 0; i < 3; ++i) x.init[4 + i] = fr->initial_t[i];
    // Host arrays' values are checked for finiteness while they are copied
    // below (one pass); device arrays, which the host cannot read without a
    // copy, are not.
    TR_TRY(validate_frame(fr->points2D, fr->points3D, fr->num_points, &fr->camera, &fr->ransac,
                          &fr->bundle, false, &x));
    if (device)
        TR_TRY(validate_device_arrays(device_, fr->points2D, fr->points3D, fr->weights, nullptr));
    auto r = std::make_shared<Request>();
    r->weighted = fr->weights != nullptr;
    r->device = device;
    for (int i = 0; i < 7; ++i) r->init[i] = x.init[i];
    r->n = fr->num_points;
    r->cam = fr->camera;
    r->ro = fr->ransac;
    r->bo = fr->bundle;
    if (r->ro.adaptive_stopping &&
        (r->ro.max_total_iterations + r->ro.max_iterations - 1) / r->ro.max_iterations >
            (uint64_t)kMaxDeviceRounds)
        return Status::Err(TR_ERR_INVALID_ARGUMENT,
                           "adaptive stopping when served: max_total_iterations may ask for at "
                           "most 32 rounds of max_iterations samples");
    // A launch of few frames may score in chunks down to kMinChunk points:
    // admission charges the partial scores at that size.
    r->bytes = request_bytes(r->n, r->ro.max_iterations, kMinChunk,
                             device ? 0 : (r->weighted ? 6 : 5));
    if (r->bytes > opt_.hbm_budget_bytes)
        return Status::Err(TR_ERR_CAPACITY, "frame alone exceeds the HBM budget");
    if (device) {
        r->d_uv = fr->points2D;
        r->d_xyz = fr->points3D;
        r->d_w = fr->weights;
    } else {
        r->data.resize((r->weighted ? 6 : 5) * (size_t)r->n);
        if (!copy_finite(r->data.data(), fr->points2D, 2 * (size_t)r->n))
            return Status::Err(TR_ERR_NOT_FINITE, "points2D not finite");
        if (!copy_finite(r->data.data() + 2 * r->n, fr->points3D, 3 * (size_t)r->n))
            return Status::Err(TR_ERR_NOT_FINITE, "points3D not finite");
        if (r->weighted && !copy_finite(r->data.data() + 5 * r->n, fr->weights, (size_t)r->n))
            return Status::Err(TR_ERR_NOT_FINITE, "weights not finite");
    }
    r->arrival = Clock::now();
    {
        std::lock_guard<std::mutex> lk(mu_);
        r->ticket = next_ticket_++;
        live_[r->ticket] = r;
        queue_.push_back(r);
        *ticket = r->ticket;
    }
    cv_work_.notify_one();
    return Status::Ok();
}

Status Server::wait(uint64_t ticket, int64_t timeout_ms, tr_pose_result* res, uint8_t* mask,
                    double* latency_us) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = live_.find(ticket);
    if (it == live_.end()) return Status::Err(TR_ERR_INVALID_ARGUMENT, "unknown ticket");
    std::shared_ptr<Request> r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_ms < 0)
        cv_done_.wait(lk, pred);
    else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
        return Status::Err(TR_ERR_NOT_READY, "not finished");
    if (res) *res = r->res;
    if (mask && !r->mask.empty()) std::memcpy(mask, r->mask.data(), r->mask.size());
    if (latency_us)
        *latency_us = std::chrono::duration<double, std::micro>(r->done_at - r->arrival).count();
    live_.erase(it);
    return r->res.status == TR_OK ? Status::Ok()
                                  : Status::Err((tr_status)r->res.status, "frame failed");
}

Status Server::replay(const tr_frame* frames, const double* arrival_s, int nf,
                      uint64_t* tickets, double* latency_us, double* wall_s,
                      tr_pose_result* results, uint8_t* const* masks, bool device) {
    if (!frames || !arrival_s || nf < 0 || !tickets)
        return Status::Err(TR_ERR_INVALID_ARGUMENT, "bad replay arguments");
    // The calling thread sleeps until each arrival.  A timer slack of 1 us
    // (instead of Linux's default 50 us) ends each sleep close to its
    // deadline; the thread's own slack is restored on return.
    const int old_slack = prctl(PR_GET_TIMERSLACK, 0, 0, 0, 0);
    (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);
    struct RestoreSlack {
        int v;
        ~RestoreSlack() {
            if (v > 0) (void)prctl(PR_SET_TIMERSLACK, (unsigned long)v, 0, 0, 0);
        }
    } restore{old_slack};
    const Clock::time_point t0 = Clock::now();
    for (int i = 0; i < nf; ++i) {
        const auto target =
            t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
        std::this_thread::sleep_until(target);
        TR_TRY(submit(&frames[i], &tickets[i], device));
        // Latency counts from the scheduled arrival, also when submit ran late.
        std::lock_guard<std::mutex> lk(mu_);
        live_[tickets[i]]->arrival = target;
    }
    Clock::time_point last = t0;
    Status first_err = Status::Ok();
    for (int i = 0; i < nf; ++i) {
        std::shared_ptr<Request> r;
        {
            std::lock_guard<std::mutex> lk(mu_);
            r = live_[tickets[i]];
        }
        double lat = 0.0;
        Status s = wait(tickets[i], -1, results ? &results[i] : nullptr,
                        masks ? masks[i] : nullptr, &lat);
        if (!s.ok() && first_err.ok()) first_err = s;
        if (latency_us) latency_us[i] = lat;
        if (r && r->done_at > last) last = r->done_at;
    }
    if (wall_s) *wall_s = std::chrono::duration<double>(last - t0).count();
    return first_err;
}

Status Server::launch_slot(Slot& s, std::vector<std::shared_ptr<Request>>& take) {
    const int nf = (int)take.size();
    size_t total = 0;
    for (auto& r : take) total += r->data.size();
    TR_TRY(s.in_host.reserve(sizeof(double) * total + 256));
    TR_TRY(s.in.reserve(sizeof(double) * total + 256));
    std::vector<FrameInput> in(nf);
    double* host = static_cast<double*>(s.in_host.ptr);
    const double* dev = static_cast<const double*>(s.in.ptr);
    size_t o = 0;
    for (int f = 0; f < nf; ++f) {
        const Request& r = *take[f];
        if (r.device) {
            // Read in place from the caller's device arrays.
            in[f].uv = r.d_uv;
            in[f].xyz = r.d_xyz;
            in[f].w = r.d_w;
        } else {
            std::memcpy(host + o, r.data.data(), sizeof(double) * r.data.size());
            in[f].uv = dev + o;
            in[f].xyz = dev + o + 2 * r.n;
            in[f].w = r.weighted ? dev + o + 5 * r.n : nullptr;
        }
        in[f].n = r.n;
        in[f].cam = r.cam;
        in[f].ro = r.ro;
        in[f].bo = r.bo;
        for (int i = 0; i < 7; ++i) in[f].init[i] = r.init[i];
        o += r.data.size();
    }
    int64_t max_n = 0;
    for (auto& r : take) max_n = std::max<int64_t>(max_n, r->n);
    const int si = pick_shape(nf, max_n);
    const ShapeChoice& sc = shapes_[si];
    TR_TRY(s.plan.build(in.data(), nf, sc.tuning, true, num_cus_, sc.occ));
    TR_TRY(s.arena.reserve(s.plan.arena_bytes()));
    TR_TRY(s.meta.reserve(s.plan.meta_bytes()));
    TR_TRY(s.meta_host.reserve(s.plan.meta_bytes()));
    const size_t masks_bytes =
        s.plan.mask_offset(nf - 1) + align_up_s((size_t)take.back()->n, 256) - s.plan.mask_offset(0);
    const size_t states_bytes = align_up_s(sizeof(FrameState) * nf, 256);
    TR_TRY(s.out_host.reserve(states_bytes + masks_bytes));
    s.plan.bind(static_cast<char*>(s.arena.ptr), static_cast<char*>(s.meta.ptr),
                static_cast<char*>(s.meta_host.ptr));
    if (o)
        TR_HIP(hipMemcpyAsync(s.in.ptr, s.in_host.ptr, sizeof(double) * o, hipMemcpyHostToDevice,
                              s.stream));
    TR_HIP(hipMemcpyAsync(s.meta.ptr, s.meta_host.ptr, s.plan.meta_bytes(), hipMemcpyHostToDevice,
                          s.stream));
    if (s.graph_meta != s.meta.ptr) {
        for (auto& kv : s.exec)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
        s.exec.clear();
        s.graph_meta = s.meta.ptr;
    }
    const int lml = s.plan.grid().lm_launches, rounds = s.plan.grid().rounds;
    hipGraphExec_t& exec = s.exec[{si, rounds, lml}];
    if (!exec) {
        LaunchGrid g = grid_;
        g.lm_launches = lml;
        g.rounds = rounds;
        g.score = num_cus_ * sc.occ;
        TR_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
        Status ls = launch_pipeline_fixed(s.plan.dev(), sc.shape, g, s.stream);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(s.stream, &gr);
        if (!ls.ok()) {
            if (gr) (void)hipGraphDestroy(gr);
            return ls;
        }
        TR_HIP(e);
        const hipError_t ei = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        TR_HIP(ei);
    }
    TR_HIP(hipGraphLaunch(exec, s.stream));
    char* out = static_cast<char*>(s.out_host.ptr);
    TR_HIP(hipMemcpyAsync(out, static_cast<char*>(s.arena.ptr) + s.plan.state_offset(),
                          sizeof(FrameState) * nf, hipMemcpyDeviceToHost, s.stream));
    if (opt_.return_masks)
        TR_HIP(hipMemcpyAsync(out + states_bytes,
                              static_cast<char*>(s.arena.ptr) + s.plan.mask_offset(0), masks_bytes,
                              hipMemcpyDeviceToHost, s.stream));
    TR_HIP(hipEventRecord(s.done, s.stream));
    s.reqs = take;
    return Status::Ok();
}

// Completes the requests of a slot whose launch finished (the waiter's
// thread): results and masks from the slot's host staging, then the
// waiting clients are woken.  The slot stays busy until the caller frees it.
Status Server::finish_slot(Slot& s) {
    const int nf = (int)s.reqs.size();
    const char* out = static_cast<const char*>(s.out_host.ptr);
    const FrameState* st = reinterpret_cast<const FrameState*>(out);
    const size_t states_bytes = align_up_s(sizeof(FrameState) * nf, 256);
    const auto now = Clock::now();
    for (int f = 0; f < nf; ++f) {
        Request& r = *s.reqs[f];
        r.res = st[f].result;
        if (opt_.return_masks) {
            const uint8_t* m = reinterpret_cast<const uint8_t*>(
                out + states_bytes + (s.plan.mask_offset(f) - s.plan.mask_offset(0)));
            r.mask.assign(m, m + r.n);
        }
        r.data.clear();
        r.data.shrink_to_fit();
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& r : s.reqs) {
            r->done = true;
            r->done_at = now;
            admitted_bytes_ -= r->bytes;
        }
        frames_done_ += (uint64_t)nf;
    }
    cv_done_.notify_all();
    s.reqs.clear();
    return Status::Ok();
}

// Waiter of one slot: sleeps until the slot has a launch outstanding,
// then polls the launch's event every kPollMicros (timed sleeps, with the
// thread's timer slack at 1 us so that a sleep ends close to its deadline),
// completes the requests itself and frees the slot.  Polling in sleeps
// costs a few percent of a core while a launch runs and nothing otherwise;
// waking the client straight from here saves the hand-off through the
// scheduler.  A slot has at most one launch outstanding.
static constexpr int kPollMicros = 10;

void Server::wait_slot(Slot* s) {
    (void)hipSetDevice(device_);
    (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        cv_slot_.wait(lk, [&] { return waiters_stop_ || s->launched != s->completed; });
        if (s->launched == s->completed) break;  // stopping, nothing outstanding
        const uint64_t gen = s->launched;
        lk.unlock();
        hipError_t e;
        while ((e = hipEventQuery(s->done)) == hipErrorNotReady)
            std::this_thread::sleep_for(std::chrono::microseconds(kPollMicros));
        Status fin = Status::Ok();
        if (e == hipSuccess) {
            try {
                fin = finish_slot(*s);
            } catch (const std::exception& ex) {  // e.g. a mask's host copy
                fin = Status::Err(TR_ERR_OUT_OF_MEMORY, ex.what());
            }
        }
        if (e != hipSuccess || !fin.ok()) {
            set_error(e != hipSuccess ? std::string("launch failed: ") + hipGetErrorString(e)
                                      : "completing a launch: " + fin.msg);
            std::lock_guard<std::mutex> g(mu_);
            for (auto& r : s->reqs) {
                if (r->done) continue;  // completed before the failure
                r->res.status = e != hipSuccess ? TR_ERR_DEVICE : fin.code;
                r->done = true;
                r->done_at = Clock::now();
                admitted_bytes_ -= r->bytes;
            }
            s->reqs.clear();
            cv_done_.notify_all();
        }
        lk.lock();
        s->completed = gen;
        s->busy = false;
        cv_work_.notify_all();
    }
}

// The scheduler: admits queued frames first-in first-out into a free slot
// while the HBM budget allows, and otherwise sleeps on cv_work_ until a
// frame arrives or a waiter frees a slot.  A slot's `busy` flag changes
// under the mutex only: set here with the launch, cleared by its waiter.
void Server::run() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    auto admissible = [&] {
        return !queue_.empty() && admitted_bytes_ + queue_.front()->bytes <= opt_.hbm_budget_bytes;
    };
    for (;;) {
        bool any_busy = false;
        Slot* free_slot = nullptr;
        for (auto& sp : slots_) {
            if (sp->busy)
                any_busy = true;
            else if (!free_slot)
                free_slot = sp.get();
        }
        std::vector<std::shared_ptr<Request>> take;
        if (free_slot) {
            while ((int)take.size() < opt_.max_batch_frames && admissible()) {
                admitted_bytes_ += queue_.front()->bytes;
                take.push_back(queue_.front());
                queue_.pop_front();
            }
            peak_bytes_ = std::max<uint64_t>(peak_bytes_, admitted_bytes_);
        }
        if (take.empty()) {
            if (stop_ && queue_.empty() && !any_busy) break;
            cv_work_.wait(lk, [&] {
                bool has_free = false;
                bool busy = false;
                for (auto& sp : slots_) (sp->busy ? busy : has_free) = true;
                return (has_free && admissible()) || (stop_ && queue_.empty() && !busy);
            });
            continue;
        }
        free_slot->busy = true;  // no other launch takes it meanwhile
        lk.unlock();
        Status s = Status::Ok();
        // No exception may leave the scheduler's thread: a failed host
        // allocation while planning fails the launch's requests instead.
        try {
            s = launch_slot(*free_slot, take);
        } catch (const std::bad_alloc&) {
            s = Status::Err(TR_ERR_OUT_OF_MEMORY, "host allocation failed while planning a launch");
        } catch (const std::exception& e) {
            s = Status::Err(TR_ERR_INTERNAL, e.what());
        }
        lk.lock();
        if (!s.ok()) {
            error_ = s.msg;
            for (auto& r : take) {
                r->res.status = s.code;
                r->done = true;
                r->done_at = Clock::now();
                admitted_bytes_ -= r->bytes;
            }
            free_slot->reqs.clear();
            free_slot->busy = false;
            cv_done_.notify_all();
        } else {
            ++launches_;
            ++free_slot->launched;
            cv_slot_.notify_all();
        }
    }
}

}  // namespace tr
