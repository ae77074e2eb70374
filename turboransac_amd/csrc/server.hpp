// server.hpp - continuous-batching server over the batch pipeline.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "engine.hpp"

namespace tr {

class Server {
  public:
    static Status create(int device, const tr_server_options& opt, Server** out);
    ~Server();
    // device: the frame's arrays are device addresses, read in place by the
    // launch (no copy); otherwise they are host arrays, copied at submit.
    Status submit(const tr_frame* frame, uint64_t* ticket, bool device = false);
    Status wait(uint64_t ticket, int64_t timeout_ms, tr_pose_result* res, uint8_t* mask,
                double* latency_us);
    Status replay(const tr_frame* frames, const double* arrival_s, int nf, uint64_t* tickets,
                  double* latency_us, double* wall_s, tr_pose_result* results,
                  uint8_t* const* masks, bool device = false);
    void stats(uint64_t* done, uint64_t* launches, uint64_t* peak) const;
    const char* last_error() const;

  private:
    using Clock = std::chrono::steady_clock;
    struct Request {
        uint64_t ticket = 0;
        std::vector<double> data;  // uv (2N), xyz (3N), then the weights (N) if given
        bool weighted = false;
        // A device-resident frame: its arrays in the caller's device memory,
        // read in place (data stays empty).
        bool device = false;
        const double* d_uv = nullptr;
        const double* d_xyz = nullptr;
        const double* d_w = nullptr;
        double init[7] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // initial pose q, t
        int64_t n = 0;
        tr_camera cam{};
        tr_ransac_options ro{};
        tr_bundle_options bo{};
        size_t bytes = 0;  // device footprint
        Clock::time_point arrival;
        Clock::time_point done_at;
        bool done = false;
        tr_pose_result res{};
        std::vector<uint8_t> mask;
    };
    // One in-flight launch: its own stream, buffers and captured graph.
    struct Slot {
        hipStream_t stream = nullptr;
        DevBuf in, arena, meta;
        HostBuf in_host, meta_host, out_host;
        // Captured graphs by (scoring shape, sampling rounds, LM launches),
        // all for meta address graph_meta.
        std::map<std::tuple<int, int, int>, hipGraphExec_t> exec;
        void* graph_meta = nullptr;
        hipEvent_t done = nullptr;  // recorded after the launch and its copies
        // Set by the scheduler with a launch, cleared by the slot's waiter
        // thread once the launch's requests are complete (under the mutex).
        bool busy = false;
        // launched counts the slot's launches, completed the ones its waiter
        // has finished (both under the server's mutex).
        std::thread waiter;
        uint64_t launched = 0, completed = 0;
        std::vector<std::shared_ptr<Request>> reqs;
        size_t bytes = 0;
        Plan plan;
    };

    void run();
    void wait_slot(Slot* s);
    Status launch_slot(Slot& s, std::vector<std::shared_ptr<Request>>& take);
    Status finish_slot(Slot& s);
    void set_error(const std::string& m);

    // The MSAC shape of a launch: [0] the tuning table's batch row; [1 + i]
    // row i of its one-frame table, for launches of at most
    // kSmallLaunchFrames frames (common.hpp; the batch kernel fills the
    // device only with many frames' hypotheses).
    struct ShapeChoice {
        int64_t max_n;      // frames up to this N
        ScoreShape shape;
        Tuning tuning;      // a table whose batch row is `shape`
        int occ;            // resident scoring workgroups per CU
    };
    int pick_shape(int nf, int64_t max_n) const;

    int device_ = 0;
    int num_cus_ = 0;
    tr_server_options opt_{};
    std::vector<ShapeChoice> shapes_;
    LaunchGrid grid_{};
    std::vector<std::unique_ptr<Slot>> slots_;

    mutable std::mutex mu_;
    // cv_work_: new requests or a freed slot for the scheduler; cv_slot_: a
    // new launch for the waiters; cv_done_: finished requests.
    std::condition_variable cv_work_, cv_done_, cv_slot_;
    std::deque<std::shared_ptr<Request>> queue_;
    std::map<uint64_t, std::shared_ptr<Request>> live_;
    uint64_t next_ticket_ = 1;
    size_t admitted_bytes_ = 0;
    uint64_t frames_done_ = 0, launches_ = 0, peak_bytes_ = 0;
    bool stop_ = false;          // no new work: drain the queue, then end the scheduler
    bool waiters_stop_ = false;  // the scheduler has ended: waiters end when idle
    std::string error_;
    std::thread worker_;
};

}  // namespace tr
