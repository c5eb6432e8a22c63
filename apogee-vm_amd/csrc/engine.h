// Shard-proving engine: a worker pool with per-worker HIP streams, admission
// against an HBM budget (FIFO or shortest-first), and per-job timing.
#pragma once
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "runtime.h"

namespace apg {

struct Job {
    uint64_t id = 0;
    size_t mem_bytes = 0;  // HBM the job holds while it runs (admission)
    double work = 0;       // size estimate (scalars), orders the queue
    std::function<void(Context &, Job &)> fn;
    int status = 1;  // 1 pending, 0 done, < 0 error code
    std::string error;
    int worker = -1;
    double t_submit = 0, t_start = 0, t_end = 0;  // ms, steady clock
    // outputs
    std::vector<uint8_t> commitments, rounds, finals, transcript_state;
    std::vector<uint8_t> transcript_blob;  // opaque copy of the final transcript
};

struct EngineStats {
    size_t queued = 0, running = 0, completed = 0;
    size_t hbm_in_use = 0, hbm_peak = 0, hbm_budget = 0, arena_bytes = 0;
};

// Which queued job a free worker takes.
//   kFifo:     arrival order; the head waits until its HBM estimate fits.
//   kShortest: the smallest job (by work) whose estimate fits, so short shards
//              do not queue behind long ones.
// With `reserve` > 0 the first `reserve` workers are a fast lane: they take
// only jobs of at most `small_work`, and the other workers leave such jobs to
// an idle fast-lane worker. Built with lane_cus > 0, the fast lane's streams
// run on CUs [0, lane_cus) and the rest on the other CUs, so a short shard's
// kernels never wait for a slot behind a long shard's.
enum class Policy { kFifo = 0, kShortest = 1 };

class Engine {
   public:
    Engine(int device, int workers, size_t hbm_budget, int lane_workers = 0, int lane_cus = 0);
    ~Engine();
    uint64_t submit(std::unique_ptr<Job> job);
    std::unique_ptr<Job> wait(uint64_t id);
    bool poll(uint64_t id);
    EngineStats stats();
    int workers() const { return (int)ctxs_.size(); }
    void set_policy(Policy p, int reserve, double small_work);
    // Kernel timing on every worker's stream (see Stats); summed over workers
    // as their jobs finish.
    void set_profiling(bool on);
    Stats kernel_stats();
    void reset_kernel_stats();

   private:
    void run(int w);
    // index into queue_ of the job free worker w takes now, or -1
    long pick_locked(int w) const;
    int device_;
    size_t budget_;
    Policy policy_ = Policy::kFifo;
    int reserve_ = 0;
    double small_work_ = 0;
    int idle_fast_ = 0;  // fast-lane workers waiting for a job
    Stats kstats_;
    std::vector<std::unique_ptr<Context>> ctxs_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::unique_ptr<Job>> queue_;
    std::unordered_map<uint64_t, std::unique_ptr<Job>> done_;
    uint64_t next_id_ = 1;
    size_t in_use_ = 0, peak_ = 0, running_ = 0, completed_ = 0;
    bool stop_ = false;
};

}  // namespace apg
