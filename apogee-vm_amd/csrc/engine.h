// Shard-proving engine: a worker pool with per-worker HIP streams, FIFO
// admission against an HBM budget, and per-job timing.
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

class Engine {
   public:
    Engine(int device, int workers, size_t hbm_budget);
    ~Engine();
    uint64_t submit(std::unique_ptr<Job> job);
    std::unique_ptr<Job> wait(uint64_t id);
    bool poll(uint64_t id);
    EngineStats stats();
    int workers() const { return (int)ctxs_.size(); }

   private:
    void run(int w);
    int device_;
    size_t budget_;
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
