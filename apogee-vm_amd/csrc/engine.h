// Shard-proving engine: a worker pool with per-worker HIP streams and scratch
// arenas, admission against an HBM budget (FIFO or shortest-first, jobqueue.h),
// and per-job timing.
#pragma once
#include <atomic>
#include <memory>
#include <vector>

#include "runtime.h"
#include "jobqueue.h"

namespace apg {

using Job = BasicJob<Context>;

// With lane_workers > 0 and lane_cus > 0, the fast lane's streams run on CUs
// [0, lane_cus) and the other workers' on the remaining CUs, so a short
// shard's kernels never wait for a slot behind a long shard's.
class Engine {
   public:
    Engine(int device, int workers, size_t hbm_budget, int lane_workers = 0, int lane_cus = 0);
    ~Engine();
    uint64_t submit(std::unique_ptr<Job> job) { return sched_->submit(std::move(job)); }
    std::unique_ptr<Job> wait(uint64_t id) { return sched_->wait(id); }
    int poll(uint64_t id) { return sched_->poll(id); }
    EngineStats stats() { return sched_->stats(); }
    int workers() const { return (int)ctxs_.size(); }
    void set_policy(Policy p, int reserve, double small_work) {
        sched_->set_policy(p, reserve, small_work);
        small_work_ = small_work;
    }
    // jobs above this size (scalars) are long: their bucket accumulation
    // yields issue slots to the other kernels (0: no such jobs)
    double small_work() const { return small_work_; }
    // fast-lane jobs count for preemption only when the lane shares its CUs
    void set_preemption(double ratio, Preempt mode = Preempt::kSize) {
        sched_->set_preemption(ratio, !lane_split_, mode);
    }
    uint64_t pauses() { return sched_->pauses(); }
    // Kernel timing on every worker's stream (see Stats); summed over workers
    // as their jobs finish.
    void set_profiling(bool on);
    Stats kernel_stats();
    void reset_kernel_stats();

   private:
    int device_;
    bool lane_split_ = false;  // the fast lane runs on CUs of its own
    std::atomic<double> small_work_{0.0};
    Stats kstats_;  // written under the scheduler's lock
    std::vector<std::unique_ptr<Context>> ctxs_;
    std::unique_ptr<Scheduler<Context>> sched_;
};

}  // namespace apg
