// Shard-proving engine: continuous batching of shard jobs on one GPU.
//
// The reference proves shards with `max_in_flight` scoped worker threads that
// claim shards as the streaming pipeline emits them
// (crates/prover/src/streaming.rs:601-680, default 8 in
// tools/bench/src/block.rs:129). Here each worker owns one HIP stream (one
// hardware queue) and a scratch arena (a Context); jobs wait in one queue and
// are admitted when a worker is free and the job's HBM charge fits the budget
// (jobqueue.h), in arrival order or shortest first (Policy). Workers run
// concurrently, so one shard's latency-bound phases (bucket-reduction trees,
// sumcheck rounds waiting on the host transcript) overlap other shards'
// throughput-bound kernels.
//
// Latency: the device services about five of a process's busy hardware queues
// at once and time-slices the rest in ~1 ms slices, so a serving engine keeps
// workers + other busy streams within that (four workers measured best), and
// a fast lane (workers on their own CUs, short jobs only) keeps short shards
// from waiting for execution slots behind long ones. A job's result depends
// only on its own inputs and transcript: the same bytes alone or in any mix.
#include <algorithm>

#include "engine.h"

namespace apg {

Engine::Engine(int device, int workers, size_t hbm_budget, int lane_workers, int lane_cus) : device_(device) {
    if (workers < 1) workers = 1;
    lane_workers = std::max(0, std::min(lane_workers, workers - 1));
    std::vector<uint32_t> fast_mask, slow_mask;
    if (lane_workers > 0 && lane_cus > 0) {
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, device));
        const int ncu = prop.multiProcessorCount;
        lane_cus = std::min(lane_cus, ncu - 1);
        fast_mask.assign((ncu + 31) / 32, 0u);
        slow_mask.assign((ncu + 31) / 32, 0u);
        for (int cu = 0; cu < ncu; cu++) (cu < lane_cus ? fast_mask : slow_mask)[cu / 32] |= 1u << (cu % 32);
        lane_split_ = true;
    }
    ctxs_.resize(workers);
    std::vector<Context *> raw;
    try {
        for (int w = 0; w < workers; w++) {
            ctxs_[w] = std::make_unique<Context>();
            context_init(*ctxs_[w], device, w < lane_workers ? fast_mask : slow_mask, /*side_stream=*/false);
            raw.push_back(ctxs_[w].get());
        }
    } catch (...) {
        for (auto &c : ctxs_)
            if (c) context_release(*c);
        throw;
    }
    WorkerHooks<Context> hooks;
    hooks.start = [this](int, Context &) { HIP_CHECK(hipSetDevice(device_)); };
    hooks.held = [](const Context &c) { return c.held_bytes.load(); };
    hooks.collect = [this](Context &c) {
        kstats_ += c.stats;
        c.stats = Stats();
    };
    hooks.bind_yield = [](Context &c, std::function<bool()> on, std::function<bool()> wanted,
                          std::function<void()> wait, std::function<void(double)> progress) {
        c.yield_on = std::move(on);
        c.yield_wanted = std::move(wanted);
        c.yield_wait = std::move(wait);
        c.yield_progress = std::move(progress);
    };
    sched_ = std::make_unique<Scheduler<Context>>(raw, hbm_budget, lane_workers, std::move(hooks));
}

Engine::~Engine() {
    sched_.reset();  // drains the queue and joins the workers
    for (auto &c : ctxs_) context_release(*c);
}

void Engine::set_profiling(bool on) {
    for (auto &c : ctxs_) c->profiling = on;
}

Stats Engine::kernel_stats() {
    Stats s;
    sched_->locked([&] { s = kstats_; });
    return s;
}

void Engine::reset_kernel_stats() {
    sched_->locked([&] { kstats_ = Stats(); });
}

}  // namespace apg
