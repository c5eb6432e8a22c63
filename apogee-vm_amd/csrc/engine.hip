// Shard-proving engine: continuous batching of shard jobs on one GPU.
//
// The reference proves shards with `max_in_flight` scoped worker threads that
// claim shards as the streaming pipeline emits them
// (crates/prover/src/streaming.rs:601-680, default 8 in
// tools/bench/src/block.rs:129). Here each worker owns one HIP stream (one
// hardware queue) and a scratch arena (a Context); jobs wait in one queue and
// are admitted when a worker is free and the job's HBM estimate fits the
// budget beside the jobs already running, in arrival order or shortest first
// (Policy). Workers run concurrently, so one shard's latency-bound phases
// (bucket-reduction trees, sumcheck rounds waiting on the host transcript)
// overlap other shards' throughput-bound kernels.
//
// Latency: the device services about five of a process's busy hardware queues
// at once and time-slices the rest in ~1 ms slices, so a serving engine keeps
// workers + other busy streams within that (four workers measured best), and
// a fast lane (workers on their own CUs, short jobs only) keeps short shards
// from waiting for execution slots behind long ones. A job's result depends
// only on its own inputs and transcript: the same bytes alone or in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

namespace apg {

namespace {
double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
}  // namespace

Engine::Engine(int device, int workers, size_t hbm_budget, int lane_workers, int lane_cus)
    : device_(device), budget_(hbm_budget) {
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
    }
    ctxs_.resize(workers);
    for (int w = 0; w < workers; w++) {
        ctxs_[w] = std::make_unique<Context>();
        context_init(*ctxs_[w], device, w < lane_workers ? fast_mask : slow_mask, /*side_stream=*/false);
    }
    reserve_ = lane_workers;
    for (int w = 0; w < workers; w++) threads_.emplace_back([this, w] { run(w); });
}

Engine::~Engine() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : threads_) t.join();
    for (auto &c : ctxs_) context_release(*c);
}

uint64_t Engine::submit(std::unique_ptr<Job> job) {
    std::lock_guard<std::mutex> lk(mu_);
    job->id = next_id_++;
    job->t_submit = now_ms();
    const uint64_t id = job->id;
    if (job->mem_bytes > budget_) {
        job->status = -2;
        job->error = "job needs more HBM than the engine's budget";
        job->t_start = job->t_end = job->t_submit;
        done_[id] = std::move(job);
        done_cv_.notify_all();
        return id;
    }
    queue_.push_back(std::move(job));
    cv_.notify_all();
    return id;
}

std::unique_ptr<Job> Engine::wait(uint64_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return done_.count(id) != 0; });
    auto j = std::move(done_[id]);
    done_.erase(id);
    return j;
}

bool Engine::poll(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    return done_.count(id) != 0;
}

EngineStats Engine::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    EngineStats s;
    s.queued = queue_.size();
    s.running = running_;
    s.completed = completed_;
    s.hbm_in_use = in_use_;
    s.hbm_peak = peak_;
    s.hbm_budget = budget_;
    size_t arena = 0;
    for (auto &c : ctxs_)
        for (auto &kv : c->dev) arena += kv.second.cap;
    s.arena_bytes = arena;
    return s;
}

void Engine::set_policy(Policy p, int reserve, double small_work) {
    std::lock_guard<std::mutex> lk(mu_);
    policy_ = p;
    reserve_ = std::max(0, std::min(reserve, (int)ctxs_.size() - 1));
    small_work_ = small_work;
    cv_.notify_all();
}

void Engine::set_profiling(bool on) {
    for (auto &c : ctxs_) c->profiling = on;
}

Stats Engine::kernel_stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return kstats_;
}

void Engine::reset_kernel_stats() {
    std::lock_guard<std::mutex> lk(mu_);
    kstats_ = Stats();
}

long Engine::pick_locked(int w) const {
    const bool fast = w < reserve_;
    long best = -1;
    for (size_t i = 0; i < queue_.size(); i++) {
        const Job &j = *queue_[i];
        const bool small = j.work <= small_work_;
        if (fast && !small) continue;                      // the fast lane runs short jobs only
        if (!fast && small && reserve_ > 0 && idle_fast_ > 0) continue;  // left to the fast lane
        if (policy_ == Policy::kFifo)  // this lane's oldest job, once it fits
            return in_use_ + j.mem_bytes <= budget_ ? (long)i : -1;
        if (in_use_ + j.mem_bytes > budget_) continue;
        if (best < 0 || j.work < queue_[best]->work) best = (long)i;
    }
    return best;
}

void Engine::run(int w) {
    Context &ctx = *ctxs_[w];
    HIP_CHECK(hipSetDevice(device_));
    for (;;) {
        std::unique_ptr<Job> job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            long at = -1;
            const bool fast = w < reserve_;
            if (fast) idle_fast_++;
            // queued jobs still run after stop_; every queued job fits an
            // idle engine (submit refuses the rest) and the non-fast workers
            // take any job, so the queue drains
            cv_.wait(lk, [&] {
                at = pick_locked(w);
                return at >= 0 || (stop_ && queue_.empty());
            });
            if (fast) idle_fast_--;
            if (at < 0) return;
            job = std::move(queue_[at]);
            queue_.erase(queue_.begin() + at);
            in_use_ += job->mem_bytes;
            peak_ = std::max(peak_, in_use_);
            running_++;
        }
        cv_.notify_all();  // idle_fast_ changed: others may now take a short job
        job->t_start = now_ms();
        job->worker = w;
        try {
            job->fn(ctx, *job);
            job->status = 0;
        } catch (const ApgError &e) {
            job->status = e.code;
            job->error = e.what();
        } catch (const std::exception &e) {
            job->status = -5;
            job->error = e.what();
        }
        job->t_end = now_ms();
        {
            std::lock_guard<std::mutex> lk(mu_);
            kstats_ += ctx.stats;
            ctx.stats = Stats();
            in_use_ -= job->mem_bytes;
            running_--;
            completed_++;
            done_[job->id] = std::move(job);
        }
        done_cv_.notify_all();
        cv_.notify_all();
    }
}

}  // namespace apg
