// Shard-proving engine: continuous batching of shard jobs on one GPU.
//
// The reference proves shards with `max_in_flight` scoped worker threads that
// claim shards as the streaming pipeline emits them
// (crates/prover/src/streaming.rs:601-680, default 8 in
// tools/bench/src/block.rs:129). Here each worker owns a HIP stream and a
// scratch arena (a Context); jobs wait in one queue and are admitted when a
// worker is free and the job's HBM estimate fits the budget beside the jobs
// already running, in arrival order or shortest first (Policy). Workers run
// concurrently, so one shard's latency-bound
// phases (bucket-reduction trees, sumcheck rounds waiting on the host
// transcript) overlap other shards' throughput-bound kernels. A job's result
// depends only on its own inputs and transcript: the same bytes alone or in
// any mix.
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

Engine::Engine(int device, int workers, size_t hbm_budget) : device_(device), budget_(hbm_budget) {
    if (workers < 1) workers = 1;
    ctxs_.resize(workers);
    for (int w = 0; w < workers; w++) {
        ctxs_[w] = std::make_unique<Context>();
        context_init(*ctxs_[w], device);
    }
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

long Engine::pick_locked() const {
    if (queue_.empty()) return -1;
    if (policy_ == Policy::kFifo)
        return in_use_ + queue_.front()->mem_bytes <= budget_ ? 0 : -1;
    const bool large_ok = running_large_ + reserve_ < ctxs_.size();
    long best = -1;
    for (size_t i = 0; i < queue_.size(); i++) {
        const Job &j = *queue_[i];
        if (in_use_ + j.mem_bytes > budget_) continue;
        if (!large_ok && j.work > small_work_) continue;
        if (best < 0 || j.work < queue_[best]->work) best = (long)i;
    }
    return best;
}

void Engine::run(int w) {
    Context &ctx = *ctxs_[w];
    HIP_CHECK(hipSetDevice(device_));
    for (;;) {
        std::unique_ptr<Job> job;
        bool large = false;
        {
            std::unique_lock<std::mutex> lk(mu_);
            long at = -1;
            // queued jobs still run after stop_; every queued job fits an
            // idle engine (submit refuses the rest), so the queue drains
            cv_.wait(lk, [&] {
                at = pick_locked();
                return at >= 0 || (stop_ && queue_.empty());
            });
            if (at < 0) return;
            job = std::move(queue_[at]);
            queue_.erase(queue_.begin() + at);
            in_use_ += job->mem_bytes;
            peak_ = std::max(peak_, in_use_);
            running_++;
            large = job->work > small_work_;
            if (large) running_large_++;
        }
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
            if (large) running_large_--;
            completed_++;
            done_[job->id] = std::move(job);
        }
        done_cv_.notify_all();
        cv_.notify_all();
    }
}

}  // namespace apg
