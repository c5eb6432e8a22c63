// Shard-proving engine: continuous batching of shard jobs on one GPU.
//
// The reference proves shards with `max_in_flight` scoped worker threads that
// claim shards as the streaming pipeline emits them
// (crates/prover/src/streaming.rs:601-680, default 8 in
// tools/bench/src/block.rs:129). Here each worker owns a HIP stream and a
// scratch arena (a Context); jobs wait in one FIFO queue and are admitted when
// a worker is free and the job's HBM estimate fits the budget beside the jobs
// already running. Workers run concurrently, so one shard's latency-bound
// phases (bucket-reduction trees, sumcheck rounds waiting on the host
// transcript) overlap other shards' throughput-bound kernels. A job's result
// depends only on its own inputs and transcript: the same bytes alone or in
// any mix.
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

void Engine::run(int w) {
    Context &ctx = *ctxs_[w];
    HIP_CHECK(hipSetDevice(device_));
    for (;;) {
        std::unique_ptr<Job> job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            // FIFO admission: the head job starts when its HBM estimate fits
            cv_.wait(lk, [&] {
                return stop_ || (!queue_.empty() && in_use_ + queue_.front()->mem_bytes <= budget_);
            });
            if (stop_ && queue_.empty()) return;
            if (queue_.empty()) continue;
            job = std::move(queue_.front());
            queue_.pop_front();
            in_use_ += job->mem_bytes;
            peak_ = std::max(peak_, in_use_);
            running_++;
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
