// The shard engine's job queue and admission, free of HIP so that the host
// part of the engine can be built alone: the engine (engine.hip) instantiates
// it over its worker contexts, and a host-only stress test builds it with the
// thread and address sanitizers over mock workers.
//
// Jobs wait in one queue. A free worker takes the job its policy picks
// (Policy) once that job's HBM charge fits the budget. A worker's charge is
// the device memory its context holds (its scratch arena, live and retired
// buffers), or, while it runs a job, the larger of that and the job's
// estimate: a job reuses its worker's arena, so it adds only what it needs
// beyond it. When nothing is running, the picked job is admitted whatever
// the arenas hold, so the queue always drains.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "error.h"

namespace apg {

// Which queued job a free worker takes.
//   kFifo:     arrival order; the head waits until its charge fits.
//   kShortest: the smallest job (by work) whose charge fits, so short shards
//              do not queue behind long ones.
//   kLongest:  the largest job whose charge fits (longest processing time
//              first): for a batch submitted at once, the long shards start
//              early and the short ones fill in around them, so the batch
//              does not end on one long shard running alone.
// With `reserve` > 0 the first `reserve` workers are a fast lane: they take
// only jobs of at most `small_work`, and the other workers leave such jobs to
// an idle fast-lane worker.
//
// Preemption at phase boundaries (set_preemption, ratio > 0): a job running
// on a worker outside the fast lane pauses at its next yield point (a phase
// boundary of the shard: between MSM phases and accumulation chunks, between
// sumcheck rounds) while another worker runs a job whose work times the ratio
// is below its own, and resumes when none does. Jobs on the fast lane count
// only when the lane shares its CUs with the other workers (count_lane). The
// device has no preemption of its own, so the job's host thread drains its
// stream at each yield point; a smaller job that starts beside it has the
// CUs to itself once the phase in flight ends. A paused job
// keeps its worker, so a job only pauses for one that already runs (never
// for a queued one, which might wait for that very worker), and the job
// that comes first in the order (the smallest, or the most urgent) never
// pauses: some job always runs, and no cycle of waits can form.
enum class Policy { kFifo = 0, kShortest = 1, kLongest = 2 };

// What a running job gives way to (set_preemption):
//   kSize:    a job whose work times the ratio is below its own;
//   kUrgency: a job whose urgency exceeds its own times the ratio, the
//             urgency of a job being the latency over its isolated time it
//             would reach if it ran alone from now on, (elapsed + remaining)
//             / isolated, with remaining = isolated x (1 - progress). A job
//             falling behind its isolated time gains priority, so the
//             latency-over-isolated ratios even out across sizes. A paused
//             job re-checks every millisecond, since urgencies move with the
//             clock.
enum class Preempt { kSize = 0, kUrgency = 1 };

struct EngineStats {
    size_t queued = 0, running = 0, completed = 0;
    size_t hbm_in_use = 0, hbm_peak = 0, hbm_budget = 0, arena_bytes = 0;
};

template <class Ctx>
struct BasicJob {
    uint64_t id = 0;
    size_t mem_bytes = 0;  // HBM the job needs on its worker while it runs (arena included)
    double work = 0;       // size estimate (scalars), orders the queue
    double iso_ms = 0;     // estimated time alone on the device (urgency preemption)
    std::function<void(Ctx &, BasicJob &)> fn;
    int status = 1;  // 1 pending, 0 done, < 0 error code
    std::string error;
    int worker = -1;
    double t_submit = 0, t_start = 0, t_end = 0;  // ms, steady clock
    // outputs
    std::vector<uint8_t> commitments, rounds, finals;
    std::vector<uint8_t> transcript_blob;  // opaque copy of the final transcript
};

// What the scheduler needs of a worker's context.
template <class Ctx>
struct WorkerHooks {
    // on the worker's thread before its first job; if it throws, every job
    // that worker takes fails with that error
    std::function<void(int, Ctx &)> start;
    // device bytes the context holds now (read from other threads)
    std::function<size_t(const Ctx &)> held;
    // after each job, under the scheduler's lock (kernel statistics)
    std::function<void(Ctx &)> collect;
    // on the worker's thread before its first job: hands the context the
    // worker's preemption calls (whether preemption is on, whether the job
    // should pause now, the pause itself, and the job's progress report)
    std::function<void(Ctx &, std::function<bool()>, std::function<bool()>, std::function<void()>,
                       std::function<void(double)>)>
        bind_yield;
};

inline double sched_now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

template <class Ctx>
class Scheduler {
   public:
    using Job = BasicJob<Ctx>;

    // ctxs outlive the scheduler; workers start at once
    Scheduler(std::vector<Ctx *> ctxs, size_t budget, int reserve, WorkerHooks<Ctx> hooks)
        : ctxs_(std::move(ctxs)), budget_(budget), hooks_(std::move(hooks)) {
        const int nw = (int)ctxs_.size();
        reserve_ = std::max(0, std::min(reserve, nw - 1));
        running_mem_.assign(nw, 0);
        running_work_.assign(nw, 0.0);
        progress_.assign(nw, 0.0);
        t_submit_.assign(nw, 0.0);
        iso_.assign(nw, 0.0);
        busy_.assign(nw, false);
        paused_.assign(nw, false);
        for (int w = 0; w < nw; w++) threads_.emplace_back([this, w] { run(w); });
    }

    // Queued jobs still run; then the workers stop.
    ~Scheduler() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        pause_cv_.notify_all();
        for (auto &t : threads_) t.join();
    }

    uint64_t submit(std::unique_ptr<Job> job) {
        std::lock_guard<std::mutex> lk(mu_);
        job->id = next_id_++;
        job->t_submit = sched_now_ms();
        const uint64_t id = job->id;
        outstanding_.insert(id);
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

    // Blocks until job `id` is done and hands it over; an id never issued or
    // already collected is an argument error.
    std::unique_ptr<Job> wait(uint64_t id) {
        std::unique_lock<std::mutex> lk(mu_);
        if (!outstanding_.count(id)) throw ApgError(-2, "engine: unknown or already collected job id");
        done_cv_.wait(lk, [&] { return done_.count(id) != 0; });
        auto j = std::move(done_[id]);
        done_.erase(id);
        outstanding_.erase(id);
        return j;
    }

    // 1 done, 0 pending, -1 unknown or already collected
    int poll(uint64_t id) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!outstanding_.count(id)) return -1;
        return done_.count(id) ? 1 : 0;
    }

    EngineStats stats() {
        std::lock_guard<std::mutex> lk(mu_);
        EngineStats s;
        s.queued = queue_.size();
        s.running = running_;
        s.completed = completed_;
        s.hbm_in_use = in_use_locked();
        peak_ = std::max(peak_, s.hbm_in_use);
        s.hbm_peak = peak_;
        s.hbm_budget = budget_;
        for (Ctx *c : ctxs_) s.arena_bytes += hooks_.held(*c);
        return s;
    }

    void set_policy(Policy p, int reserve, double small_work) {
        std::lock_guard<std::mutex> lk(mu_);
        policy_ = p;
        reserve_ = std::max(0, std::min(reserve, (int)ctxs_.size() - 1));
        small_work_ = small_work;
        cv_.notify_all();
        pause_cv_.notify_all();
    }

    // ratio 0 turns preemption off (paused jobs resume); otherwise >= 1
    void set_preemption(double ratio, bool count_lane = false, Preempt mode = Preempt::kSize) {
        std::lock_guard<std::mutex> lk(mu_);
        preempt_ratio_ = ratio > 0 ? std::max(1.0, ratio) : 0.0;
        count_lane_ = count_lane;
        mode_ = mode;
        preempt_on_ = preempt_ratio_ > 0;
        pause_cv_.notify_all();
    }

    bool preemption_on() const { return preempt_on_.load(std::memory_order_relaxed); }

    // whether the job on worker w should pause at this yield point
    bool yield_wanted(int w) {
        std::lock_guard<std::mutex> lk(mu_);
        return yield_wanted_locked(w);
    }

    // pauses the job on worker w until no job it gives way to runs beside it
    void yield_wait(int w) {
        std::unique_lock<std::mutex> lk(mu_);
        if (!yield_wanted_locked(w)) return;
        pauses_++;
        paused_[w] = true;
        // woken by every job's end and every policy change, and re-checked
        // each millisecond, since urgencies move with the clock (a timed wait
        // on the system clock: pthread_cond_timedwait, which the thread
        // sanitizer models; a steady-clock wait it does not)
        while (yield_wanted_locked(w))
            pause_cv_.wait_until(lk, std::chrono::system_clock::now() + std::chrono::milliseconds(1));
        paused_[w] = false;
    }

    // the fraction of its work the job on worker w has done
    void report_progress(int w, double p) {
        std::lock_guard<std::mutex> lk(mu_);
        progress_[w] = std::min(1.0, std::max(0.0, p));
    }

    // pauses so far (for tests and statistics)
    uint64_t pauses() {
        std::lock_guard<std::mutex> lk(mu_);
        return pauses_;
    }

    // runs f under the scheduler's lock (for state the collect hook writes)
    template <class F>
    void locked(F &&f) {
        std::lock_guard<std::mutex> lk(mu_);
        f();
    }

    int workers() const { return (int)ctxs_.size(); }

   private:
    double urgency_locked(int v, double now) const {
        const double iso = std::max(iso_[v], 0.05);
        return (now - t_submit_[v] + iso * (1.0 - progress_[v])) / iso;
    }

    bool yield_wanted_locked(int w) const {
        if (preempt_ratio_ <= 0 || stop_ || w < reserve_ || !busy_[w]) return false;
        if (mode_ == Preempt::kUrgency) {
            const double now = sched_now_ms(), uw = urgency_locked(w, now) * preempt_ratio_;
            for (int v = count_lane_ ? 0 : reserve_; v < (int)ctxs_.size(); v++)
                if (v != w && busy_[v] && !paused_[v] && urgency_locked(v, now) > uw) return true;
            return false;
        }
        for (int v = count_lane_ ? 0 : reserve_; v < (int)ctxs_.size(); v++)
            if (v != w && busy_[v] && !paused_[v] && running_work_[v] * preempt_ratio_ < running_work_[w])
                return true;
        return false;
    }

    size_t charge_locked(int w) const {
        const size_t h = hooks_.held(*ctxs_[w]);
        return busy_[w] ? std::max(h, running_mem_[w]) : h;
    }

    size_t in_use_locked() const {
        size_t s = 0;
        for (int w = 0; w < (int)ctxs_.size(); w++) s += charge_locked(w);
        return s;
    }

    // whether job j fits on free worker w beside what the others hold
    bool fits_locked(const Job &j, int w) const {
        if (running_ == 0) return true;
        const size_t h = hooks_.held(*ctxs_[w]);
        return in_use_locked() - h + std::max(h, j.mem_bytes) <= budget_;
    }

    // index into queue_ of the job free worker w takes now, or -1
    long pick_locked(int w) const {
        const bool fast = w < reserve_;
        long best = -1;
        for (size_t i = 0; i < queue_.size(); i++) {
            const Job &j = *queue_[i];
            const bool small = j.work <= small_work_;
            if (fast && !small) continue;                                    // the fast lane runs short jobs only
            if (!fast && small && reserve_ > 0 && idle_fast_ > 0) continue;  // left to the fast lane
            if (policy_ == Policy::kFifo)  // this lane's oldest job, once it fits
                return fits_locked(j, w) ? (long)i : -1;
            if (!fits_locked(j, w)) continue;
            if (best < 0 || (policy_ == Policy::kLongest ? j.work > queue_[best]->work : j.work < queue_[best]->work))
                best = (long)i;
        }
        return best;
    }

    void run(int w) {
        Ctx &ctx = *ctxs_[w];
        std::string start_error;
        int start_code = 0;
        try {
            if (hooks_.bind_yield)
                hooks_.bind_yield(
                    ctx, [this] { return preemption_on(); }, [this, w] { return yield_wanted(w); },
                    [this, w] { yield_wait(w); }, [this, w](double p) { report_progress(w, p); });
            if (hooks_.start) hooks_.start(w, ctx);
        } catch (const ApgError &e) {
            start_code = e.code;
            start_error = e.what();
        } catch (const std::exception &e) {
            start_code = -5;
            start_error = e.what();
        }
        for (;;) {
            std::unique_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                long at = -1;
                const bool fast = w < reserve_;
                if (fast) idle_fast_++;
                // every queued job fits an idle engine, and the non-fast
                // workers take any job, so the queue drains after stop_ too
                cv_.wait(lk, [&] {
                    at = pick_locked(w);
                    return at >= 0 || (stop_ && queue_.empty());
                });
                if (fast) idle_fast_--;
                if (at < 0) return;
                job = std::move(queue_[at]);
                queue_.erase(queue_.begin() + at);
                busy_[w] = true;
                running_mem_[w] = job->mem_bytes;
                running_work_[w] = job->work;
                progress_[w] = 0.0;
                t_submit_[w] = job->t_submit;
                iso_[w] = job->iso_ms;
                running_++;
                peak_ = std::max(peak_, in_use_locked());
            }
            cv_.notify_all();  // idle_fast_ changed: others may now take a short job
            job->t_start = sched_now_ms();
            job->worker = w;
            if (start_code) {
                job->status = start_code;
                job->error = "engine worker failed to start: " + start_error;
            } else {
                try {
                    job->fn(ctx, *job);
                    job->status = 0;
                } catch (const ApgError &e) {
                    job->status = e.code;
                    job->error = e.what();
                } catch (const std::exception &e) {
                    job->status = -5;
                    job->error = e.what();
                } catch (...) {
                    job->status = -5;
                    job->error = "unknown exception in an engine job";
                }
            }
            job->t_end = sched_now_ms();
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (hooks_.collect) hooks_.collect(ctx);
                busy_[w] = false;
                running_mem_[w] = 0;
                running_work_[w] = 0;
                running_--;
                completed_++;
                done_[job->id] = std::move(job);
            }
            done_cv_.notify_all();
            cv_.notify_all();
            pause_cv_.notify_all();  // a paused job may resume
        }
    }

    std::vector<Ctx *> ctxs_;
    size_t budget_;
    WorkerHooks<Ctx> hooks_;
    Policy policy_ = Policy::kFifo;
    int reserve_ = 0;
    double small_work_ = 0;
    int idle_fast_ = 0;  // fast-lane workers waiting for a job
    std::vector<size_t> running_mem_;
    std::vector<double> running_work_, progress_, t_submit_, iso_;
    std::vector<bool> busy_, paused_;
    double preempt_ratio_ = 0;
    bool count_lane_ = false;
    Preempt mode_ = Preempt::kSize;
    std::atomic<bool> preempt_on_{false};
    uint64_t pauses_ = 0;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_, pause_cv_;
    std::deque<std::unique_ptr<Job>> queue_;
    std::unordered_map<uint64_t, std::unique_ptr<Job>> done_;
    std::unordered_set<uint64_t> outstanding_;  // issued and not yet collected
    uint64_t next_id_ = 1;
    size_t peak_ = 0, running_ = 0, completed_ = 0;
    bool stop_ = false;
};

}  // namespace apg
