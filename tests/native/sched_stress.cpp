// Host-only stress test of the shard engine's scheduler (jobqueue.h) over mock
// workers, built with -fsanitize=thread or -fsanitize=address by
// tests/test_engine_sched.py. A mock worker's context holds a byte count that
// a job raises to its estimate, as a scratch arena grows; each job's output
// depends only on its own input. Checks, printing one line per check:
//   * every job completes with its own output, from several submitting
//     threads at once, under both policies and with a fast lane;
//   * with two or more jobs running, the charged HBM never exceeds the budget,
//     even though idle workers keep their arenas;
//   * fast-lane workers never run a long job;
//   * wait and poll refuse ids never issued or already collected;
//   * a worker that fails to start fails its jobs with that error while the
//     others serve;
//   * destroying the scheduler with jobs queued runs them first;
//   * with preemption on, a long job makes no progress while a much shorter
//     one runs beside it, resumes after it, and mixes with yield points in
//     every job still complete with their own outputs (no cycle of pauses);
//     with preemption off nothing pauses.
#include <atomic>
#include <chrono>
#include <functional>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "jobqueue.h"

namespace {

struct MockCtx {
    std::atomic<size_t> held{0};
    int index = 0;
    std::function<bool()> on, wanted;
    std::function<void()> wait;
    std::function<void(double)> progress;
    void yield_point(double p = -1) {
        if (!(on && on())) return;
        if (p >= 0) progress(p);
        if (wanted()) wait();
    }
};

void bind(apg::WorkerHooks<MockCtx> &hooks) {
    hooks.bind_yield = [](MockCtx &c, std::function<bool()> on, std::function<bool()> wanted,
                          std::function<void()> wait, std::function<void(double)> progress) {
        c.on = std::move(on);
        c.wanted = std::move(wanted);
        c.wait = std::move(wait);
        c.progress = std::move(progress);
    };
}

using Sched = apg::Scheduler<MockCtx>;
using Job = apg::BasicJob<MockCtx>;

int failures = 0;
#define CHECK(cond, what)                                      \
    do {                                                       \
        if (!(cond)) {                                         \
            std::printf("FAIL %s (line %d)\n", what, __LINE__); \
            failures++;                                        \
        }                                                      \
    } while (0)

uint64_t mix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

std::unique_ptr<Job> make_job(uint64_t input, size_t mem, double work, int reserve, double small,
                              std::atomic<int> *lane_violations) {
    auto j = std::make_unique<Job>();
    j->mem_bytes = mem;
    j->work = work;
    j->fn = [input, mem, work, reserve, small, lane_violations](MockCtx &c, Job &self) {
        if (c.index < reserve && work > small) (*lane_violations)++;
        // the arena grows to what the job needs
        size_t h = c.held.load();
        while (h < mem && !c.held.compare_exchange_weak(h, mem)) {
        }
        uint64_t v = input;
        const int spins = (int)(mix(input) % 2000);
        for (int i = 0; i < spins; i++) {
            v = mix(v);
            if (i % 500 == 499) c.yield_point((double)i / spins);
        }
        std::this_thread::sleep_for(std::chrono::microseconds(mix(input) % 300));
        c.yield_point();
        self.rounds.resize(8);
        for (int i = 0; i < 8; i++) self.rounds[i] = (uint8_t)(v >> (8 * i));
    };
    return j;
}

std::vector<uint8_t> expected(uint64_t input) {
    uint64_t v = input;
    const int spins = (int)(mix(input) % 2000);
    for (int i = 0; i < spins; i++) v = mix(v);
    std::vector<uint8_t> r(8);
    for (int i = 0; i < 8; i++) r[i] = (uint8_t)(v >> (8 * i));
    return r;
}

void run_mix(apg::Policy policy, int workers, int reserve, size_t budget, int submitters, int per_thread,
             double preempt = 0, apg::Preempt mode = apg::Preempt::kSize) {
    std::vector<std::unique_ptr<MockCtx>> ctxs;
    std::vector<MockCtx *> raw;
    for (int w = 0; w < workers; w++) {
        ctxs.push_back(std::make_unique<MockCtx>());
        ctxs.back()->index = w;
        raw.push_back(ctxs.back().get());
    }
    apg::WorkerHooks<MockCtx> hooks;
    hooks.held = [](const MockCtx &c) { return c.held.load(); };
    std::atomic<int> collected{0};
    hooks.collect = [&](MockCtx &) { collected++; };
    bind(hooks);
    const double small = 50;
    std::atomic<int> lane_violations{0};
    std::atomic<bool> over_budget{false}, stop_monitor{false};
    {
        Sched s(raw, budget, reserve, hooks);
        s.set_policy(policy, reserve, small);
        s.set_preemption(preempt, false, mode);
        std::thread monitor([&] {
            while (!stop_monitor) {
                const apg::EngineStats st = s.stats();
                if (st.running >= 2 && st.hbm_in_use > budget) over_budget = true;
                std::this_thread::yield();
            }
        });
        std::vector<std::thread> subs;
        std::atomic<int> wrong{0}, errors{0};
        for (int t = 0; t < submitters; t++) {
            subs.emplace_back([&, t] {
                std::mt19937_64 rng(1000 + t);
                std::vector<std::pair<uint64_t, uint64_t>> ids;
                for (int i = 0; i < per_thread; i++) {
                    const uint64_t input = rng();
                    const size_t mem = 10 + rng() % 51;  // 10..60 units
                    const double work = (double)(1 + rng() % 100);
                    ids.push_back({s.submit(make_job(input, mem, work, reserve, small, &lane_violations)), input});
                    if (i % 7 == 3) {  // collect some early, out of order
                        auto p = ids.back();
                        ids.pop_back();
                        auto j = s.wait(p.first);
                        if (j->status != 0) errors++;
                        if (j->rounds != expected(p.second)) wrong++;
                    }
                }
                for (auto &p : ids) {
                    auto j = s.wait(p.first);
                    if (j->status != 0) errors++;
                    if (j->rounds != expected(p.second)) wrong++;
                    // a second wait on a collected id is refused at once
                    bool threw = false;
                    try {
                        s.wait(p.first);
                    } catch (const apg::ApgError &e) {
                        threw = e.code == -2;
                    }
                    if (!threw) wrong++;
                    if (s.poll(p.first) != -1) wrong++;
                }
            });
        }
        for (auto &t : subs) t.join();
        stop_monitor = true;
        monitor.join();
        CHECK(wrong == 0, "every job's output is its own, and collected ids are refused");
        CHECK(errors == 0, "no job failed");
        const apg::EngineStats st = s.stats();
        CHECK(st.completed == (size_t)(submitters * per_thread), "every job completed");
        CHECK(st.queued == 0 && st.running == 0, "nothing left queued or running");
        CHECK(st.hbm_peak >= st.arena_bytes || st.hbm_peak > 0, "peak recorded");
        if (preempt == 0) CHECK(s.pauses() == 0, "nothing pauses with preemption off");
    }
    CHECK(!over_budget, "charged HBM within the budget whenever two or more jobs run");
    CHECK(lane_violations == 0, "fast-lane workers ran short jobs only");
    CHECK(collected == submitters * per_thread, "collect hook ran once per job");
    std::printf("ok mix policy=%d workers=%d reserve=%d budget=%zu preempt=%g mode=%d\n", (int)policy, workers,
                reserve, budget, preempt, (int)mode);
}

void unknown_ids() {
    MockCtx c;
    std::vector<MockCtx *> raw{&c};
    apg::WorkerHooks<MockCtx> hooks;
    hooks.held = [](const MockCtx &x) { return x.held.load(); };
    Sched s(raw, 100, 0, hooks);
    bool threw = false;
    try {
        s.wait(12345);
    } catch (const apg::ApgError &e) {
        threw = e.code == -2;
    }
    CHECK(threw, "wait on an id never issued is refused");
    CHECK(s.poll(12345) == -1, "poll on an id never issued is refused");
    std::atomic<int> lv{0};
    auto big = make_job(1, 1000, 1, 0, 0, &lv);
    const uint64_t id = s.submit(std::move(big));
    auto j = s.wait(id);
    CHECK(j->status == -2, "a job above the budget is refused at submission");
    std::printf("ok unknown ids\n");
}

void failing_worker() {
    std::vector<std::unique_ptr<MockCtx>> ctxs;
    std::vector<MockCtx *> raw;
    for (int w = 0; w < 3; w++) {
        ctxs.push_back(std::make_unique<MockCtx>());
        ctxs.back()->index = w;
        raw.push_back(ctxs.back().get());
    }
    apg::WorkerHooks<MockCtx> hooks;
    hooks.held = [](const MockCtx &x) { return x.held.load(); };
    hooks.start = [](int w, MockCtx &) {
        if (w == 0) throw apg::ApgError(-3, "device unavailable");
    };
    std::atomic<int> lv{0};
    Sched s(raw, 1000, 0, hooks);
    std::vector<std::pair<uint64_t, uint64_t>> ids;
    for (int i = 0; i < 60; i++) ids.push_back({s.submit(make_job(100 + i, 10, 1, 0, 0, &lv)), 100 + i});
    int failed = 0, good = 0;
    for (auto &p : ids) {
        auto j = s.wait(p.first);
        if (j->status == -3 && j->worker == 0) failed++;
        if (j->status == 0 && j->worker != 0 && j->rounds == expected(p.second)) good++;
    }
    CHECK(failed + good == 60, "jobs on the failed worker carry its error, the others their outputs");
    std::printf("ok failing worker (%d failed on worker 0, %d served)\n", failed, good);
}

void drain_on_destroy() {
    MockCtx c;
    std::vector<MockCtx *> raw{&c};
    apg::WorkerHooks<MockCtx> hooks;
    hooks.held = [](const MockCtx &x) { return x.held.load(); };
    std::atomic<int> ran{0};
    {
        Sched s(raw, 100, 0, hooks);
        for (int i = 0; i < 20; i++) {
            auto j = std::make_unique<Job>();
            j->mem_bytes = 5;
            j->fn = [&ran](MockCtx &, Job &) {
                std::this_thread::sleep_for(std::chrono::microseconds(200));
                ran++;
            };
            s.submit(std::move(j));
        }
    }
    CHECK(ran == 20, "queued jobs run before the scheduler is destroyed");
    std::printf("ok drain on destroy\n");
}

// A long job (work 1000) steps with a yield point per step; a short job (work
// 10) submitted while it runs must run with the long job paused: the long
// job's step count may move by at most one (the step in flight) between the
// short job's start and end.
void preemption_order() {
    std::vector<std::unique_ptr<MockCtx>> ctxs;
    std::vector<MockCtx *> raw;
    for (int w = 0; w < 3; w++) {
        ctxs.push_back(std::make_unique<MockCtx>());
        ctxs.back()->index = w;
        raw.push_back(ctxs.back().get());
    }
    apg::WorkerHooks<MockCtx> hooks;
    hooks.held = [](const MockCtx &x) { return x.held.load(); };
    bind(hooks);
    std::atomic<int> steps{0}, at_start{-1}, at_end{-1};
    std::atomic<bool> long_started{false};
    {
        Sched s(raw, 1000, 0, hooks);
        s.set_policy(apg::Policy::kShortest, 0, 0);
        s.set_preemption(4.0);
        auto lj = std::make_unique<Job>();
        lj->mem_bytes = 10;
        lj->work = 1000;
        lj->fn = [&](MockCtx &c, Job &) {
            long_started = true;
            for (int i = 0; i < 400; i++) {
                std::this_thread::sleep_for(std::chrono::microseconds(100));
                steps++;
                c.yield_point();
            }
        };
        const uint64_t lid = s.submit(std::move(lj));
        while (!long_started) std::this_thread::yield();
        std::this_thread::sleep_for(std::chrono::milliseconds(2));
        auto sj = std::make_unique<Job>();
        sj->mem_bytes = 10;
        sj->work = 10;
        sj->fn = [&](MockCtx &, Job &) {
            at_start = steps.load();
            std::this_thread::sleep_for(std::chrono::milliseconds(5));  // ~50 of the long job's steps
            at_end = steps.load();
        };
        const uint64_t sid = s.submit(std::move(sj));
        auto sr = s.wait(sid);
        auto lr = s.wait(lid);
        CHECK(sr->status == 0 && lr->status == 0, "both jobs completed");
        CHECK(steps == 400, "the long job finished all its steps after resuming");
        CHECK(at_start >= 0 && at_end - at_start <= 1, "the long job paused while the short one ran");
        CHECK(s.pauses() >= 1, "the long job paused");
        // a job of similar size does not pause it (ratio 4)
        const int before = (int)s.pauses();
        steps = 0;
        auto a = std::make_unique<Job>();
        a->mem_bytes = 10;
        a->work = 1000;
        a->fn = [&](MockCtx &c, Job &) {
            for (int i = 0; i < 100; i++) {
                std::this_thread::sleep_for(std::chrono::microseconds(100));
                steps++;
                c.yield_point();
            }
        };
        auto b = std::make_unique<Job>();
        b->mem_bytes = 10;
        b->work = 400;
        b->fn = [&](MockCtx &c, Job &) {
            for (int i = 0; i < 100; i++) {
                std::this_thread::sleep_for(std::chrono::microseconds(100));
                c.yield_point();
            }
        };
        const uint64_t ia = s.submit(std::move(a)), ib = s.submit(std::move(b));
        CHECK(s.wait(ia)->status == 0 && s.wait(ib)->status == 0, "similar jobs completed");
        CHECK((int)s.pauses() == before, "a job less than the ratio smaller causes no pause");
    }
    std::printf("ok preemption order (long job steps %d..%d while the short one ran)\n", at_start.load(),
                at_end.load());
}

// Urgency: a long job (isolated estimate 100 ms) that has run a while and a
// short job (1 ms) submitted now: the short one is more urgent at once (its
// remaining time over its isolated time), so the long job pauses; once the
// short one ends, the long one resumes. And a job that has waited long
// enough becomes the more urgent one: two jobs of equal estimates swap as
// their urgencies cross, and both finish.
void urgency_order() {
    std::vector<std::unique_ptr<MockCtx>> ctxs;
    std::vector<MockCtx *> raw;
    for (int w = 0; w < 3; w++) {
        ctxs.push_back(std::make_unique<MockCtx>());
        ctxs.back()->index = w;
        raw.push_back(ctxs.back().get());
    }
    apg::WorkerHooks<MockCtx> hooks;
    hooks.held = [](const MockCtx &x) { return x.held.load(); };
    bind(hooks);
    std::atomic<int> steps{0}, at_start{-1}, at_end{-1};
    std::atomic<bool> started{false};
    {
        Sched s(raw, 1000, 0, hooks);
        s.set_policy(apg::Policy::kShortest, 0, 0);
        s.set_preemption(1.0, false, apg::Preempt::kUrgency);
        auto lj = std::make_unique<Job>();
        lj->mem_bytes = 10;
        lj->work = 1000;
        lj->iso_ms = 100;
        lj->fn = [&](MockCtx &c, Job &) {
            started = true;
            for (int i = 0; i < 300; i++) {
                std::this_thread::sleep_for(std::chrono::microseconds(100));
                steps++;
                c.yield_point(i / 300.0);
            }
        };
        const uint64_t lid = s.submit(std::move(lj));
        while (!started) std::this_thread::yield();
        std::this_thread::sleep_for(std::chrono::milliseconds(2));
        auto sj = std::make_unique<Job>();
        sj->mem_bytes = 10;
        sj->work = 10;
        sj->iso_ms = 1;
        sj->fn = [&](MockCtx &, Job &) {
            at_start = steps.load();
            std::this_thread::sleep_for(std::chrono::milliseconds(5));
            at_end = steps.load();
        };
        const uint64_t sid = s.submit(std::move(sj));
        CHECK(s.wait(sid)->status == 0 && s.wait(lid)->status == 0, "both jobs completed (urgency)");
        CHECK(steps == 300, "the long job finished after resuming (urgency)");
        CHECK(at_start >= 0 && at_end - at_start <= 1, "the long job paused for the more urgent short one");
        // two equal jobs: whichever waits becomes the more urgent; both end
        std::atomic<int> a_steps{0}, b_steps{0};
        auto mk = [&](std::atomic<int> &ctr) {
            auto j = std::make_unique<Job>();
            j->mem_bytes = 10;
            j->work = 500;
            j->iso_ms = 10;
            j->fn = [&ctr](MockCtx &c, Job &) {
                for (int i = 0; i < 100; i++) {
                    std::this_thread::sleep_for(std::chrono::microseconds(100));
                    ctr++;
                    c.yield_point(i / 100.0);
                }
            };
            return j;
        };
        const uint64_t ia = s.submit(mk(a_steps)), ib = s.submit(mk(b_steps));
        CHECK(s.wait(ia)->status == 0 && s.wait(ib)->status == 0, "equal jobs completed (urgency)");
        CHECK(a_steps == 100 && b_steps == 100, "equal jobs ran every step");
    }
    std::printf("ok urgency order (long job steps %d..%d while the short one ran)\n", at_start.load(), at_end.load());
}

}  // namespace

namespace apg {
void set_last_error(const std::string &) {}
}  // namespace apg

int main() {
    unknown_ids();
    failing_worker();
    drain_on_destroy();
    // budgets below, at and above what the arenas reach (workers x 60 units)
    run_mix(apg::Policy::kFifo, 4, 0, 100, 3, 40);
    run_mix(apg::Policy::kShortest, 4, 0, 130, 3, 40);
    run_mix(apg::Policy::kShortest, 5, 2, 150, 4, 30);
    run_mix(apg::Policy::kFifo, 6, 1, 1000, 4, 30);
    run_mix(apg::Policy::kLongest, 4, 0, 130, 3, 40);
    run_mix(apg::Policy::kLongest, 5, 2, 150, 4, 30);
    preemption_order();
    run_mix(apg::Policy::kShortest, 5, 1, 1000, 4, 30, 1.0);
    run_mix(apg::Policy::kShortest, 6, 0, 150, 4, 30, 2.0);
    run_mix(apg::Policy::kFifo, 4, 1, 130, 3, 40, 4.0);
    run_mix(apg::Policy::kShortest, 6, 2, 1000, 4, 30, 1.0, apg::Preempt::kUrgency);
    run_mix(apg::Policy::kShortest, 5, 0, 150, 4, 30, 1.5, apg::Preempt::kUrgency);
    urgency_order();
    if (failures) {
        std::printf("%d checks failed\n", failures);
        return 1;
    }
    std::printf("all checks passed\n");
    return 0;
}
