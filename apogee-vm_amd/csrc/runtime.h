// Library runtime: per-context HIP stream, a grow-only device arena of named
// scratch buffers, pinned host staging, and the error channel of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "error.h"

namespace apg {

#define HIP_CHECK(x)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            throw ::apg::ApgError(-3, std::string("HIP error at " __FILE__ ":") +           \
                                          std::to_string(__LINE__) + ": " + #x + ": " +      \
                                          hipGetErrorString(e_));                            \
    } while (0)

// A device buffer that only grows. Reused across calls on one context. A
// buffer it outgrows is retired, not freed: hipFree waits for the whole
// device, which would stall this context behind every other stream's work
// (other engine workers' long kernels). Retired buffers are freed when the
// context is released.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    // held: the owning context's count of device bytes (live and retired),
    // which the shard engine charges against its HBM budget
    void *get(size_t bytes, std::vector<void *> &retired, std::atomic<size_t> &held) {
        if (bytes > cap) {
            if (ptr) retired.push_back(ptr);
            ptr = nullptr;
            size_t c = bytes + bytes / 4 + 256;
            HIP_CHECK(hipMalloc(&ptr, c));
            cap = c;
            held += c;
        }
        return ptr;
    }
    ~DevBuf() {
        if (ptr) (void)hipFree(ptr);
    }
};

struct HostBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    void *get(size_t bytes, std::vector<void *> &retired, unsigned flags = hipHostMallocDefault) {
        if (bytes > cap) {
            if (ptr) retired.push_back(ptr);
            ptr = nullptr;
            size_t c = bytes + bytes / 4 + 256;
            HIP_CHECK(hipHostMalloc(&ptr, c, flags));
            cap = c;
        }
        return ptr;
    }
    ~HostBuf() {
        if (ptr) (void)hipHostFree(ptr);
    }
};

// Per-context statistics: the host wall clock of every commit MSM and
// sumcheck, always; kernel timings recorded with HIP events on the context's
// stream only when profiling is on (apg_ctx_set_profiling), since the events
// and their reads cost a few microseconds per launch.
struct Stats {
    double accumulate_ms = 0;        // k_accumulate (MSM bucket accumulation), profiling;
                                     // not a preemptible job's chunked launches
    uint64_t accumulate_launches = 0;
    uint64_t accumulate_entries = 0;  // nonzero signed digits processed
    double msm_ms = 0;                // whole msm_run, host wall clock
    uint64_t msm_calls = 0;
    uint64_t msm_points = 0;          // scalars (points x columns)
    double sumcheck_kernel_ms = 0;    // round kernels (fold + evaluate), profiling
    uint64_t sumcheck_launches = 0;
    uint64_t sumcheck_bytes = 0;      // algorithmic HBM bytes of those launches
    double sumcheck_ms = 0;           // whole sumcheck, host wall clock

    Stats &operator+=(const Stats &o) {
        accumulate_ms += o.accumulate_ms;
        accumulate_launches += o.accumulate_launches;
        accumulate_entries += o.accumulate_entries;
        msm_ms += o.msm_ms;
        msm_calls += o.msm_calls;
        msm_points += o.msm_points;
        sumcheck_kernel_ms += o.sumcheck_kernel_ms;
        sumcheck_launches += o.sumcheck_launches;
        sumcheck_bytes += o.sumcheck_bytes;
        sumcheck_ms += o.sumcheck_ms;
        return *this;
    }
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // fork/join partner of `stream` within one call
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::atomic<bool> profiling{false};  // set from other threads (engine)
    Stats stats;
    std::mutex mu;  // one call at a time per context
    std::unordered_map<std::string, DevBuf> dev;
    std::unordered_map<std::string, HostBuf> host;
    std::vector<void *> retired_dev, retired_host;  // outgrown buffers (see DevBuf)
    // device bytes this context holds (arena buffers, live and retired);
    // written by the thread that owns the context, read by others
    std::atomic<size_t> held_bytes{0};
    int cu_count = 256;
    bool p2_rc_ready = false;  // the Poseidon2 constants sit in the arena ("p2.rc")
    // this job's bucket accumulation runs below other kernels' waves (a long
    // job of a served stream, set per job by the engine)
    bool low_acc_prio = false;
    // Phase-boundary preemption, bound by the shard engine to its scheduler
    // (jobqueue.h); empty for plain contexts, where yield_point does nothing.
    std::function<bool()> yield_on, yield_wanted;
    std::function<void()> yield_wait;
    std::function<void(double)> yield_progress;
    double prog_lo = 0, prog_hi = 1;  // the running phase's share of the job (ProgressSpan)

    bool preemptible() const { return yield_on && yield_on(); }

    // A phase boundary of a job: the host waits for the stream to drain (so a
    // job that starts now waits for at most the phase in flight), reports the
    // job's progress (`done`: the fraction of the current phase done, or < 0
    // for none), then pauses while the scheduler asks it to.
    void yield_point(double done = -1) {
        if (!preemptible()) return;
        sync();
        if (done >= 0 && yield_progress) yield_progress(prog_lo + (prog_hi - prog_lo) * done);
        if (yield_wanted()) yield_wait();
    }

    template <class T>
    T *d(const char *name, size_t count) {
        return static_cast<T *>(dev[name].get(count * sizeof(T), retired_dev, held_bytes));
    }
    // A buffer that is zero when first handed out; kernels that use it as a
    // counter leave it zero again when they finish.
    template <class T>
    T *zeroed(const char *name, size_t count) {
        DevBuf &b = dev[name];
        const bool fresh = b.cap < count * sizeof(T);
        T *p = static_cast<T *>(b.get(count * sizeof(T), retired_dev, held_bytes));
        if (fresh) {
            HIP_CHECK(hipMemsetAsync(p, 0, b.cap, stream));
            sync();
        }
        return p;
    }
    template <class T>
    T *h(const char *name, size_t count) {
        return static_cast<T *>(host[name].get(count * sizeof(T), retired_host));
    }
    // pinned host memory that device stores reach without a stream
    // synchronisation (coherent mapping): completion flags and what they guard
    template <class T>
    T *hc(const char *name, size_t count) {
        return static_cast<T *>(host[name].get(count * sizeof(T), retired_host, hipHostMallocCoherent));
    }
    void sync() { HIP_CHECK(hipStreamSynchronize(stream)); }
};

// A phase of a job as the fraction [a, b) of the enclosing phase's share,
// for the progress its yield points report; restores the share on exit.
struct ProgressSpan {
    Context &c;
    double lo, hi;
    ProgressSpan(Context &ctx, double a, double b) : c(ctx), lo(ctx.prog_lo), hi(ctx.prog_hi) {
        c.prog_lo = lo + (hi - lo) * a;
        c.prog_hi = lo + (hi - lo) * b;
    }
    ~ProgressSpan() {
        c.prog_lo = lo;
        c.prog_hi = hi;
    }
};

// Streams (default priority; non-blocking, or restricted to the CUs set in
// cu_mask, 32 CUs per word) and fork/join events of a context. With
// side_stream = false the side work runs in order on the one stream: engine
// workers use one stream each, because the device time-slices a process's
// hardware queues once more than about five are busy, and a slice is ~1 ms.
inline void context_init(Context &c, int device, const std::vector<uint32_t> &cu_mask = {},
                         bool side_stream = true) {
    c.device = device;
    HIP_CHECK(hipSetDevice(device));
    if (cu_mask.empty())
        HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    else
        HIP_CHECK(hipExtStreamCreateWithCUMask(&c.stream, (uint32_t)cu_mask.size(), cu_mask.data()));
    if (!side_stream)
        c.side = c.stream;
    else if (cu_mask.empty())
        HIP_CHECK(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    else
        HIP_CHECK(hipExtStreamCreateWithCUMask(&c.side, (uint32_t)cu_mask.size(), cu_mask.data()));
    HIP_CHECK(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, device));
    c.cu_count = prop.multiProcessorCount;
}

inline void context_release(Context &c) {
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    if (c.side) (void)hipStreamSynchronize(c.side);
    c.dev.clear();
    c.host.clear();
    for (void *p : c.retired_dev) (void)hipFree(p);
    for (void *p : c.retired_host) (void)hipHostFree(p);
    c.retired_dev.clear();
    c.retired_host.clear();
    c.held_bytes = 0;
    if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
    if (c.ev_join) (void)hipEventDestroy(c.ev_join);
    if (c.side && c.side != c.stream) (void)hipStreamDestroy(c.side);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.stream = c.side = nullptr;
    c.ev_fork = c.ev_join = nullptr;
}

// Wave priority of the shard path's kernels: 1, above the bucket
// accumulation of a long job in a served stream (0), so that when both have
// waves ready on a SIMD a short job's kernels issue first instead of getting
// the long job's issue share.
__device__ __forceinline__ void raise_wave_priority() { __builtin_amdgcn_s_setprio(1); }
// Priority 2: a few waves of dependent additions on a phase's critical path
// (the MSM's heavy-bucket trees) ahead of the wide kernel beside them.
__device__ __forceinline__ void raise_wave_priority_critical() { __builtin_amdgcn_s_setprio(2); }

inline unsigned grid_for(size_t n, unsigned block) {
    return (unsigned)((n + block - 1) / block);
}

}  // namespace apg
