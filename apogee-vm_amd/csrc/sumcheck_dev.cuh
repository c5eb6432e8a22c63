// Device helpers shared by the sumcheck round kernels (sumcheck.hip, gkr.hip):
// Fr loads and stores as 16-byte vectors, the suffix-eq factors of a round,
// and the grid-wide reduction of per-thread node sums.
#pragma once
#include "field.cuh"
#include "runtime.h"

namespace apg {
namespace {

constexpr int RTPB = 256;
constexpr int RWAVES = RTPB / 64;
constexpr int MAX_RBLOCKS = 2048;

// E_s(i) for the pairs of one round: lo[i mod 2^lo_bits] * hi[i >> lo_bits],
// or hi[i] when lo is null.
struct EqArgs {
    const Fe *lo;
    const Fe *hi;
    uint32_t lo_bits;
};

__device__ __forceinline__ Fe ldfe(const Fe *p, size_t i) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p + i);
    uint4 a = q[0], b = q[1];
    Fe r;
    r.l[0] = a.x; r.l[1] = a.y; r.l[2] = a.z; r.l[3] = a.w;
    r.l[4] = b.x; r.l[5] = b.y; r.l[6] = b.z; r.l[7] = b.w;
    return r;
}

__device__ __forceinline__ void stfe(Fe *p, size_t i, const Fe &v) {
    uint4 *q = reinterpret_cast<uint4 *>(p + i);
    q[0] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
    q[1] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
}

__device__ __forceinline__ Fe eq_factor(const EqArgs &q, size_t i) {
    if (!q.lo) return ldfe(q.hi, i);
    return f_mul<Fr>(ldfe(q.lo, i & ((size_t(1) << q.lo_bits) - 1)), ldfe(q.hi, i >> q.lo_bits));
}

__device__ __forceinline__ Fe wave_sum(Fe v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        Fe o;
#pragma unroll
        for (int q = 0; q < 8; q++) o.l[q] = __shfl_xor(v.l[q], off, 64);
        v = f_add<Fr>(v, o);
    }
    return v;
}

// The sums over the even and over the odd lanes of the wave: lane 0 ends
// with the even lanes' sum, lane 1 with the odd lanes'.
__device__ __forceinline__ Fe wave_sum_parity(Fe v) {
#pragma unroll
    for (int off = 32; off > 1; off >>= 1) {
        Fe o;
#pragma unroll
        for (int q = 0; q < 8; q++) o.l[q] = __shfl_xor(v.l[q], off, 64);
        v = f_add<Fr>(v, o);
    }
    return v;
}

// With `done` set (pinned host memory, like out), the block that writes out[]
// then sets *done = 1 behind a system-scope fence, so the host can wait for
// the node sums without synchronising the stream.
__device__ __forceinline__ void signal_done(unsigned *done) {
    if (!done) return;
    __syncthreads();  // every out[] store of the block issued
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The block's sums of nn values, given each wave's sums in lds[wave][x]
// (written before a barrier): into out[] (then the completion flag) when the
// grid is one block, else into the block's row of partials (row stride NN),
// which k_finish_nodes adds up in a launch of its own after the kernel. The
// kernel boundary orders every row's stores before that pass's loads, so no
// block fences its stores or counts itself done: a device-scope fence writes
// back the XCD's L2 and waits, and with many blocks those fences queue (at 4
// blocks per CU they made round 0 of 2^22 x 2 tables 300 us against 202).
template <int NN>
__device__ void finish_nodes(Fe (&lds)[RWAVES][NN], int nn, Fe *partials, Fe *out, unsigned *done) {
    const int tid = threadIdx.x;
    const bool single = gridDim.x == 1;
    if (tid < nn) {
        Fe s = lds[0][tid];
        for (int q = 1; q < RWAVES; q++) s = f_add<Fr>(s, lds[q][tid]);
        if (single)
            out[tid] = s;
        else
            stfe(partials, (size_t)blockIdx.x * NN + tid, s);
    }
    if (single) signal_done(done);
}

// Sum of the first nn of NN per-thread accumulators over the block into
// out[] or the block's row of partials (finish_nodes).
template <int NN>
__device__ void reduce_nodes(Fe (&acc)[NN], int nn, Fe *partials, Fe *out, unsigned *done = nullptr) {
    __shared__ Fe lds[RWAVES][NN];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int x = 0; x < NN; x++) {
        if (x < nn) {
            const Fe v = wave_sum(acc[x]);
            if (lane == 0) lds[w][x] = v;
        }
    }
    __syncthreads();
    finish_nodes<NN>(lds, nn, partials, out, done);
}

// The second pass of a multi-block reduction: the rows of `blocks` blocks'
// partials (row stride NN) added into out[0..nn), then the completion flag.
// One block; all outputs at once, output x by the tpx lanes tid / tpx == x
// (an aligned group inside one wave), each over every tpx-th row.
template <int NN>
__global__ void __launch_bounds__(RTPB) k_finish_nodes(const Fe *__restrict__ partials, unsigned blocks, int nn,
                                                     Fe *out, unsigned *done) {
    raise_wave_priority();
    const int tid = threadIdx.x;
    int tpx = 64;
    while (tpx * nn > RTPB) tpx >>= 1;
    const int x = tid / tpx, sub = tid % tpx;
    Fe a = f_zero<Fr>();
    if (x < nn)
        for (unsigned b = sub; b < blocks; b += tpx) a = f_add<Fr>(a, ldfe(partials, (size_t)b * NN + x));
    for (int off = tpx / 2; off > 0; off >>= 1) {
        Fe o;
#pragma unroll
        for (int q = 0; q < 8; q++) o.l[q] = __shfl_xor(a.l[q], off, 64);
        a = f_add<Fr>(a, o);
    }
    if (x < nn && sub == 0) out[x] = a;
    signal_done(done);
}

// After a round kernel of `blocks` blocks whose rows of NN partials are in
// `partials`: the pass that adds them (a one-block grid wrote out[] itself).
template <int NN>
inline void launch_finish(hipStream_t st, unsigned blocks, int nn, const Fe *partials, Fe *out, unsigned *done) {
    if (blocks <= 1) return;
    hipLaunchKernelGGL(k_finish_nodes<NN>, dim3(1), dim3(RTPB), 0, st, partials, blocks, nn, out, done);
}

// The suffix-eq tables of one proof, as products over the variables they
// cover (crates/poly/src/lib.rs:233-248: eq(r, x) = prod_j (r_j x_j +
// (1 - r_j)(1 - x_j)), variable j = bit j):
//   A_s (1 <= s < h0): 2^(h0-s) entries over variables s..h0-1, at A + (s-1) 2^(h0-1);
//   B_s (h0 <= s <= n): 2^(n-s) entries over variables s..n-1, at B + (s-h0) 2^(n-h0).
__global__ void k_eq_suffix(const Fe *__restrict__ r, int n, int h0, Fe *__restrict__ A,
                            Fe *__restrict__ B) {
    raise_wave_priority();
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t astride = size_t(1) << (h0 - 1), bstride = size_t(1) << (n - h0);
    const size_t asz = (size_t)(h0 - 1) * astride, bsz = (size_t)(n - h0 + 1) * bstride;
    int s, top;
    size_t x;
    Fe *dst;
    if (tid < asz) {
        s = 1 + (int)(tid / astride);
        x = tid % astride;
        top = h0;
        dst = A + tid;
    } else if (tid < asz + bsz) {
        const size_t u = tid - asz;
        s = h0 + (int)(u / bstride);
        x = u % bstride;
        top = n;
        dst = B + u;
    } else {
        return;
    }
    if (x >= (size_t(1) << (top - s))) return;
    const Fe one = f_one<Fr>();
    Fe acc = one;
    for (int l = s; l < top; l++) {
        const Fe rl = ldfe(r, l);
        acc = f_mul<Fr>(acc, ((x >> (l - s)) & 1) ? rl : f_sub<Fr>(one, rl));
    }
    stfe(dst, 0, acc);
}

}  // namespace
}  // namespace apg
