// Multi-scalar multiplication over BN254 G1 for gfx950.
//
// Replaces curve::msm::{msm, msm_small_u32} (crates/curve/src/msm.rs:48-100).
// The value is the group element sum_i s_i B_i, so any correct bucket schedule
// returns the reference's exact affine bytes; this one is built for the MI355X:
//
//  1. digits  : one thread per scalar converts Montgomery -> canonical, recodes
//               into signed c-bit digits exactly as msm.rs:180-224 does (same
//               carry rule) and emits one (bucket key, table index | sign)
//               pair per nonzero digit, packed densely in scalar order (a
//               count pass, a scan of the blocks' counts, a write pass).
//  2. sort    : an LSD radix sort of the pairs (hipCUB): by the bucket bits
//               alone when every window has its table copy (stable, so the
//               columns stay apart within a bucket), else by the whole key;
//               then a binary search per key for its range.
//  3. accumulate: the sorted entries are cut into equal slices, one thread
//               per slice summing its points with mixed XYZZ adds, so the
//               lanes of a wave finish together whatever the digit skew (the
//               witness-like mix piles 10% of all points into one bucket). A
//               bucket inside one slice is written directly; one that crosses
//               slices leaves a head and a tail partial sum per slice. The
//               gather of 64-byte affine bases is the only large HBM traffic.
//  4. buckets : crossing buckets combine their partials: a thread per light
//               bucket; heavy buckets cut into pieces over the whole grid
//               (block slices + LDS tree, then the few piece sums), on a
//               forked stream.
//  5. reduce  : sum_b (b+1) S_b per group by a two-level split
//               (b = 256 r + t: 256 * sum_r r Z_r + sum_t t U_t + sum_r Z_r):
//               row and column sums by lane groups (serial slices, then wave
//               shuffles), then an LDS suffix scan per 256-item problem:
//               ~30 dependent adds instead of the 2 * 2^(c-1) serial running
//               sum of msm.rs:304-310.
//  6. host    : groups combine with 2^(c P) doublings (none when the table holds
//               every window shifted), one batched inversion gives affine.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cstdio>

#include "curve29.cuh"
#include "msm.h"

namespace apg {

namespace {

constexpr uint32_t SIGN = 0x80000000u;
constexpr int TPB = 256;

__device__ __forceinline__ uint32_t window_of32(const uint32_t l[8], int off, int c) {
    const int limb = off >> 5, sh = off & 31;
    uint32_t v = l[limb] >> sh;
    if (sh && limb + 1 < 8) v |= l[limb + 1] << (32 - sh);
    return v & ((1u << c) - 1u);
}

__device__ __forceinline__ void load_scalar(const Fe *fr, const uint32_t *u32, size_t idx,
                                            uint32_t l[8]) {
    if (fr) {
        Fe c = f_from_mont<Fr>(fr[idx]);
#pragma unroll
        for (int i = 0; i < 8; i++) l[i] = c.l[i];
    } else {
        l[0] = u32[idx];
#pragma unroll
        for (int i = 1; i < 8; i++) l[i] = 0;
    }
}

// Signed-digit recode of one scalar (crates/curve/src/msm.rs:180-198): digit
// j in [-2^(c-1), 2^(c-1)], carry 1 out of any window whose value reaches
// 2^(c-1), the top window swallowing the last carry. Calls f(copy, group,
// |d|, d < 0) for every nonzero digit, where window j reads table copy
// j mod P of bucket group j / P (kept as counters: no division per digit).
// CT > 0 fixes c and the window count at compile time, so the loop unrolls
// and every window is read at a constant bit offset.
template <int CT, int WT, class F>
__device__ __forceinline__ void recode_digits(const uint32_t l[8], int c_rt, int windows_rt, int P, F &&f) {
    const int c = CT ? CT : c_rt;
    const int windows = CT ? WT : windows_rt;
    const uint32_t half = 1u << (c - 1);
    uint32_t carry = 0;
    int copy = 0, group = 0;
#pragma unroll
    for (int j = 0; j < windows; j++) {
        const uint32_t coef = window_of32(l, j * c, c) + carry;
        uint32_t mag;
        bool neg = false;
        if (j == windows - 1) {
            mag = coef;  // coef <= half here
        } else if (coef >= half) {
            mag = (1u << c) - coef;  // digit coef - 2^c; 0 when coef == 2^c
            neg = mag != 0;
            carry = 1;
        } else {
            mag = coef;
            carry = 0;
        }
        if (mag) f(copy, group, mag, neg);
        if (++copy == P) {
            copy = 0;
            group++;
        }
    }
}

// Emits one (bucket key, table index | sign) pair per nonzero digit. Each
// block packs its pairs contiguously at its offset from k_emit_count's scan,
// so the output is dense (zero digits, 20-70% of all digits for witness-like
// scalars, never reach the sort) and in scalar order. (One atomic per block
// on a shared counter instead serialised the ~32K blocks of a 2^20 x 8 shard:
// 387 us against 157 us without it.) With stage != 0 (windows <=
// STAGE_WINDOWS) the block first gathers its pairs in LDS, keys then values,
// and writes them out with consecutive lanes on consecutive words; otherwise
// each thread writes its own run.
constexpr int STAGE_WINDOWS = 31;  // 256 x 31 x 8 B + the scan: within 64 KiB of LDS

// Digits per block of scalars (the first pass of the emission): with the
// blocks' exclusive scan as output offsets, the pairs land densely in scalar
// order (column-major) with no atomics.
template <int CT, int WT>
__global__ void __launch_bounds__(TPB) k_emit_count(const Fe *__restrict__ fr, const uint32_t *__restrict__ u32,
                                                  uint32_t n, uint32_t K, int c, int windows, int P,
                                                  uint32_t *__restrict__ counts) {
    __shared__ uint32_t wsum[TPB / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + tid;
    uint32_t cnt = 0;
    if (idx < (size_t)n * K) {
        uint32_t l[8];
        load_scalar(fr, u32, idx, l);
        recode_digits<CT, WT>(l, c, windows, P, [&](int, int, uint32_t, bool) { cnt++; });
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if (lane == 0) wsum[wave] = cnt;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < TPB / 64; w++) t += wsum[w];
        counts[blockIdx.x] = t;
    }
}

template <int CT, int WT>
__global__ void __launch_bounds__(TPB) k_emit(const Fe *__restrict__ fr,
                                            const uint32_t *__restrict__ u32, uint32_t n,
                                            uint32_t K, int c, int windows, int P, int G,
                                            uint32_t B, uint32_t tab_stride, int stage,
                                            const uint32_t *__restrict__ offsets,
                                            uint32_t *__restrict__ keys,
                                            uint32_t *__restrict__ vals,
                                            uint32_t *__restrict__ total_out) {
    extern __shared__ uint32_t staged[];  // [TPB * windows] keys, then as many values
    __shared__ uint32_t wsum[TPB / 64];
    __shared__ uint32_t base;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the scan's last element is the pair count (to pinned host memory)
    if (blockIdx.x == 0 && tid == 0) *total_out = offsets[gridDim.x];
    const size_t idx = (size_t)blockIdx.x * blockDim.x + tid;
    const bool live = idx < (size_t)n * K;
    uint32_t l[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (live) load_scalar(fr, u32, idx, l);
    uint32_t cnt = 0;
    if (live) recode_digits<CT, WT>(l, c, windows, P, [&](int, int, uint32_t, bool) { cnt++; });
    // inclusive scan of the counts: within each wave by shuffles, then across
    // the block's waves (one barrier)
    uint32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, 64);
        if ((int)lane >= off) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t btotal = 0;
#pragma unroll
    for (int w = 0; w < TPB / 64; w++) {
        if (w < (int)wave) incl += wsum[w];
        btotal += wsum[w];
    }
    if (tid == 0) base = offsets[blockIdx.x];
    const uint32_t col = (uint32_t)(idx / n);
    const uint32_t i = (uint32_t)(idx - (size_t)col * n);
    const uint32_t key0 = col * (uint32_t)G;
    if (stage) {
        const uint32_t cap = TPB * (uint32_t)windows;
        uint32_t pos = incl - cnt;
        if (live && cnt)
            recode_digits<CT, WT>(l, c, windows, P, [&](int copy, int group, uint32_t mag, bool neg) {
                staged[pos] = (key0 + (uint32_t)group) * B + (mag - 1);
                staged[cap + pos] = ((uint32_t)copy * tab_stride + i) | (neg ? SIGN : 0u);
                pos++;
            });
        __syncthreads();
        const uint32_t b0 = base;
        for (uint32_t e = tid; e < btotal; e += TPB) {
            keys[b0 + e] = staged[e];
            vals[b0 + e] = staged[cap + e];
        }
        return;
    }
    __syncthreads();
    if (!live || !cnt) return;
    uint32_t pos = base + incl - cnt;
    recode_digits<CT, WT>(l, c, windows, P, [&](int copy, int group, uint32_t mag, bool neg) {
        keys[pos] = (key0 + (uint32_t)group) * B + (mag - 1);
        vals[pos] = ((uint32_t)copy * tab_stride + i) | (neg ? SIGN : 0u);
        pos++;
    });
}

// off[2k], off[2k + 1] = the range of key k in the sorted key array, which
// holds each key's entries contiguously (sorted fully, or by the bucket bits
// alone with the columns in emission order within a bucket). One thread per
// entry marks where the key changes; keys without entries keep the (0, 0)
// the array is cleared to.
__global__ void k_bounds(const uint32_t *__restrict__ skeys, uint32_t total, uint32_t *__restrict__ off) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t k = skeys[i];
    if (i == 0 || skeys[i - 1] != k) {
        off[2 * (size_t)k] = i;
        if (i > 0) off[2 * (size_t)skeys[i - 1] + 1] = i;
    }
    if (i == total - 1) off[2 * (size_t)k + 1] = total;
}

__device__ __forceinline__ Affine load_affine(const Affine *__restrict__ p, uint32_t idx) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p + idx);
    uint4 a = q[0], b = q[1], c = q[2], d = q[3];
    Affine r;
    r.x.l[0] = a.x; r.x.l[1] = a.y; r.x.l[2] = a.z; r.x.l[3] = a.w;
    r.x.l[4] = b.x; r.x.l[5] = b.y; r.x.l[6] = b.z; r.x.l[7] = b.w;
    r.y.l[0] = c.x; r.y.l[1] = c.y; r.y.l[2] = c.z; r.y.l[3] = c.w;
    r.y.l[4] = d.x; r.y.l[5] = d.y; r.y.l[6] = d.z; r.y.l[7] = d.w;
    return r;
}

// Accumulation over equal slices: thread t sums the sorted entries
// [t L, min((t + 1) L, total)), so every lane of a wave runs the same number
// of additions. A bucket that lies inside one slice is written to buckets[]
// directly. A bucket that crosses slice boundaries leaves partial sums:
// part[2t] (head) is the slice's first run when that bucket started in an
// earlier slice (the whole slice when it also continues past it), and
// part[2t + 1] (tail) its last run when that bucket continues into the next
// slice; unused slots hold the identity. Bucket k's partials are then the
// contiguous part[2 t0 + 1 .. 2 t1], t0 and t1 the slices of its first and last
// entries.
__global__ void __launch_bounds__(TPB) k_accumulate(const uint32_t *__restrict__ skeys,
                                                  const uint32_t *__restrict__ entries, uint32_t total,
                                                  uint32_t L, const Affine *__restrict__ bases,
                                                  Xyzz29 *__restrict__ buckets, Xyzz29 *__restrict__ part) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if ((size_t)t * L >= total) return;
    const uint32_t lo = t * L, hi = lo + min(total - lo, L);
    const bool head_open = lo > 0 && skeys[lo - 1] == skeys[lo];
    const bool tail_open = hi < total && skeys[hi] == skeys[hi - 1];
    uint32_t cur = skeys[lo];
    bool in_head = head_open, head_done = false;
    // lazy 29-bit Fq (curve29.cuh): no carry chains in the multiplies; sums
    // are stored in that form
    Xyzz29 acc{};
    bool id = true;
    // keys and entries arrive four at a time as 16-byte vectors (L is a
    // multiple of 4, so lo is 16-byte aligned; the arrays carry 4 words of
    // slack past total): one line fetch serves four iterations
    uint4 kv = make_uint4(0, 0, 0, 0), ev = kv;
    for (uint32_t e = lo; e < hi; e++) {
        const uint32_t q = (e - lo) & 3;
        if (q == 0) {
            kv = *reinterpret_cast<const uint4 *>(skeys + e);
            ev = *reinterpret_cast<const uint4 *>(entries + e);
        }
        const uint32_t key = q == 0 ? kv.x : q == 1 ? kv.y : q == 2 ? kv.z : kv.w;
        const uint32_t ent = q == 0 ? ev.x : q == 1 ? ev.y : q == 2 ? ev.z : ev.w;
        if (key != cur) {
            if (in_head) {
                st29_acc(part + 2 * (size_t)t, acc, id);
                head_done = true;
                in_head = false;
            } else {
                st29_acc(buckets + cur, acc, id);
            }
            cur = key;
            id = true;
        }
        xyzz29_accumulate(acc, id, load_affine(bases, ent & ~SIGN), (ent & SIGN) != 0);
    }
    bool tail_done = false;
    if (in_head) {
        st29_acc(part + 2 * (size_t)t, acc, id);  // the whole slice belongs to one crossing bucket
        head_done = true;
    } else if (tail_open) {
        st29_acc(part + 2 * (size_t)t + 1, acc, id);
        tail_done = true;
    } else {
        st29_acc(buckets + cur, acc, id);
    }
    if (!head_done) st29(part + 2 * (size_t)t, xyzz29_identity());
    if (!tail_done) st29(part + 2 * (size_t)t + 1, xyzz29_identity());
}

// Partial sums of bucket k in part[]: count 0 when it is empty or lies in
// one slice (then k_accumulate wrote it, or it stays empty).
struct Span {
    uint32_t first, count;
};

__device__ __forceinline__ Span merge_span(const uint32_t *__restrict__ off, uint32_t k, uint32_t L) {
    const uint32_t a = off[2 * (size_t)k], b = off[2 * (size_t)k + 1];
    if (a == b) return {0, 0};
    const uint32_t t0 = a / L, t1 = (b - 1) / L;
    if (t0 == t1) return {0, 0};
    return {2 * t0 + 1, 2 * (t1 - t0)};
}

constexpr uint32_t LIGHT_ITEMS = 32;  // more partials than this: a heavy bucket
constexpr size_t ACC_ROUNDS = 4, MIN_SLICE = 16, MAX_SLICE = 4096;

// One thread per bucket: empty buckets become the identity, buckets with at
// most LIGHT_ITEMS partials are summed here, heavier ones by the pieces below.
__global__ void __launch_bounds__(TPB, 2) k_bucket_light(const uint32_t *__restrict__ off, uint32_t nkeys,
                                                       uint32_t L, const Xyzz29 *__restrict__ part,
                                                       Xyzz29 *__restrict__ buckets) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    if (off[2 * (size_t)k] == off[2 * (size_t)k + 1]) {
        st29(buckets + k, xyzz29_identity());
        return;
    }
    const Span sp = merge_span(off, k, L);
    if (sp.count == 0 || sp.count > LIGHT_ITEMS) return;
    Xyzz29 acc = xyzz29_identity();
    for (uint32_t s = 0; s < sp.count; s++) acc = xyzz29_add(acc, ld29(part + sp.first + s));
    st29(buckets + k, acc);
}

// Point exchange between lanes of one wave.
__device__ __forceinline__ Xyzz29 shfl_xor_xyzz(const Xyzz29 &p, int mask) {
    Xyzz29 r;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        r.x.v[q] = __shfl_xor(p.x.v[q], mask, 64);
        r.y.v[q] = __shfl_xor(p.y.v[q], mask, 64);
        r.zz.v[q] = __shfl_xor(p.zz.v[q], mask, 64);
        r.zzz.v[q] = __shfl_xor(p.zzz.v[q], mask, 64);
    }
    return r;
}

// Heavy buckets (more than LIGHT_ITEMS partials; the witness-like mix puts
// every scalar equal to 1 into one bucket per column, spanning many slices)
// are cut into pieces of PIECE partials spread over the whole grid: each piece
// is one block's serial slices plus a shuffle/LDS tree, and a last pass adds
// each bucket's few piece sums. k_classify allocates the pieces as it finds
// the buckets.
constexpr uint32_t PIECE = 4 * TPB;

struct HeavyPlan {
    uint32_t nheavy, npieces;
};

// heavy[h] = (bucket, first piece, pieces) of heavy bucket h; piece_of[p] =
// (heavy index, piece within the bucket).
__global__ void k_classify(const uint32_t *__restrict__ off, uint32_t nkeys, uint32_t L, HeavyPlan *plan,
                           uint4 *__restrict__ heavy, uint2 *__restrict__ piece_of) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const uint32_t q = merge_span(off, k, L).count;
    if (q <= LIGHT_ITEMS) return;
    const uint32_t np = (q + PIECE - 1) / PIECE;
    const uint32_t h = atomicAdd(&plan->nheavy, 1u);
    const uint32_t p0 = atomicAdd(&plan->npieces, np);
    heavy[h] = make_uint4(k, p0, np, 0);
    for (uint32_t j = 0; j < np; j++) piece_of[p0 + j] = make_uint2(h, j);
}

// Sum of the 256 lanes' points of a block (all lanes call it).
__device__ __forceinline__ Xyzz29 block_sum256(Xyzz29 acc, Xyzz29 *lds) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) acc = xyzz29_add(acc, shfl_xor_xyzz(acc, m));
    const uint32_t tid = threadIdx.x;
    if ((tid & 63) == 0) lds[tid >> 6] = acc;
    __syncthreads();
    Xyzz29 r = lds[0];
    for (int w = 1; w < TPB / 64; w++) r = xyzz29_add(r, lds[w]);
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(TPB, 2) k_heavy_pieces(const uint32_t *__restrict__ off, uint32_t L,
                                                       const Xyzz29 *__restrict__ part,
                                                       const HeavyPlan *plan,
                                                       const uint4 *__restrict__ heavy,
                                                       const uint2 *__restrict__ piece_of,
                                                       Xyzz29 *__restrict__ piece_sums) {
    __shared__ Xyzz29 lds[TPB / 64];
    const uint32_t np = plan->npieces, tid = threadIdx.x;
    for (uint32_t pc = blockIdx.x; pc < np; pc += gridDim.x) {
        const uint2 po = piece_of[pc];
        const Span sp = merge_span(off, heavy[po.x].x, L);
        const uint32_t a = sp.first + po.y * PIECE;
        const uint32_t end = min(sp.first + sp.count, a + PIECE);
        Xyzz29 acc = xyzz29_identity();
        for (uint32_t c = a + tid; c < end; c += TPB) acc = xyzz29_add(acc, ld29(part + c));
        const Xyzz29 r = block_sum256(acc, lds);
        if (tid == 0) st29(piece_sums + pc, r);
    }
}

__global__ void __launch_bounds__(TPB, 2) k_heavy_final(const HeavyPlan *plan,
                                                      const uint4 *__restrict__ heavy,
                                                      const Xyzz29 *__restrict__ piece_sums,
                                                      Xyzz29 *__restrict__ buckets) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= plan->nheavy) return;
    const uint4 e = heavy[h];
    Xyzz29 acc = xyzz29_identity();
    for (uint32_t j = 0; j < e.z; j++) acc = xyzz29_add(acc, ld29(piece_sums + e.y + j));
    st29(buckets + e.x, acc);
}

constexpr int RGS = 32;  // threads per row / column sum

// Level 1 of the bucket reduction for B = 256 R buckets per problem, read as
// an R x 256 matrix S[r][t] = S[256 r + t]: row sums Z_r and column sums U_t.
// RGS consecutive lanes share one output: each sums a strided slice of it
// serially (full waves), then a shuffle tree across the RGS lanes.
__global__ void __launch_bounds__(TPB, 2) k_reduce_rows(const Xyzz29 *__restrict__ S, uint32_t R,
                                                      uint32_t nprob, Xyzz29 *__restrict__ Z,
                                                      Xyzz29 *__restrict__ U) {
    const size_t gt = (size_t)blockIdx.x * TPB + threadIdx.x;
    const uint32_t per = R + 256;
    const size_t o = gt / RGS;
    const uint32_t j = (uint32_t)(gt % RGS);
    if (o >= (size_t)nprob * per) return;  // whole groups only: RGS divides TPB
    const uint32_t p = (uint32_t)(o / per), w = (uint32_t)(o % per);
    const Xyzz29 *base = S + (size_t)p * R * 256;
    Xyzz29 acc = xyzz29_identity();
    if (w < R) {
        for (uint32_t t = j; t < 256; t += RGS) acc = xyzz29_add(acc, ld29(base + (size_t)w * 256 + t));
    } else {
        const uint32_t t = w - R;
        for (uint32_t r = j; r < R; r += RGS) acc = xyzz29_add(acc, ld29(base + (size_t)r * 256 + t));
    }
#pragma unroll
    for (int m = 1; m < RGS; m <<= 1) acc = xyzz29_add(acc, shfl_xor_xyzz(acc, m));
    if (j == 0) {
        if (w < R)
            st29(Z + (size_t)p * R + w, acc);
        else
            st29(U + (size_t)p * 256 + (w - R), acc);
    }
}

// For m <= 256 items X_0..X_{m-1}: WS = sum_i i X_i and TOT = sum_i X_i, by an
// LDS suffix scan (Suf_i = sum_{j >= i} X_j; WS = sum_{i >= 1} Suf_i).
// Block b < na reads problem b of a (ma items each), block na + b problem b of
// b_src (mb items each); block b writes out[2b] = WS, out[2b + 1] = TOT.
__global__ void __launch_bounds__(TPB) k_weighted(const Xyzz29 *__restrict__ a_src, uint32_t ma,
                                                uint32_t na, const Xyzz29 *__restrict__ b_src,
                                                uint32_t mb, Xyzz *__restrict__ out) {
    __shared__ Xyzz29 lds[TPB];
    const uint32_t tid = threadIdx.x;
    const bool first = blockIdx.x < na;
    const uint32_t m = first ? ma : mb;
    const Xyzz29 *src = first ? a_src : b_src;
    const size_t base = (size_t)(first ? blockIdx.x : blockIdx.x - na) * m;
    lds[tid] = tid < m ? ld29(src + base + tid) : xyzz29_identity();
    __syncthreads();
    for (uint32_t s = 1; s < m; s <<= 1) {
        Xyzz29 v = lds[tid];
        const bool take = tid + s < m;
        Xyzz29 o = take ? lds[tid + s] : v;
        __syncthreads();
        if (take) lds[tid] = xyzz29_add(v, o);
        __syncthreads();
    }
    const Xyzz29 tot = lds[0];
    __syncthreads();
    if (tid == 0) lds[0] = xyzz29_identity();
    __syncthreads();
    for (uint32_t s = TPB / 2; s > 0; s >>= 1) {
        if (tid < s && tid + s < m) lds[tid] = xyzz29_add(lds[tid], lds[tid + s]);
        __syncthreads();
    }
    if (tid == 0) {
        out[2 * blockIdx.x] = xyzz29_store(lds[0]);
        out[2 * blockIdx.x + 1] = xyzz29_store(tot);
    }
}

// Window-shifted copies: copy w of base i is 2^(c w) B_i, affine, written in
// the bucket phase's domain (coordinates times 2^5 mod p, affine_to_table).
__global__ void k_expand(const Affine *__restrict__ src, size_t n, int c, int copies,
                         Affine *__restrict__ dst, size_t stride) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Affine a = src[i];
    dst[i] = affine_to_table(a);
    Xyzz p = xyzz_from_affine(a);
    for (int w = 1; w < copies; w++) {
        for (int d = 0; d < c; d++) p = xyzz_dbl(p);
        dst[(size_t)w * stride + i] = affine_to_table(xyzz_to_affine(p));
    }
}

__global__ void k_table_convert(Affine *__restrict__ pts, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pts[i] = affine_to_table(pts[i]);
}

}  // namespace

Xyzz host_xyzz_mul_pow2(Xyzz p, int k) {
    for (int i = 0; i < k; i++) p = xyzz_dbl(p);
    return p;
}

void host_batch_to_affine(const Xyzz *in, Affine *out, size_t k) {
    // Montgomery's trick over the nonzero ZZ * ZZZ products, one inversion.
    std::vector<Fe> t(k), pre(k + 1);
    pre[0] = f_one<Fq>();
    for (size_t i = 0; i < k; i++) {
        t[i] = xyzz_is_id(in[i]) ? f_one<Fq>() : f_mul<Fq>(in[i].zz, in[i].zzz);
        pre[i + 1] = f_mul<Fq>(pre[i], t[i]);
    }
    Fe inv = f_inv<Fq>(pre[k]);
    for (size_t i = k; i-- > 0;) {
        Fe ti = f_mul<Fq>(inv, pre[i]);  // 1 / t[i]
        inv = f_mul<Fq>(inv, t[i]);
        if (xyzz_is_id(in[i])) {
            out[i].x = f_zero<Fq>();
            out[i].y = f_zero<Fq>();
            continue;
        }
        Fe zz_inv = f_mul<Fq>(ti, in[i].zzz);
        Fe zzz_inv = f_mul<Fq>(ti, in[i].zz);
        out[i].x = f_mul<Fq>(in[i].x, zz_inv);
        out[i].y = f_mul<Fq>(in[i].y, zzz_inv);
    }
}

namespace {
std::atomic<int> g_window_override[64];  // by floor(log2 n); 0 = the default rule
}

void msm_set_window_bits(int log2n, int c) {
    if (log2n < 0 || log2n >= 64) throw ApgError(-2, "window table: log2 n out of range");
    if (c != 0 && (c < MSM_MIN_C || c > MSM_MAX_C)) throw ApgError(-2, "window table: width out of range");
    g_window_override[log2n] = c;
}

// Built-in widths for floor(log2 n) = 8..22, the fastest of c in default +- 3
// for 8 witness-like columns on MI355X (apogee-vm_amd/autotune.py); outside
// that range, lg n - 2 clamped to [4, 17].
static const int kTunedC[23] = {0, 0, 0, 0, 0, 0, 0, 0, 9, 9, 8, 8, 9, 12, 12, 14, 14, 17, 17, 17, 17, 17, 17};

int msm_pick_c(size_t n, size_t columns) {
    (void)columns;
    int lg = 0;
    while (lg < 63 && (size_t(1) << (lg + 1)) <= n) lg++;
    if (const int o = g_window_override[lg]) return o;
    if (n < 32) return 3;
    if (lg < 23 && kTunedC[lg]) return kTunedC[lg];
    int c = lg - 2;
    if (c < 4) c = 4;
    if (c > 17) c = 17;
    return c;
}

void msm_table_convert(Context &ctx, Affine *points, size_t n) {
    if (!n) return;
    hipLaunchKernelGGL(k_table_convert, dim3(grid_for(n, TPB)), dim3(TPB), 0, ctx.stream, points, n);
    HIP_CHECK(hipGetLastError());
}

std::unique_ptr<BaseTable> msm_expand_bases(Context &ctx, const Affine *src, size_t n, int c,
                                            int copies) {
    auto t = std::make_unique<BaseTable>();
    t->c = c;
    t->copies = copies;
    t->stride = n;
    t->owned = true;
    HIP_CHECK(hipMalloc(&t->points, std::max<size_t>(1, (size_t)copies * n) * sizeof(Affine)));
    if (n) {
        hipLaunchKernelGGL(k_expand, dim3(grid_for(n, 128)), dim3(128), 0, ctx.stream, src, n, c,
                           copies, t->points, n);
        HIP_CHECK(hipGetLastError());
    }
    return t;
}

static void msm_run_group(Context &ctx, const BaseTable &tab, size_t n, size_t K, const Fe *fr_dev,
                          const uint32_t *u32_dev, Affine *out_host);

// Columns are processed in groups of at most MAX_PAIRS digit slots, which
// bounds the scratch one call holds (keys, entries and their sorted copies).
void msm_run(Context &ctx, const BaseTable &tab, size_t n, size_t K, const Fe *fr_dev,
             const uint32_t *u32_dev, Affine *out_host) {
    if (K == 0) return;
    if (n == 0) {
        for (size_t k = 0; k < K; k++) {
            out_host[k].x = f_zero<Fq>();
            out_host[k].y = f_zero<Fq>();
        }
        return;
    }
    constexpr size_t MAX_PAIRS = size_t(1) << 28;
    const size_t windows = (size_t)msm_window_count(fr_dev ? 254 : 32, tab.c);
    size_t per = std::max<size_t>(1, MAX_PAIRS / (n * windows));
    for (size_t k0 = 0; k0 < K; k0 += per) {
        const size_t kg = std::min(per, K - k0);
        msm_run_group(ctx, tab, n, kg, fr_dev ? fr_dev + k0 * n : nullptr,
                      u32_dev ? u32_dev + k0 * n : nullptr, out_host + k0);
    }
}

static void msm_run_group(Context &ctx, const BaseTable &tab, size_t n, size_t K, const Fe *fr_dev,
                          const uint32_t *u32_dev, Affine *out_host) {
    if (n > tab.stride) throw ApgError(-2, "msm: more scalars than table points");
    if (n * K >= (size_t(1) << 31)) throw ApgError(-2, "msm: batch too large");
    MsmShape sh;
    sh.c = tab.c;
    sh.windows = msm_window_count(fr_dev ? 254 : 32, sh.c);
    sh.P = std::min(tab.copies, sh.windows);
    sh.G = (sh.windows + sh.P - 1) / sh.P;
    sh.B = 1u << (sh.c - 1);
    if ((size_t)tab.copies * tab.stride >= (size_t(1) << 31))
        throw ApgError(-2, "msm: table too large for 31-bit entry indices");
    const size_t nkeys64 = K * (size_t)sh.G * sh.B;
    if (nkeys64 >= (size_t(1) << 31)) throw ApgError(-2, "msm: too many buckets");
    const uint32_t nkeys = (uint32_t)nkeys64;
    hipStream_t st = ctx.stream;

    // 1. digits -> dense (key, entry) pairs
    const size_t nthreads = n * K;
    const size_t max_pairs = nthreads * (size_t)sh.windows;
    if (max_pairs >= (size_t(1) << 32)) throw ApgError(-2, "msm: too many digits");
    uint32_t *keys_in = ctx.d<uint32_t>("msm.keys_in", max_pairs);
    uint32_t *vals_in = ctx.d<uint32_t>("msm.vals_in", max_pairs);
    const unsigned eblocks = grid_for(nthreads, TPB);
    uint32_t *counts = ctx.d<uint32_t>("msm.counts", (size_t)eblocks + 1);
    uint32_t *offs = ctx.d<uint32_t>("msm.offs", (size_t)eblocks + 1);
    size_t scan_bytes = 0;
    HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, counts, offs, (int)eblocks + 1, st));
    void *scan_tmp = ctx.d<uint8_t>("msm.scantmp", std::max<size_t>(scan_bytes, 1));
    const int stage = sh.windows <= STAGE_WINDOWS ? 1 : 0;
    const size_t stage_lds = stage ? (size_t)TPB * sh.windows * 2 * sizeof(uint32_t) : 0;
    // compile-time window widths for the built-in size classes, Fr or u32 scalars
#define APG_EMIT(CT, WT)                                                                                        \
    do {                                                                                                        \
        hipLaunchKernelGGL((k_emit_count<CT, WT>), dim3(eblocks), dim3(TPB), 0, st, fr_dev, u32_dev, (uint32_t)n, \
                           (uint32_t)K, sh.c, sh.windows, sh.P, counts);                                        \
        HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, counts, offs, (int)eblocks + 1, st));  \
        hipLaunchKernelGGL((k_emit<CT, WT>), dim3(eblocks), dim3(TPB), stage_lds, st, fr_dev, u32_dev,           \
                           (uint32_t)n, (uint32_t)K, sh.c, sh.windows, sh.P, sh.G, sh.B, (uint32_t)tab.stride,  \
                           stage, offs, keys_in, vals_in, h_total_dev);                                         \
    } while (0)
#define APG_EMIT_C(CT)                                                  \
    case CT:                                                            \
        if (fr_dev)                                                     \
            APG_EMIT(CT, msm_window_count(254, CT));                    \
        else                                                            \
            APG_EMIT(CT, msm_window_count(32, CT));                     \
        break;
    // counts[eblocks] = 0, so the scan's last element is the total; k_emit
    // writes it into pinned host memory, visible after the stream sync
    uint32_t *h_total = ctx.h<uint32_t>("msm.total", 1);
    uint32_t *h_total_dev = nullptr;
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&h_total_dev), h_total, 0));
    HIP_CHECK(hipMemsetAsync(counts + eblocks, 0, sizeof(uint32_t), st));
    switch (sh.c) {
        APG_EMIT_C(7)
        APG_EMIT_C(8)
        APG_EMIT_C(9)
        APG_EMIT_C(11)
        APG_EMIT_C(12)
        APG_EMIT_C(14)
        APG_EMIT_C(17)
        default: APG_EMIT(0, 0); break;
    }
#undef APG_EMIT_C
#undef APG_EMIT
    HIP_CHECK(hipGetLastError());
    ctx.sync();
    const uint32_t total = *h_total;

    // 2. one radix sort of all pairs by bucket key (per-column sorts of the
    // low key bits measured slower: small onesweep passes lose more than the
    // saved pass), then bucket boundaries
    int key_bits = 1;
    while ((1ull << key_bits) < (uint64_t)nkeys) key_bits++;
    // With one bucket group per column (every window has its table copy) the
    // sort keys only the bucket's c - 1 bits: it is stable and the emission is
    // column-major, so each (column, bucket) stays contiguous, in bucket-major
    // order. 16 bits at c = 17 is two onesweep passes instead of three.
    const int bbits = sh.c - 1;
    const bool bucket_major = sh.G == 1 && K > 1 && bbits < key_bits;
    const int sort_bits = bucket_major ? bbits : key_bits;
    // 4 words of slack: the accumulation reads keys and entries as 16-byte vectors
    uint32_t *skeys = ctx.d<uint32_t>("msm.skeys", (size_t)total + 4);
    uint32_t *entries = ctx.d<uint32_t>("msm.entries", (size_t)total + 4);
    size_t tmp_bytes = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys_in, skeys, vals_in, entries,
                                                 (int)total, 0, sort_bits, st));
    void *tmp = ctx.d<uint8_t>("msm.cubtmp", std::max<size_t>(tmp_bytes, 1));
    if (total)
        HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, skeys, vals_in, entries,
                                                     (int)total, 0, sort_bits, st));
    uint32_t *off = ctx.d<uint32_t>("msm.off", 2 * (size_t)nkeys);
    HIP_CHECK(hipMemsetAsync(off, 0, 2 * (size_t)nkeys * sizeof(uint32_t), st));
    if (total) hipLaunchKernelGGL(k_bounds, dim3(grid_for(total, TPB)), dim3(TPB), 0, st, skeys, total, off);
    HIP_CHECK(hipGetLastError());

    // 3. accumulate over equal slices of the sorted entries: about ACC_ROUNDS
    // slices per resident thread (k_accumulate holds 3 waves per SIMD), at
    // least MIN_SLICE entries and an eighth of the mean bucket, so a typical
    // bucket crosses about eight slice boundaries and stays light (16 partials
    // of at most LIGHT_ITEMS)
    const size_t resident = (size_t)ctx.cu_count * 4 * 3 * 64;
    const size_t by_grid = (total + ACC_ROUNDS * resident - 1) / (ACC_ROUNDS * resident);
    const size_t by_bucket = total / (8 * (size_t)std::max<uint32_t>(nkeys, 1));
    const uint32_t L = (uint32_t)(std::min<size_t>(MAX_SLICE, std::max({MIN_SLICE, by_grid, by_bucket})) + 3) & ~3u;
    const uint32_t nslices = (total + L - 1) / L;
    Xyzz29 *buckets = ctx.d<Xyzz29>("msm.buckets", nkeys);
    Xyzz29 *part = ctx.d<Xyzz29>("msm.part", 2 * (size_t)std::max<uint32_t>(nslices, 1));
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (ctx.profiling) {
        HIP_CHECK(hipEventCreate(&ev0));
        HIP_CHECK(hipEventCreate(&ev1));
        HIP_CHECK(hipEventRecord(ev0, st));
    }
    if (nslices)
        hipLaunchKernelGGL(k_accumulate, dim3(grid_for(nslices, TPB)), dim3(TPB), 0, st, skeys, entries,
                           total, L, tab.points, buckets, part);
    HIP_CHECK(hipGetLastError());
    if (ctx.profiling) HIP_CHECK(hipEventRecord(ev1, st));

    // 4. buckets: empty ones set, crossing ones merged from their partials. A
    // heavy bucket spans more than LIGHT_ITEMS / 2 slices, each slice belongs
    // to at most one heavy span, and spans hold 2 partials per slice.
    const uint32_t max_heavy = std::min<uint32_t>(nkeys, nslices / (LIGHT_ITEMS / 2) + 1);
    const uint32_t max_pieces = 2 * nslices / PIECE + max_heavy + 1;
    HeavyPlan *plan = ctx.d<HeavyPlan>("msm.plan", 1);
    uint4 *heavy = ctx.d<uint4>("msm.heavy", max_heavy);
    uint2 *piece_of = ctx.d<uint2>("msm.piece_of", max_pieces);
    Xyzz29 *piece_sums = ctx.d<Xyzz29>("msm.piece_sums", max_pieces);
    HIP_CHECK(hipMemsetAsync(plan, 0, sizeof(HeavyPlan), st));
    hipLaunchKernelGGL(k_classify, dim3(grid_for(nkeys, TPB)), dim3(TPB), 0, st, off, nkeys, L, plan, heavy,
                       piece_of);
    // heavy pieces (a few latency-bound blocks) on the side stream, queued
    // first, beside the light buckets (throughput-bound) here
    HIP_CHECK(hipEventRecord(ctx.ev_fork, st));
    HIP_CHECK(hipStreamWaitEvent(ctx.side, ctx.ev_fork, 0));
    hipLaunchKernelGGL(k_heavy_pieces, dim3(std::min<uint32_t>(max_pieces, 4u * ctx.cu_count)), dim3(TPB), 0,
                       ctx.side, off, L, part, plan, heavy, piece_of, piece_sums);
    hipLaunchKernelGGL(k_heavy_final, dim3(grid_for(max_heavy, TPB)), dim3(TPB), 0, ctx.side, plan, heavy,
                       piece_sums, buckets);
    HIP_CHECK(hipEventRecord(ctx.ev_join, ctx.side));
    hipLaunchKernelGGL(k_bucket_light, dim3(grid_for(nkeys, TPB)), dim3(TPB), 0, st, off, nkeys, L, part,
                       buckets);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamWaitEvent(st, ctx.ev_join, 0));

    // 5. reduce: per problem (column, group): W = sum_b (b + 1) S_b
    const uint32_t nprob = (uint32_t)(K * sh.G);
    const uint32_t B = sh.B;
    // k_weighted writes the window sums straight into pinned host memory
    Xyzz *h_w = ctx.h<Xyzz>("msm.hw", (size_t)nprob * 4);
    Xyzz *wout = nullptr;
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&wout), h_w, 0));
    if (B > (uint32_t)TPB) {
        const uint32_t R = B / TPB;
        Xyzz29 *Z = ctx.d<Xyzz29>("msm.Z", (size_t)nprob * R);
        Xyzz29 *U = ctx.d<Xyzz29>("msm.U", (size_t)nprob * TPB);
        const size_t threads = (size_t)nprob * (R + TPB) * RGS;
        hipLaunchKernelGGL(k_reduce_rows, dim3(grid_for(threads, TPB)), dim3(TPB), 0, st, buckets, R,
                           nprob, Z, U);
        // WS/TOT of the Z_r (blocks 0..nprob) and of the U_t (nprob..2 nprob)
        hipLaunchKernelGGL(k_weighted, dim3(2 * nprob), dim3(TPB), 0, st, Z, R, nprob, U,
                           (uint32_t)TPB, wout);
    } else {
        hipLaunchKernelGGL(k_weighted, dim3(nprob), dim3(TPB), 0, st, buckets, B, nprob, buckets, B,
                           wout);
    }
    HIP_CHECK(hipGetLastError());
    ctx.sync();

    // 6. host tail
    std::vector<Xyzz> res(K);
    for (size_t col = 0; col < K; col++) {
        Xyzz acc = xyzz_identity();
        for (int g = sh.G - 1; g >= 0; g--) {
            acc = host_xyzz_mul_pow2(acc, sh.c * sh.P);
            const uint32_t p = (uint32_t)(col * sh.G + g);
            Xyzz W;
            if (B > (uint32_t)TPB) {
                // 256 * WS(Z) + WS(U) + TOT(Z)
                W = xyzz_add(host_xyzz_mul_pow2(h_w[2 * p], 8), h_w[2 * nprob + 2 * p]);
                W = xyzz_add(W, h_w[2 * p + 1]);
            } else {
                W = xyzz_add(h_w[2 * p], h_w[2 * p + 1]);
            }
            acc = xyzz_add(acc, W);
        }
        res[col] = acc;
    }
    host_batch_to_affine(res.data(), out_host, K);
    if (ctx.profiling) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
        ctx.stats.accumulate_ms += ms;
        ctx.stats.accumulate_launches += nslices ? 1 : 0;
        ctx.stats.accumulate_entries += total;
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    }
}

}  // namespace apg
