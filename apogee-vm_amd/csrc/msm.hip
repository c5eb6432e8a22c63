// Multi-scalar multiplication over BN254 G1 for gfx950.
//
// Replaces curve::msm::{msm, msm_small_u32} (crates/curve/src/msm.rs:48-100).
// The value is the group element sum_i s_i B_i, so any correct bucket schedule
// returns the reference's exact affine bytes; this one is built for the MI355X:
//
//  1. digits  : one thread per scalar converts Montgomery -> canonical and
//               recodes into signed c-bit digits exactly as msm.rs:180-224
//               does (same carry rule); every nonzero digit is one (bucket
//               key, table index | sign) pair. The recode runs twice (count,
//               then scatter) instead of storing an unsorted pair array.
//  2. sort    : a two-level counting sort over the known key range (the
//               "digit sort" section below): per-block coarse-bin counts in
//               LDS, row scans and the accumulation plan on the device, an
//               LDS-staged scatter writing each bin contiguously, then a
//               chunked fine pass writing each key's run and its range.
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

// ---- digit sort: a two-level counting sort over the known key range ----
//
// A digit's local key is lk = group * B + (|digit| - 1) < G B = lkeys, its key
// column * lkeys + lk. The sort leaves every key's pairs (key, table index |
// sign) in one contiguous run, keys ascending, and writes each key's range
// [off[2k], off[2k + 1]); within a key the order is arbitrary, since the
// bucket's sum does not depend on it.
//  1. k_sort_count: blocks of SORT_SPB scalars, never straddling a column,
//     recode and count their digits per coarse bin (lk >> fine_bits, at most
//     256 bins per column) in LDS.
//  2. k_sort_rowscan: for each (column, coarse bin) row, the exclusive scan of
//     its blocks' counts. k_sort_base: the scan of the rows' totals, the pair
//     total, the accumulation's slice length (the plan) and the fine pass's
//     extra chunks, on the device.
//  3. k_sort_scatter: the blocks recode again, stage their pairs in LDS grouped
//     by coarse bin and write each bin's segment contiguously.
//  4. k_sort_fine: each coarse bin is ordered by the key's low fine_bits, in
//     chunks of at most `cap` pairs, one chunk per workgroup, staged in LDS and
//     written contiguously per fine bucket; a row of several chunks (the
//     witness mix puts every scalar equal to one in one bucket per column) is
//     counted first (k_sort_fine_count) and its chunks claim their places per
//     bucket with one global atomic each. The ranges of the keys come out of
//     the same counts. When lkeys <= 256 the coarse bins are the keys and this
//     pass is not needed (k_sort_base writes the ranges).
// Nothing is read back to the host before the MSM's end.
constexpr uint32_t LIGHT_ITEMS = 32;  // more partials than this: a heavy bucket
constexpr size_t ACC_ROUNDS = 4, MIN_SLICE = 16, MAX_SLICE = 4096;

constexpr int SORT_TPB = 256;
constexpr int SORT_SPT = 4;                                // scalars per thread
constexpr uint32_t SORT_SPB = (uint32_t)SORT_TPB * SORT_SPT;  // scalars per block
constexpr uint32_t MAX_COARSE = 256;
constexpr int MAX_FINE_BITS = 12;
// LDS: a scatter block stages up to SCATTER_CAP pairs (40 KiB); a fine
// workgroup holds a chunk of at most FINE_CAP pairs, FINE_PER_THREAD to a
// thread, and stages it (60 KiB, less when its counters need more room, so
// that it stays within 64 KiB). Longer fine chunks write longer runs per key:
// 7680 against 4096 measured the 2^20 x 8 commit 2% faster and the mixed
// batch 1-2% (3072 / 7168 for the scatter stage measured no gain).
constexpr uint32_t SCATTER_CAP = 5120;
constexpr uint32_t FINE_CAP = 7680;
// the fine pass's workgroups: 512 threads (two per CU fit the LDS), so that
// its phases between barriers, each short, have eight waves per workgroup
constexpr int FINE_TPB = 512;
constexpr int FINE_PER_THREAD = FINE_CAP / FINE_TPB;
static_assert(FINE_CAP % FINE_TPB == 0, "fine chunks hold whole rows of the workgroup");
constexpr size_t LDS_LIMIT = 65536;
constexpr size_t MAX_GROUP_COLUMNS = 64;  // keeps k_sort_base's rows <= 64 x 256

struct SortGeom {
    uint32_t n, K, nbc;  // scalars per column, columns, blocks per column
    int c, windows, P;
    uint32_t B, lkeys;   // buckets per group, local keys per column (G B)
    uint32_t CB;         // coarse bins per column: ceil(lkeys / 2^fine_bits) <= 256
    int fine_bits;
    uint32_t tab_stride;
};

// Device-side plan of one MSM group, written by k_sort_base: the pair total,
// the accumulation's slice length and slice count, the fine pass's extra
// chunks, and the heavy-bucket allocation k_classify fills.
struct HeavyPlan {
    uint32_t nheavy, npieces;
};
struct MsmPlan {
    uint32_t total, L, nslices, nextra;
    HeavyPlan heavy;
};

// Exclusive scan of one value per thread over a TPB-thread block; `total`
// receives the block's sum. Every thread of the block calls it.
template <int TPB = SORT_TPB>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t &total) {
    __shared__ uint32_t ws[TPB / 64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, 64);
        if ((int)lane >= off) incl += o;
    }
    if (lane == 63) ws[wave] = incl;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < TPB / 64; w++) {
        if (w < (int)wave) before += ws[w];
        tot += ws[w];
    }
    __syncthreads();  // ws is reused by the next call
    total = tot;
    return before + incl - v;
}

// A block's scalars, loaded before any is recoded so the loads overlap.
__device__ __forceinline__ void load_block_scalars(const Fe *fr, const uint32_t *u32, const SortGeom &g,
                                                   uint32_t col, uint32_t i0, uint32_t (&l)[SORT_SPT][8]) {
#pragma unroll
    for (int s = 0; s < SORT_SPT; s++) {
        const uint32_t i = i0 + (uint32_t)s * SORT_TPB;
#pragma unroll
        for (int w = 0; w < 8; w++) l[s][w] = 0;
        if (i < g.n) load_scalar(fr, u32, (size_t)col * g.n + i, l[s]);
    }
}

template <int CT, int WT>
__global__ void __launch_bounds__(SORT_TPB) k_sort_count(const Fe *__restrict__ fr, const uint32_t *__restrict__ u32,
                                                       SortGeom g, uint32_t *__restrict__ counts) {
    raise_wave_priority();
    __shared__ uint32_t hist[MAX_COARSE];
    const uint32_t col = blockIdx.x / g.nbc, b = blockIdx.x - col * g.nbc;
    for (uint32_t i = threadIdx.x; i < g.CB; i += SORT_TPB) hist[i] = 0;
    const uint32_t i0 = b * SORT_SPB + threadIdx.x;
    uint32_t l[SORT_SPT][8];
    load_block_scalars(fr, u32, g, col, i0, l);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SORT_SPT; s++) {
        if (i0 + (uint32_t)s * SORT_TPB >= g.n) break;
        recode_digits<CT, WT>(l[s], g.c, g.windows, g.P, [&](int, int group, uint32_t mag, bool) {
            atomicAdd(&hist[((uint32_t)group * g.B + (mag - 1)) >> g.fine_bits], 1u);
        });
    }
    __syncthreads();
    for (uint32_t bin = threadIdx.x; bin < g.CB; bin += SORT_TPB)
        counts[((size_t)col * g.CB + bin) * g.nbc + b] = hist[bin];
}

// Row r = column * CB + bin: counts[r][0..nbc) becomes its exclusive scan,
// rowtot[r] its sum. A row of more than cap pairs (fine pass, several chunks)
// gets its fine counters and claims cleared.
__global__ void __launch_bounds__(SORT_TPB) k_sort_rowscan(uint32_t *__restrict__ counts, uint32_t nbc,
                                                         uint32_t *__restrict__ rowtot, int fine_bits, uint32_t cap,
                                                         uint32_t *__restrict__ fcount, uint32_t *__restrict__ fclaim) {
    raise_wave_priority();
    uint32_t *row = counts + (size_t)blockIdx.x * nbc;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nbc; base += SORT_TPB) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nbc ? row[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, tot);
        if (i < nbc) row[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) rowtot[blockIdx.x] = carry;
    if (fine_bits > 0 && carry > cap) {
        const size_t FB = size_t(1) << fine_bits, f0 = (size_t)blockIdx.x * FB;
        for (size_t f = threadIdx.x; f < FB; f += SORT_TPB) {
            fcount[f0 + f] = 0;
            fclaim[f0 + f] = 0;
        }
    }
}

// One block: rowbase[r] = the pairs of the rows before r (every row of the
// group in column-major, bin-ascending order: the final key order), the plan,
// the fine pass's extra chunks (chunks 1.. of every row of more than cap
// pairs, as (row, chunk)), and, when rows are keys, the key ranges.
// `target` is the slice count wanted (the size class's slice rounds, twice
// that for a job that pauses between accumulation chunks, per thread
// k_accumulate keeps resident); slices are at least MIN_SLICE long, at most total / target and an
// eighth of the mean bucket, a multiple of 4, and never more than slice_cap.
__global__ void __launch_bounds__(SORT_TPB) k_sort_base(const uint32_t *__restrict__ rowtot, uint32_t rows,
                                                      uint32_t nkeys, uint32_t target, uint32_t slice_cap,
                                                      int rows_are_keys, uint32_t cap,
                                                      uint32_t *__restrict__ rowbase, uint32_t *__restrict__ off,
                                                      uint2 *__restrict__ extra, MsmPlan *__restrict__ plan,
                                                      uint32_t *__restrict__ total_host) {
    raise_wave_priority();
    uint32_t carry = 0, xcarry = 0;
    for (uint32_t base = 0; base < rows; base += SORT_TPB) {
        const uint32_t r = base + threadIdx.x;
        const uint32_t v = r < rows ? rowtot[r] : 0u;
        const uint32_t xc = (!rows_are_keys && v > cap) ? (v + cap - 1) / cap - 1 : 0u;
        uint32_t tot, xtot;
        const uint32_t ex = carry + block_exclusive_scan(v, tot);
        const uint32_t xex = xcarry + block_exclusive_scan(xc, xtot);
        if (r < rows) {
            rowbase[r] = ex;
            if (rows_are_keys) {
                off[2 * (size_t)r] = ex;
                off[2 * (size_t)r + 1] = ex + v;
            }
            for (uint32_t c = 0; c < xc; c++) extra[xex + c] = make_uint2(r, c + 1);
        }
        carry += tot;
        xcarry += xtot;
    }
    if (threadIdx.x == 0) {
        const uint64_t total = carry;
        uint64_t L = MIN_SLICE;
        L = max(L, (total + (uint64_t)target - 1) / (uint64_t)target);
        L = max(L, total / (8 * (uint64_t)max(nkeys, 1u)));
        L = min<uint64_t>(L, MAX_SLICE);
        L = max(L, (total + slice_cap - 1) / slice_cap);
        L = (L + 3) & ~uint64_t(3);
        plan->total = (uint32_t)total;
        plan->L = (uint32_t)L;
        plan->nslices = (uint32_t)((total + L - 1) / L);
        plan->nextra = xcarry;
        plan->heavy.nheavy = 0;
        plan->heavy.npieces = 0;
        *total_host = (uint32_t)total;
    }
}

// The block's pairs are staged in LDS grouped by coarse bin (each bin's
// region at its exclusive prefix of the block's counts) and then written out
// with consecutive lanes on consecutive pairs of a bin, so every bin's
// segment of the block is one contiguous write. Bins are taken in ranges
// whose pairs fit the stage, one recode pass per range (one pass unless the
// block holds more than SCATTER_CAP pairs); a single bin above the stage is
// written directly.
template <int CT, int WT>
__global__ void __launch_bounds__(SORT_TPB) k_sort_scatter(const Fe *__restrict__ fr,
                                                         const uint32_t *__restrict__ u32, SortGeom g,
                                                         const uint32_t *__restrict__ counts,
                                                         const uint32_t *__restrict__ rowtot,
                                                         const uint32_t *__restrict__ rowbase,
                                                         uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    raise_wave_priority();
    __shared__ uint2 stage[SCATTER_CAP];
    __shared__ uint32_t cnt[MAX_COARSE], lbase[MAX_COARSE], gbase[MAX_COARSE], cur[MAX_COARSE];
    __shared__ uint32_t range_count;
    const uint32_t tid = threadIdx.x;
    const uint32_t col = blockIdx.x / g.nbc, b = blockIdx.x - col * g.nbc;
    const uint32_t i0 = b * SORT_SPB + tid;
    for (uint32_t bin = tid; bin < g.CB; bin += SORT_TPB) {
        const size_t r = (size_t)col * g.CB + bin;
        const uint32_t p = counts[r * g.nbc + b];
        const uint32_t nx = b + 1 < g.nbc ? counts[r * g.nbc + b + 1] : rowtot[r];
        cnt[bin] = nx - p;
        gbase[bin] = rowbase[r] + p;
    }
    __syncthreads();
    const uint32_t key0 = col * g.lkeys;
    uint32_t lo = 0;
    while (lo < g.CB) {  // uniform: every thread derives the same ranges
        // bins [lo, hi): as many as fit the stage, at least one (CB <= SORT_TPB:
        // one bin per thread)
        const uint32_t bin = tid;
        const bool live = bin >= lo && bin < g.CB;
        const uint32_t v = live ? cnt[bin] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, tot);
        const int nfit = __syncthreads_count(live && ex + v <= SCATTER_CAP);
        const uint32_t hi = lo + (uint32_t)max(nfit, 1);
        if (bin == hi - 1) range_count = ex + v;
        if (live && bin < hi) {
            lbase[bin] = ex;
            cur[bin] = ex;
        }
        __syncthreads();
        const uint32_t staged = range_count;
        const bool direct = staged > SCATTER_CAP;  // one bin larger than the stage
        if (direct && tid == 0) cur[lo] = gbase[lo];
        __syncthreads();
        // one scalar at a time (unrolled, the recode's registers quadruple)
#pragma unroll 1
        for (int sc = 0; sc < SORT_SPT; sc++) {
            const uint32_t i = i0 + (uint32_t)sc * SORT_TPB;
            if (i >= g.n) break;
            uint32_t l[8];
            load_scalar(fr, u32, (size_t)col * g.n + i, l);
            recode_digits<CT, WT>(l, g.c, g.windows, g.P, [&](int copy, int group, uint32_t mag, bool neg) {
                const uint32_t lk = (uint32_t)group * g.B + (mag - 1);
                const uint32_t bn = lk >> g.fine_bits;
                if (bn < lo || bn >= hi) return;
                const uint32_t val = ((uint32_t)copy * g.tab_stride + i) | (neg ? SIGN : 0u);
                const uint32_t slot = atomicAdd(&cur[bn], 1u);
                if (direct) {
                    keys[slot] = key0 + lk;
                    vals[slot] = val;
                } else {
                    stage[slot] = make_uint2(key0 + lk, val);
                }
            });
        }
        __syncthreads();
        if (!direct)
            for (uint32_t e = tid; e < staged; e += SORT_TPB) {
                const uint2 kv = stage[e];
                const uint32_t bn = (kv.x - key0) >> g.fine_bits;
                const uint32_t dest = gbase[bn] + (e - lbase[bn]);
                keys[dest] = kv.x;
                vals[dest] = kv.y;
            }
        __syncthreads();
        lo = hi;
    }
}

// The fine pass's work items: workgroup w < rows takes chunk 0 of row w; the
// others take the plan's extra chunks, striding over them. Calls f(row,
// chunk) per item.
template <class F>
__device__ __forceinline__ void fine_items(uint32_t rows, const MsmPlan *plan, const uint2 *extra, F &&f) {
    const uint32_t w = blockIdx.x;
    if (w < rows) {
        f(w, 0u);
        return;
    }
    const uint32_t nx = plan->nextra, stride = gridDim.x - rows;
    for (uint32_t x = w - rows; x < nx; x += stride) {
        const uint2 it = extra[x];
        f(it.x, it.y);
    }
}

// A chunk of pairs into registers (FINE_PER_THREAD to a thread, m of them
// live).
__device__ __forceinline__ void load_chunk(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t at,
                                           uint32_t m, uint2 (&v)[FINE_PER_THREAD]) {
#pragma unroll
    for (int q = 0; q < FINE_PER_THREAD; q++) {
        const uint32_t e = threadIdx.x + (uint32_t)q * FINE_TPB;
        v[q] = make_uint2(0u, 0u);
        if (e < m) v[q] = make_uint2(keys_in[at + e], vals_in[at + e]);
    }
}

// Fine counts of every row of more than cap pairs: each chunk counts in LDS
// and adds its counts to the row's (global atomics, one per bucket).
__global__ void __launch_bounds__(FINE_TPB) k_sort_fine_count(const uint32_t *__restrict__ keys_in, SortGeom g,
                                                            uint32_t cap, uint32_t rows,
                                                            const uint32_t *__restrict__ rowbase,
                                                            const uint32_t *__restrict__ rowtot,
                                                            const MsmPlan *__restrict__ plan,
                                                            const uint2 *__restrict__ extra,
                                                            uint32_t *__restrict__ fcount) {
    raise_wave_priority();
    extern __shared__ uint32_t sm[];
    const uint32_t FB = 1u << g.fine_bits, mask = FB - 1;
    fine_items(rows, plan, extra, [&](uint32_t r, uint32_t chunk) {
        const uint32_t E = rowtot[r];
        if (E <= cap) return;  // uniform per workgroup
        const uint32_t c0 = chunk * cap, m = min(cap, E - c0);
        __syncthreads();
        for (uint32_t f = threadIdx.x; f < FB; f += FINE_TPB) sm[f] = 0;
        __syncthreads();
        uint32_t k[FINE_PER_THREAD];
#pragma unroll
        for (int q = 0; q < FINE_PER_THREAD; q++) {  // every load in flight before the counts
            const uint32_t e = threadIdx.x + (uint32_t)q * FINE_TPB;
            k[q] = 0;
            if (e < m) k[q] = keys_in[rowbase[r] + c0 + e];
        }
#pragma unroll
        for (int q = 0; q < FINE_PER_THREAD; q++)
            if (threadIdx.x + (uint32_t)q * FINE_TPB < m) atomicAdd(&sm[k[q] & mask], 1u);
        __syncthreads();
        for (uint32_t f = threadIdx.x; f < FB; f += FINE_TPB)
            if (sm[f]) atomicAdd(&fcount[(size_t)r * FB + f], sm[f]);
    });
}

// Dynamic LDS: 3 x 2^fine_bits words of counters, then cap <= FINE_CAP pairs
// (cap a multiple of FINE_TPB).
__global__ void __launch_bounds__(FINE_TPB) k_sort_fine(const uint32_t *__restrict__ keys_in,
                                                      const uint32_t *__restrict__ vals_in, SortGeom g, uint32_t cap,
                                                      uint32_t rows, const uint32_t *__restrict__ rowbase,
                                                      const uint32_t *__restrict__ rowtot,
                                                      const MsmPlan *__restrict__ plan,
                                                      const uint2 *__restrict__ extra,
                                                      const uint32_t *__restrict__ fcount,
                                                      uint32_t *__restrict__ fclaim, uint32_t *__restrict__ skeys,
                                                      uint32_t *__restrict__ entries, uint32_t *__restrict__ off) {
    raise_wave_priority();
    extern __shared__ uint32_t sm[];
    const uint32_t FB = 1u << g.fine_bits, mask = FB - 1;
    uint32_t *rstart = sm;                                 // the row's bucket starts (several chunks)
    uint32_t *lcur = sm + FB;                              // the chunk's counts, then its cursors
    uint32_t *dl = sm + 2 * FB;                            // stage index -> destination, per bucket
    uint2 *stage = reinterpret_cast<uint2 *>(sm + 3 * FB);  // cap pairs
    const uint32_t tid = threadIdx.x;
    fine_items(rows, plan, extra, [&](uint32_t r, uint32_t chunk) {
        const uint32_t col = r / g.CB, bin = r - col * g.CB;
        const uint32_t start = rowbase[r], E = rowtot[r];
        const uint32_t lk0 = bin << g.fine_bits;
        const bool single = E <= cap;
        const uint32_t c0 = chunk * cap, m = min(cap, E - c0);
        __syncthreads();
        for (uint32_t f = tid; f < FB; f += FINE_TPB) lcur[f] = 0;
        uint2 v[FINE_PER_THREAD];
        load_chunk(keys_in, vals_in, start + c0, m, v);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < FINE_PER_THREAD; q++)
            if (tid + (uint32_t)q * FINE_TPB < m) atomicAdd(&lcur[v[q].x & mask], 1u);
        __syncthreads();
        // the keys' ranges: from the chunk's counts (a one-chunk row) or the
        // row's (written by chunk 0); the row's bucket starts in rstart
        if (single || chunk == 0) {
            uint32_t carry = start;
            for (uint32_t b0 = 0; b0 < FB; b0 += FINE_TPB) {
                const uint32_t f = b0 + tid;
                const uint32_t cnt = f < FB ? (single ? lcur[f] : fcount[(size_t)r * FB + f]) : 0u;
                uint32_t tot;
                const uint32_t at = carry + block_exclusive_scan<FINE_TPB>(cnt, tot);
                if (f < FB) {
                    rstart[f] = at;
                    if (lk0 + f < g.lkeys) {
                        const size_t key = (size_t)col * g.lkeys + lk0 + f;
                        off[2 * key] = at;
                        off[2 * key + 1] = at + cnt;
                    }
                }
                carry += tot;
            }
        } else {
            uint32_t carry = start;
            for (uint32_t b0 = 0; b0 < FB; b0 += FINE_TPB) {
                const uint32_t f = b0 + tid;
                const uint32_t cnt = f < FB ? fcount[(size_t)r * FB + f] : 0u;
                uint32_t tot;
                const uint32_t at = carry + block_exclusive_scan<FINE_TPB>(cnt, tot);
                if (f < FB) rstart[f] = at;
                carry += tot;
            }
        }
        __syncthreads();
        // chunk cursors (exclusive prefix of the chunk's counts) and where each
        // bucket's staged pairs go: right after the row's earlier pairs of the
        // bucket (one chunk: at once; several: claimed per bucket)
        {
            uint32_t carry = 0;
            for (uint32_t b0 = 0; b0 < FB; b0 += FINE_TPB) {
                const uint32_t f = b0 + tid;
                const uint32_t cnt = f < FB ? lcur[f] : 0u;
                uint32_t tot;
                const uint32_t lb = carry + block_exclusive_scan<FINE_TPB>(cnt, tot);
                if (f < FB) {
                    uint32_t at = rstart[f];
                    if (!single && cnt) at += atomicAdd(&fclaim[(size_t)r * FB + f], cnt);
                    dl[f] = at - lb;
                    lcur[f] = lb;
                }
                carry += tot;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < FINE_PER_THREAD; q++)
            if (tid + (uint32_t)q * FINE_TPB < m) stage[atomicAdd(&lcur[v[q].x & mask], 1u)] = v[q];
        __syncthreads();
        for (uint32_t e = tid; e < m; e += FINE_TPB) {
            const uint2 kv = stage[e];
            const uint32_t dest = e + dl[kv.x & mask];
            skeys[dest] = kv.x;
            entries[dest] = kv.y;
        }
    });
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
// The launches cover the plan's largest possible slice count (one launch, or
// consecutive chunks of slices starting at t0 where a job must be able to
// pause between them); threads past the pair total (plan->total, plan->L:
// k_sort_base) return at once.
__global__ void __launch_bounds__(TPB) k_accumulate(const uint32_t *__restrict__ skeys,
                                                  const uint32_t *__restrict__ entries,
                                                  const MsmPlan *__restrict__ plan, const Affine *__restrict__ bases,
                                                  Xyzz29 *__restrict__ buckets, Xyzz29 *__restrict__ part,
                                                  uint32_t t0, int raise_prio) {
    if (raise_prio) raise_wave_priority();
    const uint32_t total = plan->total, L = plan->L;
    const uint32_t t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
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


// One thread per bucket: empty buckets become the identity, buckets with at
// most LIGHT_ITEMS partials are summed here, heavier ones by the pieces below.
__global__ void __launch_bounds__(TPB, 2) k_bucket_light(const uint32_t *__restrict__ off, uint32_t nkeys,
                                                       const MsmPlan *__restrict__ plan,
                                                       const Xyzz29 *__restrict__ part,
                                                       Xyzz29 *__restrict__ buckets) {
    raise_wave_priority();
    const uint32_t L = plan->L;
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

// heavy[h] = (bucket, first piece, pieces) of heavy bucket h; piece_of[p] =
// (heavy index, piece within the bucket).
__global__ void k_classify(const uint32_t *__restrict__ off, uint32_t nkeys, MsmPlan *mplan,
                           uint4 *__restrict__ heavy, uint2 *__restrict__ piece_of) {
    raise_wave_priority();
    const uint32_t L = mplan->L;
    HeavyPlan *plan = &mplan->heavy;
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

__global__ void __launch_bounds__(TPB, 2) k_heavy_pieces(const uint32_t *__restrict__ off,
                                                       const Xyzz29 *__restrict__ part,
                                                       const MsmPlan *mplan,
                                                       const uint4 *__restrict__ heavy,
                                                       const uint2 *__restrict__ piece_of,
                                                       Xyzz29 *__restrict__ piece_sums) {
    raise_wave_priority_critical();
    __shared__ Xyzz29 lds[TPB / 64];
    const uint32_t L = mplan->L;
    const uint32_t np = mplan->heavy.npieces, tid = threadIdx.x;
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

__global__ void __launch_bounds__(TPB, 2) k_heavy_final(const MsmPlan *mplan,
                                                      const uint4 *__restrict__ heavy,
                                                      const Xyzz29 *__restrict__ piece_sums,
                                                      Xyzz29 *__restrict__ buckets) {
    raise_wave_priority_critical();
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= mplan->heavy.nheavy) return;
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
    raise_wave_priority();
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
    raise_wave_priority();
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
std::atomic<int> g_rounds_override[64];  // by floor(log2 n); 0 = the built-in table

int floor_log2(size_t n) {
    int lg = 0;
    while (lg < 63 && (size_t(1) << (lg + 1)) <= n) lg++;
    return lg;
}
}

void msm_set_slice_rounds(int log2n, int rounds) {
    if (log2n < 0 || log2n >= 64) throw ApgError(-2, "slice table: log2 n out of range");
    if (rounds != 0 && (rounds < MSM_MIN_ROUNDS || rounds > MSM_MAX_ROUNDS))
        throw ApgError(-2, "slice table: rounds out of range");
    g_rounds_override[log2n] = rounds;
}

// Built-in slices per resident thread for floor(log2 n) = 8..22. The sweep of
// {2, 3, 4, 6, 8} for 8 witness-like columns on MI355X
// (apogee-vm_amd/autotune.py --what rounds) found every class within 2% of
// ACC_ROUNDS = 4 (the accumulation is bound by VALU issue, not by the tail
// of its slices), so the table keeps 4.
static const int kTunedRounds[23] = {0, 0, 0, 0, 0, 0, 0, 0, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};

int msm_pick_rounds(size_t n) {
    const int lg = floor_log2(n);
    if (const int o = g_rounds_override[lg]) return o;
    if (lg < 23 && kTunedRounds[lg]) return kTunedRounds[lg];
    return (int)ACC_ROUNDS;
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

namespace {

// Host-side record of one column group, kept until the MSM's single sync.
struct GroupRun {
    size_t k0, K;
    double p0 = 0, p1 = 1;     // this group's share of the MSM (progress reports)
    MsmShape sh;
    size_t w_off;              // first window sum of this group in the pinned array
    uint32_t *total_host;      // the group's pair count (pinned)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

int ceil_log2(uint64_t x) {
    int b = 0;
    while ((uint64_t(1) << b) < x) b++;
    return b;
}

// Enqueues one column group: digits, sort, accumulation, bucket sums and the
// window sums (written by k_weighted into pinned host memory at w_dev).
void msm_enqueue_group(Context &ctx, const BaseTable &tab, size_t n, GroupRun &g, const Fe *fr_dev,
                       const uint32_t *u32_dev, Xyzz *w_dev, uint32_t *total_dev) {
    const size_t K = g.K;
    MsmShape &sh = g.sh;
    sh.c = tab.c;
    sh.windows = msm_window_count(fr_dev ? 254 : 32, sh.c);
    sh.P = std::min(tab.copies, sh.windows);
    sh.G = (sh.windows + sh.P - 1) / sh.P;
    sh.B = 1u << (sh.c - 1);
    const uint64_t lkeys = (uint64_t)sh.G * sh.B;
    const size_t nkeys64 = K * lkeys;
    if (nkeys64 >= (size_t(1) << 31)) throw ApgError(-2, "msm: too many buckets");
    const uint32_t nkeys = (uint32_t)nkeys64;
    hipStream_t st = ctx.stream;

    // 1-2. the digit sort (k_sort_*), geometry first
    SortGeom geo;
    geo.n = (uint32_t)n;
    geo.K = (uint32_t)K;
    geo.nbc = (uint32_t)((n + SORT_SPB - 1) / SORT_SPB);
    geo.c = sh.c;
    geo.windows = sh.windows;
    geo.P = sh.P;
    geo.B = sh.B;
    geo.lkeys = (uint32_t)lkeys;
    geo.fine_bits = std::max(0, ceil_log2(lkeys) - 8);
    if (geo.fine_bits > MAX_FINE_BITS) throw ApgError(-2, "msm: window geometry outside the sort's range");
    geo.CB = (uint32_t)((lkeys + (uint64_t(1) << geo.fine_bits) - 1) >> geo.fine_bits);
    geo.tab_stride = (uint32_t)tab.stride;
    const bool fine = geo.fine_bits > 0;
    const uint32_t rows = (uint32_t)(K * geo.CB);
    const size_t max_pairs = n * K * (size_t)sh.windows;

    uint32_t *counts = ctx.d<uint32_t>("msm.counts", (size_t)rows * geo.nbc);
    uint32_t *rowtot = ctx.d<uint32_t>("msm.rowtot", rows);
    uint32_t *rowbase = ctx.d<uint32_t>("msm.rowbase", rows);
    MsmPlan *plan = ctx.d<MsmPlan>("msm.plan", 1);
    // 4 words of slack: the accumulation reads keys and entries as 16-byte vectors
    uint32_t *skeys = ctx.d<uint32_t>("msm.skeys", max_pairs + 4);
    uint32_t *entries = ctx.d<uint32_t>("msm.entries", max_pairs + 4);
    uint32_t *keys_c = fine ? ctx.d<uint32_t>("msm.keys_c", max_pairs) : skeys;
    uint32_t *vals_c = fine ? ctx.d<uint32_t>("msm.vals_c", max_pairs) : entries;
    uint32_t *off = ctx.d<uint32_t>("msm.off", 2 * (size_t)nkeys);
    // the fine pass: chunk capacity FINE_CAP, or what 64 KiB leaves beside the
    // counters (256 bytes kept for the scans' static words); per-row fine
    // counts and claims, and the list of extra chunks
    const size_t counters = 3 * (size_t(4) << geo.fine_bits);
    const uint32_t cap =
        (uint32_t)std::min<size_t>(FINE_CAP, (LDS_LIMIT - 256 - counters) / sizeof(uint2) / FINE_TPB * FINE_TPB);
    const size_t xmax = max_pairs / cap + 1;
    const uint32_t xgroups = (uint32_t)std::min<size_t>(xmax, 4 * (size_t)ctx.cu_count);
    uint32_t *fcount = fine ? ctx.d<uint32_t>("msm.fcount", (size_t)rows << geo.fine_bits) : nullptr;
    uint32_t *fclaim = fine ? ctx.d<uint32_t>("msm.fclaim", (size_t)rows << geo.fine_bits) : nullptr;
    uint2 *extra = ctx.d<uint2>("msm.extra", xmax);

    // k_accumulate holds 3 waves per SIMD; slices follow k_sort_base's rule,
    // so at most `rounds` per resident thread (+1) and at most one per
    // MIN_SLICE pairs. A preemptible job launches one resident wave of slices
    // at a time and pauses between launches, so its slices are half as long.
    const size_t resident = (size_t)ctx.cu_count * 4 * 3 * 64;
    const size_t rounds = (size_t)msm_pick_rounds(n) * (ctx.preemptible() ? 2 : 1);
    const size_t slice_cap = std::min(rounds * resident + 1, max_pairs / MIN_SLICE + 1);

    const unsigned sblocks = (unsigned)(K * geo.nbc);
#define APG_SORT(CT, WT)                                                                                        \
    do {                                                                                                        \
        hipLaunchKernelGGL((k_sort_count<CT, WT>), dim3(sblocks), dim3(SORT_TPB), 0, st, fr_dev, u32_dev, geo,   \
                           counts);                                                                             \
        hipLaunchKernelGGL(k_sort_rowscan, dim3(rows), dim3(SORT_TPB), 0, st, counts, geo.nbc, rowtot,           \
                           geo.fine_bits, cap, fcount, fclaim);                                                 \
        hipLaunchKernelGGL(k_sort_base, dim3(1), dim3(SORT_TPB), 0, st, rowtot, rows, nkeys,                    \
                           (uint32_t)(rounds * resident),                                                       \
                           (uint32_t)slice_cap, fine ? 0 : 1, cap, rowbase, off, extra, plan, total_dev);       \
        hipLaunchKernelGGL((k_sort_scatter<CT, WT>), dim3(sblocks), dim3(SORT_TPB), 0, st, fr_dev, u32_dev, geo, \
                           counts, rowtot, rowbase, keys_c, vals_c);                                            \
    } while (0)
#define APG_SORT_C(CT)                                                  \
    case CT:                                                            \
        if (fr_dev)                                                     \
            APG_SORT(CT, msm_window_count(254, CT));                    \
        else                                                            \
            APG_SORT(CT, msm_window_count(32, CT));                     \
        break;
    // compile-time window widths for the built-in size classes, Fr or u32 scalars
    switch (sh.c) {
        APG_SORT_C(8)
        APG_SORT_C(9)
        APG_SORT_C(12)
        APG_SORT_C(14)
        APG_SORT_C(17)
        default: APG_SORT(0, 0); break;
    }
#undef APG_SORT_C
#undef APG_SORT
    const auto at = [&](double f) { return g.p0 + (g.p1 - g.p0) * f; };  // progress within the MSM
    if (fine) {
        ctx.yield_point(at(0.1));
        const dim3 fg(rows + xgroups);
        hipLaunchKernelGGL(k_sort_fine_count, fg, dim3(FINE_TPB), size_t(4) << geo.fine_bits, st, keys_c, geo, cap,
                           rows, rowbase, rowtot, plan, extra, fcount);
        hipLaunchKernelGGL(k_sort_fine, fg, dim3(FINE_TPB), counters + (size_t)cap * sizeof(uint2), st, keys_c,
                           vals_c, geo, cap, rows, rowbase, rowtot, plan, extra, fcount, fclaim, skeys, entries,
                           off);
    }
    HIP_CHECK(hipGetLastError());

    // 3. accumulate over equal slices of the sorted entries; a preemptible
    // job launches one resident wave of slices at a time, with a yield point
    // between launches (the slices are independent)
    Xyzz29 *buckets = ctx.d<Xyzz29>("msm.buckets", nkeys);
    Xyzz29 *part = ctx.d<Xyzz29>("msm.part", 2 * slice_cap);
    const bool chunked = ctx.preemptible();
    if (chunked) ctx.yield_point(at(0.18));
    // kernel timing brackets a single launch only: a chunked (preemptible)
    // accumulation may pause between its launches
    if (ctx.profiling && !chunked) {
        HIP_CHECK(hipEventCreate(&g.ev0));
        HIP_CHECK(hipEventCreate(&g.ev1));
        HIP_CHECK(hipEventRecord(g.ev0, st));
    }
    const size_t chunk = chunked ? (resident + TPB - 1) / TPB * TPB : slice_cap;
    for (size_t t0 = 0; t0 < slice_cap; t0 += chunk) {
        if (t0 > 0) ctx.yield_point(at(0.18 + 0.7 * (double)t0 / (double)slice_cap));
        hipLaunchKernelGGL(k_accumulate, dim3(grid_for(std::min(chunk, slice_cap - t0), TPB)), dim3(TPB), 0, st,
                           skeys, entries, plan, tab.points, buckets, part, (uint32_t)t0, ctx.low_acc_prio ? 0 : 1);
        HIP_CHECK(hipGetLastError());
    }
    if (g.ev1) HIP_CHECK(hipEventRecord(g.ev1, st));
    if (chunked) ctx.yield_point(at(0.88));

    // 4. buckets: empty ones set, crossing ones merged from their partials. A
    // heavy bucket spans more than LIGHT_ITEMS / 2 slices, each slice belongs
    // to at most one heavy span, and spans hold 2 partials per slice.
    const uint32_t max_heavy = (uint32_t)std::min<size_t>(nkeys, slice_cap / (LIGHT_ITEMS / 2) + 1);
    const uint32_t max_pieces = (uint32_t)(2 * slice_cap / PIECE + max_heavy + 1);
    uint4 *heavy = ctx.d<uint4>("msm.heavy", max_heavy);
    uint2 *piece_of = ctx.d<uint2>("msm.piece_of", max_pieces);
    Xyzz29 *piece_sums = ctx.d<Xyzz29>("msm.piece_sums", max_pieces);
    hipLaunchKernelGGL(k_classify, dim3(grid_for(nkeys, TPB)), dim3(TPB), 0, st, off, nkeys, plan, heavy, piece_of);
    // heavy pieces (a few latency-bound blocks) on the side stream, queued
    // first, beside the light buckets (throughput-bound) here
    HIP_CHECK(hipEventRecord(ctx.ev_fork, st));
    HIP_CHECK(hipStreamWaitEvent(ctx.side, ctx.ev_fork, 0));
    hipLaunchKernelGGL(k_heavy_pieces, dim3(std::min<uint32_t>(max_pieces, 4u * ctx.cu_count)), dim3(TPB), 0,
                       ctx.side, off, part, plan, heavy, piece_of, piece_sums);
    hipLaunchKernelGGL(k_heavy_final, dim3(grid_for(max_heavy, TPB)), dim3(TPB), 0, ctx.side, plan, heavy,
                       piece_sums, buckets);
    HIP_CHECK(hipEventRecord(ctx.ev_join, ctx.side));
    hipLaunchKernelGGL(k_bucket_light, dim3(grid_for(nkeys, TPB)), dim3(TPB), 0, st, off, nkeys, plan, part,
                       buckets);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamWaitEvent(st, ctx.ev_join, 0));
    ctx.yield_point(at(0.94));

    // 5. reduce: per problem (column, group): W = sum_b (b + 1) S_b, written
    // by k_weighted straight into pinned host memory
    const uint32_t nprob = (uint32_t)(K * sh.G);
    const uint32_t B = sh.B;
    if (B > (uint32_t)TPB) {
        const uint32_t R = B / TPB;
        Xyzz29 *Z = ctx.d<Xyzz29>("msm.Z", (size_t)nprob * R);
        Xyzz29 *U = ctx.d<Xyzz29>("msm.U", (size_t)nprob * TPB);
        const size_t threads = (size_t)nprob * (R + TPB) * RGS;
        hipLaunchKernelGGL(k_reduce_rows, dim3(grid_for(threads, TPB)), dim3(TPB), 0, st, buckets, R, nprob, Z, U);
        // WS/TOT of the Z_r (blocks 0..nprob) and of the U_t (nprob..2 nprob)
        hipLaunchKernelGGL(k_weighted, dim3(2 * nprob), dim3(TPB), 0, st, Z, R, nprob, U, (uint32_t)TPB, w_dev);
    } else {
        hipLaunchKernelGGL(k_weighted, dim3(nprob), dim3(TPB), 0, st, buckets, B, nprob, buckets, B, w_dev);
    }
    HIP_CHECK(hipGetLastError());
}

// Groups combine their windows with 2^(c P) doublings (none when the table
// holds every window shifted); one batched inversion gives affine.
void msm_group_tail(const GroupRun &g, const Xyzz *h_w, Xyzz *res) {
    const MsmShape &sh = g.sh;
    const uint32_t nprob = (uint32_t)(g.K * sh.G);
    const Xyzz *w = h_w + g.w_off;
    for (size_t col = 0; col < g.K; col++) {
        Xyzz acc = xyzz_identity();
        for (int grp = sh.G - 1; grp >= 0; grp--) {
            acc = host_xyzz_mul_pow2(acc, sh.c * sh.P);
            const uint32_t p = (uint32_t)(col * sh.G + grp);
            Xyzz W;
            if (sh.B > (uint32_t)TPB) {
                // 256 * WS(Z) + WS(U) + TOT(Z)
                W = xyzz_add(host_xyzz_mul_pow2(w[2 * p], 8), w[2 * nprob + 2 * p]);
                W = xyzz_add(W, w[2 * p + 1]);
            } else {
                W = xyzz_add(w[2 * p], w[2 * p + 1]);
            }
            acc = xyzz_add(acc, W);
        }
        res[g.k0 + col] = acc;
    }
}

}  // namespace

// Columns are processed in groups of at most MAX_PAIRS digit slots (and
// MAX_GROUP_COLUMNS columns), which bounds the scratch one call holds. All
// groups are enqueued on the context's stream with nothing read back until
// the one synchronisation at the end.
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
    if (n > tab.stride) throw ApgError(-2, "msm: more scalars than table points");
    if ((size_t)tab.copies * tab.stride >= (size_t(1) << 31))
        throw ApgError(-2, "msm: table too large for 31-bit entry indices");
    constexpr size_t MAX_PAIRS = size_t(1) << 28;
    const int windows = msm_window_count(fr_dev ? 254 : 32, tab.c);
    if (n * (size_t)windows >= MAX_PAIRS * 2) throw ApgError(-2, "msm: too many digits for one column");
    const size_t per = std::min(MAX_GROUP_COLUMNS, std::max<size_t>(1, MAX_PAIRS / (n * windows)));
    std::vector<GroupRun> groups;
    size_t wtot = 0;
    for (size_t k0 = 0; k0 < K; k0 += per) {
        GroupRun g;
        g.k0 = k0;
        g.K = std::min(per, K - k0);
        g.w_off = wtot;
        const int P = std::min(tab.copies, windows);
        wtot += 4 * g.K * (size_t)((windows + P - 1) / P);
        groups.push_back(g);
    }
    Xyzz *h_w = ctx.h<Xyzz>("msm.hw", wtot);
    uint32_t *h_tot = ctx.h<uint32_t>("msm.total", groups.size());
    Xyzz *w_dev = nullptr;
    uint32_t *tot_dev = nullptr;
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&w_dev), h_w, 0));
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&tot_dev), h_tot, 0));
    struct EvGuard {
        std::vector<GroupRun> &gs;
        ~EvGuard() {
            for (auto &g : gs) {
                if (g.ev0) (void)hipEventDestroy(g.ev0);
                if (g.ev1) (void)hipEventDestroy(g.ev1);
            }
        }
    } evg{groups};
    for (size_t gi = 0; gi < groups.size(); gi++) {
        GroupRun &g = groups[gi];
        g.total_host = h_tot + gi;
        g.p0 = (double)gi / (double)groups.size();
        g.p1 = (double)(gi + 1) / (double)groups.size();
        if (gi > 0) ctx.yield_point(g.p0);
        msm_enqueue_group(ctx, tab, n, g, fr_dev ? fr_dev + g.k0 * n : nullptr,
                          u32_dev ? u32_dev + g.k0 * n : nullptr, w_dev + g.w_off, tot_dev + gi);
    }
    ctx.sync();
    std::vector<Xyzz> res(K);
    for (const GroupRun &g : groups) msm_group_tail(g, h_w, res.data());
    host_batch_to_affine(res.data(), out_host, K);
    if (ctx.profiling) {
        for (const GroupRun &g : groups) {
            if (!g.ev0) continue;  // a chunked accumulation, not timed
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, g.ev0, g.ev1));
            ctx.stats.accumulate_ms += ms;
            ctx.stats.accumulate_launches += 1;
            ctx.stats.accumulate_entries += *g.total_host;
        }
    }
}

}  // namespace apg
