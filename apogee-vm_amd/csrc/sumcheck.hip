// Sumcheck round evaluation and fold for gfx950.
//
// Replaces the round loop of gkr::prove_sumcheck (crates/gkr/src/lib.rs:792-899)
// and sumcheck::prove_zerocheck (crates/sumcheck/src/lib.rs:359-408): per
// round, the values of g(X) = sum_pairs eq_X * G(lines_X) at X = 0..d, then the
// bind f'(i) = f(2i) + r (f(2i+1) - f(2i)) (crates/poly/src/lib.rs:156-185) of
// every table and of eq.
//
// The fold of round j-1 is fused into the evaluation of round j: one pass per
// round reads each live table once (4 entries per thread = one pair of the
// folded table), writes the half-size fold, and accumulates the node sums, so
// each round streams every table from HBM exactly once. Node 1 is taken from
// the previous round's claim (g_j(1) = g_{j-1}(r_{j-1}) - g_j(0)) from round 1
// on, which is the identity the verifier checks (crates/sumcheck/src/lib.rs:441)
// and saves one product per pair. All arithmetic is exact, so the round
// polynomials are the reference's, coefficient for coefficient.
#include <algorithm>
#include <cstring>

#include "sumcheck.h"

namespace apg {

namespace {

constexpr int RTPB = 256;
constexpr int MAX_RBLOCKS = 2048;

struct TablePtrs {
    const Fe *src[SC_MAX_TABLES + 1];  // last slot: eq
    Fe *dst[SC_MAX_TABLES + 1];
};

struct GateArgs {
    int nterms;
    Term terms[SC_MAX_TERMS];
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

// lo/hi of pair i of table t, folding the previous round's table if asked.
__device__ __forceinline__ void pair_of(const TablePtrs &tp, int t, size_t i, bool fold,
                                        const Fe &r, Fe &lo, Fe &hi) {
    if (fold) {
        const Fe *s = tp.src[t];
        Fe a0 = ldfe(s, 4 * i), a1 = ldfe(s, 4 * i + 1), a2 = ldfe(s, 4 * i + 2),
           a3 = ldfe(s, 4 * i + 3);
        lo = f_add<Fr>(a0, f_mul<Fr>(r, f_sub<Fr>(a1, a0)));
        hi = f_add<Fr>(a2, f_mul<Fr>(r, f_sub<Fr>(a3, a2)));
        stfe(tp.dst[t], 2 * i, lo);
        stfe(tp.dst[t], 2 * i + 1, hi);
    } else {
        lo = ldfe(tp.src[t], 2 * i);
        hi = ldfe(tp.src[t], 2 * i + 1);
    }
}

// Block-level sum of ND Fr accumulators, then the last block to finish sums
// all block partials into out[0..ND).
template <int ND>
__device__ void reduce_nodes(Fe (&acc)[ND], Fe *partials, unsigned *counter, Fe *out) {
    __shared__ Fe lds[RTPB];
    __shared__ bool is_last;
    const int tid = threadIdx.x;
#pragma unroll
    for (int x = 0; x < ND; x++) {
        lds[tid] = acc[x];
        __syncthreads();
        for (int s = RTPB / 2; s > 0; s >>= 1) {
            if (tid < s) lds[tid] = f_add<Fr>(lds[tid], lds[tid + s]);
            __syncthreads();
        }
        if (tid == 0) partials[(size_t)blockIdx.x * ND + x] = lds[0];
        __syncthreads();
    }
    __threadfence();
    if (tid == 0) is_last = atomicAdd(counter, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    for (int x = 0; x < ND; x++) {
        Fe a = f_zero<Fr>();
        for (unsigned b = tid; b < gridDim.x; b += RTPB) {
            // partials written by other blocks: read through the L2
            const Fe *pp = partials + (size_t)b * ND + x;
            Fe v;
            for (int q = 0; q < 8; q++) v.l[q] = __atomic_load_n(&pp->l[q], __ATOMIC_RELAXED);
            a = f_add<Fr>(a, v);
        }
        lds[tid] = a;
        __syncthreads();
        for (int s = RTPB / 2; s > 0; s >>= 1) {
            if (tid < s) lds[tid] = f_add<Fr>(lds[tid], lds[tid + s]);
            __syncthreads();
        }
        if (tid == 0) out[x] = lds[0];
        __syncthreads();
    }
    if (tid == 0) *counter = 0;
}

// Product summand over KT tables (coef 1): the configs' eq * a * b (* c).
template <int KT, int ND>
__global__ void __launch_bounds__(RTPB) k_round_product(TablePtrs tp, size_t m, int fold, Fe r,
                                                      int skip1, Fe *partials,
                                                      unsigned *counter, Fe *out) {
    Fe acc[ND];
#pragma unroll
    for (int x = 0; x < ND; x++) acc[x] = f_zero<Fr>();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        Fe cur[KT], step[KT], e_lo, e_hi;
#pragma unroll
        for (int t = 0; t < KT; t++) {
            Fe lo, hi;
            pair_of(tp, t, i, fold, r, lo, hi);
            cur[t] = lo;
            step[t] = f_sub<Fr>(hi, lo);
        }
        pair_of(tp, SC_MAX_TABLES, i, fold, r, e_lo, e_hi);
        const Fe e_step = f_sub<Fr>(e_hi, e_lo);
        Fe e = e_lo;
#pragma unroll
        for (int x = 0; x < ND; x++) {
            if (!(skip1 && x == 1)) {
                Fe g = cur[0];
#pragma unroll
                for (int t = 1; t < KT; t++) g = f_mul<Fr>(g, cur[t]);
                acc[x] = f_add<Fr>(acc[x], f_mul<Fr>(e, g));
            }
            if (x + 1 < ND) {
                e = f_add<Fr>(e, e_step);
#pragma unroll
                for (int t = 0; t < KT; t++) cur[t] = f_add<Fr>(cur[t], step[t]);
            }
        }
    }
    reduce_nodes<ND>(acc, partials, counter, out);
}

// Any gate: terms read from the kernel argument, table lines kept in a local
// array (dynamic factor indices).
template <int ND>
__global__ void __launch_bounds__(RTPB) k_round_gate(TablePtrs tp, int k, GateArgs gate, size_t m,
                                                   int fold, Fe r, int skip1, Fe *partials,
                                                   unsigned *counter, Fe *out) {
    Fe acc[ND];
#pragma unroll
    for (int x = 0; x < ND; x++) acc[x] = f_zero<Fr>();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        Fe cur[SC_MAX_TABLES], step[SC_MAX_TABLES], e_lo, e_hi;
        for (int t = 0; t < k; t++) {
            Fe lo, hi;
            pair_of(tp, t, i, fold, r, lo, hi);
            cur[t] = lo;
            step[t] = f_sub<Fr>(hi, lo);
        }
        pair_of(tp, SC_MAX_TABLES, i, fold, r, e_lo, e_hi);
        const Fe e_step = f_sub<Fr>(e_hi, e_lo);
        Fe e = e_lo;
        for (int x = 0; x < ND; x++) {
            if (!(skip1 && x == 1)) {
                Fe g = f_zero<Fr>();
                for (int q = 0; q < gate.nterms; q++) {
                    const Term &tm = gate.terms[q];
                    Fe v = f_mul<Fr>(tm.coef, cur[tm.f[0]]);
                    if (tm.f[1] >= 0) v = f_mul<Fr>(v, cur[tm.f[1]]);
                    if (tm.f[2] >= 0) v = f_mul<Fr>(v, cur[tm.f[2]]);
                    g = f_add<Fr>(g, v);
                }
                acc[x] = f_add<Fr>(acc[x], f_mul<Fr>(e, g));
            }
            if (x + 1 < ND) {
                e = f_add<Fr>(e, e_step);
                for (int t = 0; t < k; t++) cur[t] = f_add<Fr>(cur[t], step[t]);
            }
        }
    }
    reduce_nodes<ND>(acc, partials, counter, out);
}

// eq tables: lo half by bits of x & (2^h - 1), hi half by x >> h.
__global__ void k_eq_half(const Fe *__restrict__ r, int nbits, Fe *__restrict__ out) {
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (size_t(1) << nbits)) return;
    Fe acc = f_one<Fr>();
    const Fe one = f_one<Fr>();
    for (int j = 0; j < nbits; j++) {
        const Fe f = ((x >> j) & 1) ? r[j] : f_sub<Fr>(one, r[j]);
        acc = f_mul<Fr>(acc, f);
    }
    out[x] = acc;
}

__global__ void k_eq_full(const Fe *__restrict__ lo, const Fe *__restrict__ hi, int h, size_t len,
                          Fe *__restrict__ out) {
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= len) return;
    stfe(out, x, f_mul<Fr>(ldfe(lo, x & ((size_t(1) << h) - 1)), ldfe(hi, x >> h)));
}

struct CombineArgs {
    Fe w[4][SC_MAX_TABLES];
    Fe *out[4];
};

template <int OUTS>
__global__ void k_combine(const Fe *__restrict__ cols, size_t n, int K, CombineArgs a) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fe s[OUTS];
#pragma unroll
    for (int o = 0; o < OUTS; o++) s[o] = f_zero<Fr>();
    for (int c = 0; c < K; c++) {
        const Fe v = ldfe(cols, (size_t)c * n + i);
#pragma unroll
        for (int o = 0; o < OUTS; o++) s[o] = f_add<Fr>(s[o], f_mul<Fr>(a.w[o][c], v));
    }
#pragma unroll
    for (int o = 0; o < OUTS; o++) stfe(a.out[o], i, s[o]);
}

}  // namespace

int Gate::degree() const {
    int mx = 1;
    for (int q = 0; q < nterms; q++) {
        int f = 1 + (terms[q].f[1] >= 0) + (terms[q].f[2] >= 0);
        mx = std::max(mx, f);
    }
    return std::max(3, mx + 1);
}

bool Gate::is_plain_product(int k) const {
    if (nterms != 1) return false;
    const Term &t = terms[0];
    if (!f_eq(t.coef, f_one<Fr>())) return false;
    int n = 1 + (t.f[1] >= 0) + (t.f[2] >= 0);
    if (n != k) return false;
    for (int q = 0; q < n; q++)
        if (t.f[q] != q) return false;
    return true;
}

void interpolate_host(const Fe *v, int d, Fe *c) {
    if (d == 3) {
        // crates/sumcheck/src/lib.rs:294-305
        static const Fe inv2 = f_inv<Fr>(f_from_u64<Fr>(2));
        static const Fe inv3 = f_inv<Fr>(f_from_u64<Fr>(3));
        static const Fe inv6 = f_inv<Fr>(f_from_u64<Fr>(6));
        Fe d1 = f_sub<Fr>(v[1], v[0]);
        Fe t = f_sub<Fr>(v[2], v[1]);
        Fe d2 = f_sub<Fr>(t, d1);
        Fe d3 = f_sub<Fr>(f_sub<Fr>(f_sub<Fr>(v[3], v[2]), t), d2);
        c[0] = v[0];
        c[1] = f_add<Fr>(f_sub<Fr>(d1, f_mul<Fr>(d2, inv2)), f_mul<Fr>(d3, inv3));
        c[2] = f_mul<Fr>(f_sub<Fr>(d2, d3), inv2);
        c[3] = f_mul<Fr>(d3, inv6);
        return;
    }
    // Newton forward differences on nodes 0..d, expanded into monomials.
    Fe diff[SC_MAX_NODES], delta[SC_MAX_NODES];
    for (int i = 0; i <= d; i++) diff[i] = v[i];
    for (int k = 0; k <= d; k++) {
        delta[k] = diff[0];
        for (int i = 0; i + k < d; i++) diff[i] = f_sub<Fr>(diff[i + 1], diff[i]);
    }
    Fe basis[SC_MAX_NODES], acc[SC_MAX_NODES];
    for (int i = 0; i <= d; i++) basis[i] = acc[i] = f_zero<Fr>();
    basis[0] = f_one<Fr>();
    acc[0] = delta[0];
    for (int k = 1; k <= d; k++) {
        Fe km1 = f_from_u64<Fr>(k - 1), invk = f_inv<Fr>(f_from_u64<Fr>(k));
        Fe nb[SC_MAX_NODES];
        for (int i = 0; i <= d; i++) nb[i] = f_zero<Fr>();
        for (int i = 0; i < k; i++) {
            nb[i + 1] = f_add<Fr>(nb[i + 1], basis[i]);
            nb[i] = f_sub<Fr>(nb[i], f_mul<Fr>(basis[i], km1));
        }
        for (int i = 0; i <= k; i++) basis[i] = f_mul<Fr>(nb[i], invk);
        for (int i = 0; i <= k; i++) acc[i] = f_add<Fr>(acc[i], f_mul<Fr>(basis[i], delta[k]));
    }
    for (int i = 0; i <= d; i++) c[i] = acc[i];
}

static Fe horner(const Fe *c, int d, const Fe &x) {
    Fe acc = c[d];
    for (int i = d - 1; i >= 0; i--) acc = f_add<Fr>(f_mul<Fr>(acc, x), c[i]);
    return acc;
}

void eq_table_device(Context &ctx, const Fe *r_host, int n, Fe *out_dev) {
    const int h = n / 2, hn = n - h;
    Fe *rd = ctx.d<Fe>("eq.r", std::max(n, 1));
    Fe *lo = ctx.d<Fe>("eq.lo", size_t(1) << h);
    Fe *hi = ctx.d<Fe>("eq.hi", size_t(1) << hn);
    if (n) HIP_CHECK(hipMemcpyAsync(rd, r_host, n * sizeof(Fe), hipMemcpyHostToDevice, ctx.stream));
    hipLaunchKernelGGL(k_eq_half, dim3(grid_for(size_t(1) << h, 256)), dim3(256), 0, ctx.stream, rd, h,
                       lo);
    hipLaunchKernelGGL(k_eq_half, dim3(grid_for(size_t(1) << hn, 256)), dim3(256), 0, ctx.stream,
                       rd + h, hn, hi);
    const size_t len = size_t(1) << n;
    hipLaunchKernelGGL(k_eq_full, dim3(grid_for(len, 256)), dim3(256), 0, ctx.stream, lo, hi, h, len,
                       out_dev);
    HIP_CHECK(hipGetLastError());
    // the host copy of r must outlive the async copy
    ctx.sync();
}

void combine_columns_device(Context &ctx, const Fe *cols, size_t n, int K, const Fe *w, int outs,
                            Fe *const *out) {
    if (K > SC_MAX_TABLES || outs > 4) throw ApgError(-2, "combine: too many columns/outputs");
    CombineArgs a;
    for (int o = 0; o < outs; o++) {
        for (int c = 0; c < K; c++) a.w[o][c] = w[o * K + c];
        a.out[o] = out[o];
    }
    const dim3 g(grid_for(n, 256)), b(256);
    switch (outs) {
        case 1: hipLaunchKernelGGL(k_combine<1>, g, b, 0, ctx.stream, cols, n, K, a); break;
        case 2: hipLaunchKernelGGL(k_combine<2>, g, b, 0, ctx.stream, cols, n, K, a); break;
        case 3: hipLaunchKernelGGL(k_combine<3>, g, b, 0, ctx.stream, cols, n, K, a); break;
        default: hipLaunchKernelGGL(k_combine<4>, g, b, 0, ctx.stream, cols, n, K, a); break;
    }
    HIP_CHECK(hipGetLastError());
}

template <int ND>
static void launch_round(Context &ctx, const Gate &gate, int k, bool product, const TablePtrs &tp,
                         size_t m, bool fold, const Fe &r, bool skip1, Fe *partials,
                         unsigned *counter, Fe *out) {
    const unsigned blocks = (unsigned)std::min<size_t>(MAX_RBLOCKS, std::max<size_t>(1, (m + RTPB - 1) / RTPB));
    if (product) {
        switch (k) {
#define APG_PROD(KT)                                                                            \
    case KT:                                                                                    \
        hipLaunchKernelGGL((k_round_product<KT, ND>), dim3(blocks), dim3(RTPB), 0, ctx.stream, tp, \
                           m, (int)fold, r, (int)skip1, partials, counter, out);               \
        return;
            APG_PROD(1)
            APG_PROD(2)
            APG_PROD(3)
            APG_PROD(4)
#undef APG_PROD
            default:
                break;
        }
    }
    GateArgs ga;
    ga.nterms = gate.nterms;
    for (int q = 0; q < gate.nterms; q++) ga.terms[q] = gate.terms[q];
    hipLaunchKernelGGL((k_round_gate<ND>), dim3(blocks), dim3(RTPB), 0, ctx.stream, tp, k, ga, m,
                       (int)fold, r, (int)skip1, partials, counter, out);
}

void sumcheck_prove_device(Context &ctx, const Gate &gate, Fe *const *tables, int k, int n,
                           const Fe *eq_point, const RoundFn &round_fn, Fe *rounds_out,
                           Fe *point_out, Fe *finals_out) {
    if (k < 1 || k > SC_MAX_TABLES) throw ApgError(-2, "sumcheck: table count out of range");
    if (gate.nterms < 0 || gate.nterms > SC_MAX_TERMS) throw ApgError(-2, "sumcheck: too many terms");
    for (int q = 0; q < gate.nterms; q++)
        for (int f = 0; f < 3; f++)
            if (gate.terms[q].f[f] >= k || (f == 0 && gate.terms[q].f[0] < 0))
                throw ApgError(-2, "sumcheck: term factor out of range");
    const int d = gate.degree();
    if (d + 1 > SC_MAX_NODES) throw ApgError(-2, "sumcheck: degree too high");
    const bool product = gate.is_plain_product(k);
    hipStream_t st = ctx.stream;
    const size_t len0 = size_t(1) << n;

    // eq and two ping-pong buffers per table (the inputs are read in round 0
    // and folded into scratch from round 1 on).
    Fe *eq = ctx.d<Fe>("sc.eq", std::max<size_t>(len0, 2));
    eq_table_device(ctx, eq_point, n, eq);
    std::vector<Fe *> bufA(k + 1), bufB(k + 1);
    Fe *poolA = ctx.d<Fe>("sc.poolA", (size_t)(k + 1) * std::max<size_t>(len0 / 2, 1));
    Fe *poolB = ctx.d<Fe>("sc.poolB", (size_t)(k + 1) * std::max<size_t>(len0 / 4, 1));
    for (int t = 0; t <= k; t++) {
        bufA[t] = poolA + (size_t)t * std::max<size_t>(len0 / 2, 1);
        bufB[t] = poolB + (size_t)t * std::max<size_t>(len0 / 4, 1);
    }
    std::vector<const Fe *> cur(k + 1);
    for (int t = 0; t < k; t++) cur[t] = tables[t];
    cur[k] = eq;

    Fe *partials = ctx.d<Fe>("sc.partials", (size_t)MAX_RBLOCKS * SC_MAX_NODES);
    Fe *dout = ctx.d<Fe>("sc.out", SC_MAX_NODES);
    if (!ctx.dev["sc.counter"].ptr) {
        HIP_CHECK(hipMalloc(&ctx.dev["sc.counter"].ptr, 256));
        ctx.dev["sc.counter"].cap = 256;
        HIP_CHECK(hipMemsetAsync(ctx.dev["sc.counter"].ptr, 0, 256, st));
    }
    unsigned *counter = static_cast<unsigned *>(ctx.dev["sc.counter"].ptr);
    Fe *hvals = ctx.h<Fe>("sc.hvals", SC_MAX_NODES);

    // Round 0 reads the inputs T^0 (2^n entries). Round j >= 1 folds T^(j-1)
    // with r_(j-1) into T^j (2^(n-j) entries: odd j into bufA, even j into
    // bufB, so source and destination never alias) and evaluates T^j's pairs.
    if (n == 0) {  // nothing to bind: the finals are the single entries
        for (int t = 0; t < k; t++)
            HIP_CHECK(hipMemcpyAsync(&finals_out[t], tables[t], sizeof(Fe), hipMemcpyDeviceToHost, st));
        ctx.sync();
        return;
    }
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (ctx.profiling) {
        HIP_CHECK(hipEventCreate(&ev0));
        HIP_CHECK(hipEventCreate(&ev1));
    }
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } evg{ev0, ev1};
    Fe r = f_zero<Fr>();
    Fe claim = f_zero<Fr>();
    for (int j = 0; j < n; j++) {
        const bool fold = j > 0;
        const size_t m = len0 >> (j + 1);  // pairs of T^j
        TablePtrs tp;
        std::vector<Fe *> nxt(k + 1);
        for (int t = 0; t <= k; t++) nxt[t] = (j % 2 == 1) ? bufA[t] : bufB[t];
        const int slot = SC_MAX_TABLES;
        for (int t = 0; t < k; t++) {
            tp.src[t] = cur[t];
            tp.dst[t] = nxt[t];
        }
        tp.src[slot] = cur[k];
        tp.dst[slot] = nxt[k];
        const bool skip1 = fold;
        if (ctx.profiling) HIP_CHECK(hipEventRecord(ev0, st));
        switch (d + 1) {
            case 4: launch_round<4>(ctx, gate, k, product, tp, m, fold, r, skip1, partials, counter, dout); break;
            case 5: launch_round<5>(ctx, gate, k, product, tp, m, fold, r, skip1, partials, counter, dout); break;
            case 6: launch_round<6>(ctx, gate, k, product, tp, m, fold, r, skip1, partials, counter, dout); break;
            default: throw ApgError(-2, "sumcheck: unsupported degree");
        }
        HIP_CHECK(hipGetLastError());
        if (ctx.profiling) HIP_CHECK(hipEventRecord(ev1, st));
        HIP_CHECK(hipMemcpyAsync(hvals, dout, (d + 1) * sizeof(Fe), hipMemcpyDeviceToHost, st));
        ctx.sync();
        if (ctx.profiling) {
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
            ctx.stats.sumcheck_kernel_ms += ms;
            ctx.stats.sumcheck_launches += 1;
            // reads every live table (eq included) once; a folding round also
            // writes the half-size fold
            ctx.stats.sumcheck_bytes += (uint64_t)(k + 1) * 32 * (fold ? 6 * m : 2 * m);
        }
        Fe vals[SC_MAX_NODES];
        for (int x = 0; x <= d; x++) vals[x] = hvals[x];
        if (skip1) vals[1] = f_sub<Fr>(claim, vals[0]);
        Fe *coeffs = rounds_out + (size_t)j * (d + 1);
        interpolate_host(vals, d, coeffs);
        r = round_fn(coeffs, d + 1);
        point_out[j] = r;
        claim = horner(coeffs, d, r);
        if (fold)
            for (int t = 0; t <= k; t++) cur[t] = nxt[t];
    }
    // final fold of the two-entry tables with the last challenge
    std::vector<Fe> last(2 * (k + 1));
    for (int t = 0; t < k; t++)
        HIP_CHECK(hipMemcpyAsync(&last[2 * t], cur[t], 2 * sizeof(Fe), hipMemcpyDeviceToHost, st));
    ctx.sync();
    for (int t = 0; t < k; t++) {
        const Fe lo = last[2 * t], hi = last[2 * t + 1];
        finals_out[t] = f_add<Fr>(lo, f_mul<Fr>(r, f_sub<Fr>(hi, lo)));
    }
}

}  // namespace apg
