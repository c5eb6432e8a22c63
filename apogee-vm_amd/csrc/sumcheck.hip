// Sumcheck round evaluation and fold for gfx950.
//
// Replaces the round loop of gkr::prove_sumcheck (crates/gkr/src/lib.rs:792-899)
// and sumcheck::prove_zerocheck (crates/sumcheck/src/lib.rs:359-408): per
// round j, the values of g_j(X) = sum_pairs eq_X * G(lines_X) at X = 0..d,
// then the bind f'(i) = f(2i) + r (f(2i+1) - f(2i)) (crates/poly/src/lib.rs:
// 156-185) of every table. The round polynomials are exact field elements, so
// any evaluation order gives the reference's coefficients; this one is built
// for HBM streaming:
//
//  * eq is never materialised or folded. With variable j bound by round j,
//    eq(r, (rho_0..rho_{j-1}, X, i)) = P_j * eq(r_j, X) * E_{j+1}(i), where
//    P_j = prod_{l<j} eq(r_l, rho_l) is a host scalar and E_s(i) =
//    eq(r_s..r_{n-1}, i) is a product of two small tables (the variables
//    below and above a split point h0, built once per proof), so a round reads
//    only the gate's tables. The kernel sums S_j(X) = sum_i E(i) G(lines_X(i));
//    the host multiplies by P_j eq(r_j, X).
//  * One node of S_j comes from the claim: g_j(0) + g_j(1) = g_{j-1}(rho_{j-1})
//    (the verifier's identity, crates/sumcheck/src/lib.rs:441) gives
//    P_j S_j(1) = (claim - g_j(0)) / r_j, so a gate of degree D in X needs D
//    device nodes (X = 0, 2..D) and the rest follow by extrapolation. Round 0
//    (no claim yet) and any r_j = 0 evaluate all D + 1 nodes instead.
//  * The fold of round j-1 is fused into the evaluation of round j: one pass
//    reads each table once (4 entries per pair of the folded table), writes
//    the half-size fold and accumulates the node sums.
//  * A product of two tables runs its device rounds two per launch
//    (k_round_pair): round j+1's node sums are degree-2 polynomials in the
//    challenge of round j, kept at three points and interpolated by the host
//    once the challenge is drawn; the next launch binds both challenges.
//  * Node sums reduce by wave shuffles, then across the block's waves in LDS,
//    then (more than one block) through per-block rows of partials that a
//    one-block pass after the kernel adds up (no per-block fences).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "sumcheck.h"
#include "field29.cuh"
#include "host_fr.h"
#include "sumcheck_dev.cuh"
#include "transcript_dev.cuh"
#include "transcript_dev.h"

namespace apg {

namespace {

struct TablePtrs {
    const Fe *src[SC_MAX_TABLES];
    Fe *dst[SC_MAX_TABLES];
    uint8_t kind[SC_MAX_TABLES];  // the backing of src (SC_FR unless LIFT)
};

// Entry i of a source table of the given backing (crates/poly/src/lib.rs:
// 46-97): Fr as stored; a narrow integer v lifted on read to v R mod p (its
// Montgomery image, Fr::from_u64); one bit to 0 or 1. The first round reads
// the caller's tables through this and the second folds them through it into
// Fr scratch, so no full-size lifted copy is made.
__device__ __forceinline__ Fe ld_entry(const Fe *src, uint8_t kind, size_t i) {
    switch (kind) {
        case SC_U1:
            return ((reinterpret_cast<const uint64_t *>(src)[i >> 6] >> (i & 63)) & 1) ? f_one<Fr>() : f_zero<Fr>();
        case SC_U8: return f_from_u64<Fr>(reinterpret_cast<const uint8_t *>(src)[i]);
        case SC_U16: return f_from_u64<Fr>(reinterpret_cast<const uint16_t *>(src)[i]);
        case SC_U32: return f_from_u64<Fr>(reinterpret_cast<const uint32_t *>(src)[i]);
        default: return ldfe(src, i);
    }
}

template <bool LIFT>
__device__ __forceinline__ Fe ld_src(const TablePtrs &tp, int t, size_t i) {
    if (LIFT) return ld_entry(tp.src[t], tp.kind[t], i);
    return ldfe(tp.src[t], i);
}

// A gate's terms for the kernel: factor indices, and the coefficient in the
// R' = 2^261 domain times 2^(5 (factors - 1)), so that the chain of lazy
// 29-bit multiplies (each further factor contributes 2^-5) gives the term's
// exact 32-bit image.
struct GateArgs {
    int nterms;
    int32_t f[SC_MAX_TERMS][3];
    F29 coef[SC_MAX_TERMS];
};

// Node sums of a paired-round launch (k_round_pair): two of round j, six of
// round j+1 (three challenge values y by two nodes X)
constexpr int PAIR_OUTS = 8;
// node sums a launch may write: the partials' row and the nodes buffer
constexpr int NODE_SLOTS = PAIR_OUTS > SC_MAX_NODES ? PAIR_OUTS : SC_MAX_NODES;
// the bind weights of a paired-round launch (the challenges folded at launch,
// in the R' domain) and round j's eq factors c_0 = 1 - r_(j+1), c_1 = r_(j+1)
struct PairWeights {
    F29 w[3];
    F29 c0, c1;
};

// lo/hi of pair i of table t, folding the previous round's table if asked.
template <bool LIFT>
__device__ __forceinline__ void pair_of(const TablePtrs &tp, int t, size_t i, bool fold,
                                        const Fe &r, Fe &lo, Fe &hi) {
    if (fold) {
        Fe a0 = ld_src<LIFT>(tp, t, 4 * i), a1 = ld_src<LIFT>(tp, t, 4 * i + 1), a2 = ld_src<LIFT>(tp, t, 4 * i + 2),
           a3 = ld_src<LIFT>(tp, t, 4 * i + 3);
        lo = f_add<Fr>(a0, f_mul<Fr>(r, f_sub<Fr>(a1, a0)));
        hi = f_add<Fr>(a2, f_mul<Fr>(r, f_sub<Fr>(a3, a2)));
        stfe(tp.dst[t], 2 * i, lo);
        stfe(tp.dst[t], 2 * i + 1, hi);
    } else {
        lo = ld_src<LIFT>(tp, t, 2 * i);
        hi = ld_src<LIFT>(tp, t, 2 * i + 1);
    }
}

// Product summand over KT tables (coef 1): the shards' eq * a * b (* c).
// S(X) = sum_i E(i) prod_t line_t(i, X) at X = 0..KT (FULL) or X = 0, 2..KT.
// Its multiplies are the lazy 29-bit ones of field29.cuh on 32-bit images
// (f_mul_r261: the image of x y / 2^5, ~40% fewer instructions than the
// 32-bit CIOS): the fold takes the challenge in the R' = 2^261 domain and so
// stays exact, and each node sum carries 2^(-5 KT), which the host removes.
template <int KT, bool FULL, bool LIFT>
// rho_dev (graph-captured rounds): the challenge in device memory, written by
// the previous round's transcript step, replacing r29.
__global__ void __launch_bounds__(RTPB) k_round_product(TablePtrs tp, size_t m, int fold, F29 r29_arg,
                                                      EqArgs eq, Fe *partials, Fe *out, unsigned *done,
                                                      const Fe *rho_dev) {
    raise_wave_priority();
    constexpr int NN = FULL ? KT + 1 : KT;
    const F29 r29 = rho_dev ? f29_from_mont32<Fr>(*rho_dev) : r29_arg;
    Fe acc[NN];
#pragma unroll
    for (int x = 0; x < NN; x++) acc[x] = f_zero<Fr>();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        Fe cur[KT], step[KT];
#pragma unroll
        for (int t = 0; t < KT; t++) {
            Fe lo, hi;
            if (fold) {
                // (a1 - a0) rho: the image times rho R' is the image of the product
                const Fe a0 = ld_src<LIFT>(tp, t, 4 * i), a1 = ld_src<LIFT>(tp, t, 4 * i + 1),
                         a2 = ld_src<LIFT>(tp, t, 4 * i + 2), a3 = ld_src<LIFT>(tp, t, 4 * i + 3);
                lo = f_add<Fr>(a0, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(f_sub<Fr>(a1, a0)), r29)));
                hi = f_add<Fr>(a2, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(f_sub<Fr>(a3, a2)), r29)));
                stfe(tp.dst[t], 2 * i, lo);
                stfe(tp.dst[t], 2 * i + 1, hi);
            } else {
                lo = ld_src<LIFT>(tp, t, 2 * i);
                hi = ld_src<LIFT>(tp, t, 2 * i + 1);
            }
            cur[t] = lo;
            step[t] = f_sub<Fr>(hi, lo);
        }
        // E(i): with two factor tables, their 29-bit product (one more 2^-5)
        const Fe e = eq.lo ? f_mul_r261<Fr>(ldfe(eq.lo, i & ((size_t(1) << eq.lo_bits) - 1)),
                                            ldfe(eq.hi, i >> eq.lo_bits))
                           : ldfe(eq.hi, i);
#pragma unroll
        for (int x = 0; x < NN; x++) {
            Fe g = cur[0];
#pragma unroll
            for (int t = 1; t < KT; t++) g = f_mul_r261<Fr>(g, cur[t]);
            acc[x] = f_add<Fr>(acc[x], f_mul_r261<Fr>(e, g));
            if (x + 1 < NN) {
#pragma unroll
                for (int t = 0; t < KT; t++) {
                    cur[t] = f_add<Fr>(cur[t], step[t]);
                    if (!FULL && x == 0) cur[t] = f_add<Fr>(cur[t], step[t]);  // X: 0 -> 2
                }
            }
        }
    }
    reduce_nodes<NN>(acc, NN, partials, out, done);
}

// Entry e of a table bound by two more variables, from its four entries
// a_0..a_3 at 4e.. of the table they were stored in:
//   a_0 + ra (a_1 - a_0) + rb (a_2 - a_0) + ra rb (a_3 - a_2 - a_1 + a_0),
// the binds by ra and then rb as one dot product with one reduction.
__device__ __forceinline__ Fe fold_twice(const Fe *src, size_t e, const F29 (&w)[3]) {
    const Fe a0 = ldfe(src, 4 * e), a1 = ldfe(src, 4 * e + 1), a2 = ldfe(src, 4 * e + 2), a3 = ldfe(src, 4 * e + 3);
    const Fe d1 = f_sub<Fr>(a1, a0);
    const F29 d[3] = {f29_from32(d1), f29_from32(f_sub<Fr>(a2, a0)), f29_from32(f_sub<Fr>(f_sub<Fr>(a3, a2), d1))};
    return f_add<Fr>(a0, f29_pack_small<Fr>(f29_dot<Fr, 3>(d, w)));
}

// Rounds j and j+1 of the product summand of two tables in one pass, both
// with node 1 from the claim (r_j and r_(j+1) nonzero). Thread p takes pair p
// of T^j: its entries folded from the stored table T^(j-FD) with the FD
// challenges known at launch (w: rho_(j-1) for FD = 1; rho_(j-2), rho_(j-1)
// and their product for FD = 2), written back as the next pass's input. It adds
//  * round j at X = 0, 2: E_j(p) line_0(p, X) line_1(p, X), with E_j(2i + b) =
//    E_(j+1)(i) c_b, c_0 = 1 - r_(j+1), c_1 = r_(j+1);
//  * round j+1 as a polynomial in its unknown challenge rho: pair i of T^(j+1)
//    is (line(2i, rho), line(2i+1, rho)), so E_(j+1)(i) prod_t (lo_t + X (hi_t
//    - lo_t)) has degree 2 in rho and is kept at rho = y = 0, 1, 2; the even
//    lane of the pair (lo = its own line) takes X = 0 and the odd lane X = 2
//    (2 hi - lo, with lo from the even lane).
// out[0..1]: round j at X = 0, 2; out[2 + 2y + X/2]: round j+1's sums at
// (rho = y, X). The host takes round j+1's node sums at rho = rho_j by
// interpolation over y. Products as in k_round_product (each node sum carries
// 2^-10, and 2^-5 more when E_(j+1) = lo hi).
template <int FD>
__global__ void __launch_bounds__(RTPB) k_round_pair(TablePtrs tp, size_t m, PairWeights pw, EqArgs eq,
                                                   Fe *partials, Fe *out, unsigned *done) {
    raise_wave_priority();
    constexpr int KT = 2;
    Fe accj[2], accl[3];
#pragma unroll
    for (int x = 0; x < 2; x++) accj[x] = f_zero<Fr>();
#pragma unroll
    for (int y = 0; y < 3; y++) accl[y] = f_zero<Fr>();
    const bool odd = threadIdx.x & 1;
    F29 cb;  // c_b of this lane's pairs (p has the parity of the thread)
#pragma unroll
    for (int q = 0; q < 9; q++) cb.v[q] = odd ? pw.c1.v[q] : pw.c0.v[q];
    const size_t stride = (size_t)gridDim.x * blockDim.x;  // even: pairs p and p ^ 1 in lanes 2i, 2i+1
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < m; p += stride) {
        Fe l[3][KT];  // line_t(p, y) at y = 0, 1, 2
#pragma unroll
        for (int t = 0; t < KT; t++) {
            Fe lo, hi;
            if (FD == 1) {
                const Fe *s = tp.src[t];
                const Fe a0 = ldfe(s, 4 * p), a1 = ldfe(s, 4 * p + 1), a2 = ldfe(s, 4 * p + 2),
                         a3 = ldfe(s, 4 * p + 3);
                lo = f_add<Fr>(a0, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(f_sub<Fr>(a1, a0)), pw.w[0])));
                hi = f_add<Fr>(a2, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(f_sub<Fr>(a3, a2)), pw.w[0])));
            } else {
                lo = fold_twice(tp.src[t], 2 * p, pw.w);
                hi = fold_twice(tp.src[t], 2 * p + 1, pw.w);
            }
            stfe(tp.dst[t], 2 * p, lo);
            stfe(tp.dst[t], 2 * p + 1, hi);
            l[0][t] = lo;
            l[1][t] = hi;
            l[2][t] = f_add<Fr>(hi, f_sub<Fr>(hi, lo));
        }
        const size_t i = p >> 1;
        const Fe e1 = eq.lo ? f_mul_r261<Fr>(ldfe(eq.lo, i & ((size_t(1) << eq.lo_bits) - 1)),
                                             ldfe(eq.hi, i >> eq.lo_bits))
                            : ldfe(eq.hi, i);
        const Fe ej = f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(e1), cb));  // exactly E_(j+1)(i) c_b
        accj[0] = f_add<Fr>(accj[0], f_mul_r261<Fr>(ej, f_mul_r261<Fr>(l[0][0], l[0][1])));
        accj[1] = f_add<Fr>(accj[1], f_mul_r261<Fr>(ej, f_mul_r261<Fr>(l[2][0], l[2][1])));
#pragma unroll
        for (int y = 0; y < 3; y++) {
            Fe v[KT];
#pragma unroll
            for (int t = 0; t < KT; t++) {
                Fe o;
#pragma unroll
                for (int q = 0; q < 8; q++) o.l[q] = __shfl_xor(l[y][t].l[q], 1, 64);
                // even lane: lo; odd lane: hi + (hi - lo)
                v[t] = odd ? f_add<Fr>(l[y][t], f_sub<Fr>(l[y][t], o)) : l[y][t];
            }
            accl[y] = f_add<Fr>(accl[y], f_mul_r261<Fr>(e1, f_mul_r261<Fr>(v[0], v[1])));
        }
    }
    __shared__ Fe lds[RWAVES][PAIR_OUTS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int x = 0; x < 2; x++) {
        const Fe s = wave_sum(accj[x]);
        if (lane == 0) lds[w][x] = s;
    }
#pragma unroll
    for (int y = 0; y < 3; y++) {
        const Fe s = wave_sum_parity(accl[y]);
        if (lane < 2) lds[w][2 + 2 * y + lane] = s;
    }
    __syncthreads();
    finish_nodes<PAIR_OUTS>(lds, PAIR_OUTS, partials, out, done);
}

// Any gate: terms read from the kernel argument, table lines kept in a local
// array (dynamic factor indices). nn of NN nodes: X = 0..nn-1 (full) or
// X = 0, 2..nn.
template <int NN, bool LIFT>
__global__ void __launch_bounds__(RTPB) k_round_gate(TablePtrs tp, int k, GateArgs gate, size_t m,
                                                   int fold, Fe r, EqArgs eq, int full, int nn,
                                                   Fe *partials, Fe *out, unsigned *done) {
    raise_wave_priority();
    Fe acc[NN];
#pragma unroll
    for (int x = 0; x < NN; x++) acc[x] = f_zero<Fr>();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        Fe cur[SC_MAX_TABLES], step[SC_MAX_TABLES];
        for (int t = 0; t < k; t++) {
            Fe lo, hi;
            pair_of<LIFT>(tp, t, i, fold, r, lo, hi);
            cur[t] = lo;
            step[t] = f_sub<Fr>(hi, lo);
        }
        const Fe e = eq_factor(eq, i);
#pragma unroll
        for (int x = 0; x < NN; x++) {
            if (x < nn) {
                Fe g = f_zero<Fr>();
                for (int q = 0; q < gate.nterms; q++) {
                    Fe v = f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(cur[gate.f[q][0]]), gate.coef[q]));
                    if (gate.f[q][1] >= 0) v = f_mul_r261<Fr>(v, cur[gate.f[q][1]]);
                    if (gate.f[q][2] >= 0) v = f_mul_r261<Fr>(v, cur[gate.f[q][2]]);
                    g = f_add<Fr>(g, v);
                }
                acc[x] = f_add<Fr>(acc[x], f_mul<Fr>(e, g));
                if (x + 1 < nn) {
                    for (int t = 0; t < k; t++) {
                        cur[t] = f_add<Fr>(cur[t], step[t]);
                        if (!full && x == 0) cur[t] = f_add<Fr>(cur[t], step[t]);
                    }
                }
            }
        }
    }
    reduce_nodes<NN>(acc, nn, partials, out, done);
}

// out_o = scale_o * sum_c w_c col_c, o < OUTS (the shard's tables are
// rho^(o K) * sum_c rho^c col_c). Weights and scales are in the Fq-style lazy
// 29-bit form over Fr, in the domain R' = 2^261 (field29.cuh): a column value
// c R (its 32-bit Montgomery image, re-sliced) times w R' gives c w R. The
// sum takes up to six columns per reduction (f29_dot: ~90 reduction
// multiply-adds per six columns instead of per column), each group below
// 1.04 r; at most four groups (24 columns) stay below 4.2 r, and each scale
// multiply brings that below 1.03 r before the repack.
struct CombineArgs {
    F29 w[SC_MAX_TABLES];
    F29 scale[4];
    Fe *out[4];
};

template <int G>
__device__ __forceinline__ F29 combine_group(const Fe *__restrict__ cols, size_t n, size_t i, int c0,
                                             const CombineArgs &a) {
    F29 v[G];
#pragma unroll
    for (int g = 0; g < G; g++) v[g] = f29_from32(ldfe(cols, (size_t)(c0 + g) * n + i));
    return f29_dot<Fr, G>(v, &a.w[c0]);
}

template <int OUTS>
__global__ void __launch_bounds__(256) k_combine(const Fe *__restrict__ cols, size_t n, int K, CombineArgs a) {
    raise_wave_priority();
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F29 s{};
    int c = 0;
    for (; c + 6 <= K; c += 6) s = f29_add(s, combine_group<6>(cols, n, i, c, a));
    switch (K - c) {
        case 1: s = f29_add(s, combine_group<1>(cols, n, i, c, a)); break;
        case 2: s = f29_add(s, combine_group<2>(cols, n, i, c, a)); break;
        case 3: s = f29_add(s, combine_group<3>(cols, n, i, c, a)); break;
        case 4: s = f29_add(s, combine_group<4>(cols, n, i, c, a)); break;
        case 5: s = f29_add(s, combine_group<5>(cols, n, i, c, a)); break;
        default: break;
    }
    stfe(a.out[0], i, f29_pack_reduced<Fr>(s));
#pragma unroll
    for (int o = 1; o < OUTS; o++) stfe(a.out[o], i, f29_pack_reduced<Fr>(f29_mul<Fr>(s, a.scale[o])));
}

}  // namespace

int Gate::x_degree() const {
    int mx = 1;
    for (int q = 0; q < nterms; q++) {
        int f = 1 + (terms[q].f[1] >= 0) + (terms[q].f[2] >= 0);
        mx = std::max(mx, f);
    }
    return mx;
}

int Gate::degree() const { return std::max(3, x_degree() + 1); }

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

// Host arithmetic of a round (the node sums to the round polynomial, the
// claim and the running eq product) on four 64-bit words (host_fr.h). The
// constants 1..SC_MAX_NODES and their inverses are computed once.
namespace {
using hfr::W4;
struct SmallConsts {
    W4 num[SC_MAX_NODES + 1], inv[SC_MAX_NODES + 1];
    W4 one, zero, inv2, inv3, inv6;
    SmallConsts() {
        for (int x = 0; x <= SC_MAX_NODES; x++) {
            num[x] = hfr::w4_of(f_from_u64<Fr>((uint64_t)x));
            inv[x] = x ? hfr::w4_of(f_inv<Fr>(f_from_u64<Fr>((uint64_t)x))) : num[0];
        }
        zero = num[0];
        one = num[1];
        inv2 = inv[2];
        inv3 = inv[3];
        inv6 = hfr::w4_of(f_inv<Fr>(f_from_u64<Fr>(6)));
    }
};
const SmallConsts &smallc() {
    static const SmallConsts c;
    return c;
}

// Coefficients (ascending) of the degree-d polynomial through (x, v[x]),
// x = 0..d.
void interpolate_w4(const W4 *v, int d, W4 *c) {
    using namespace hfr;
    const SmallConsts &K = smallc();
    if (d == 3) {
        // crates/sumcheck/src/lib.rs:294-305
        const W4 d1 = w4_sub(v[1], v[0]);
        const W4 t = w4_sub(v[2], v[1]);
        const W4 d2 = w4_sub(t, d1);
        const W4 d3 = w4_sub(w4_sub(w4_sub(v[3], v[2]), t), d2);
        c[0] = v[0];
        c[1] = w4_add(w4_sub(d1, w4_mul(d2, K.inv2)), w4_mul(d3, K.inv3));
        c[2] = w4_mul(w4_sub(d2, d3), K.inv2);
        c[3] = w4_mul(d3, K.inv6);
        return;
    }
    // Newton forward differences on nodes 0..d, expanded into monomials:
    // basis_k(X) = X (X - 1) ... (X - k + 1) / k!
    W4 diff[SC_MAX_NODES], delta[SC_MAX_NODES];
    for (int i = 0; i <= d; i++) diff[i] = v[i];
    for (int k = 0; k <= d; k++) {
        delta[k] = diff[0];
        for (int i = 0; i + k < d; i++) diff[i] = w4_sub(diff[i + 1], diff[i]);
    }
    W4 basis[SC_MAX_NODES], acc[SC_MAX_NODES];
    for (int i = 0; i <= d; i++) basis[i] = acc[i] = K.zero;
    basis[0] = K.one;
    acc[0] = delta[0];
    for (int k = 1; k <= d; k++) {
        W4 nb[SC_MAX_NODES];
        for (int i = 0; i <= d; i++) nb[i] = K.zero;
        for (int i = 0; i < k; i++) {
            nb[i + 1] = w4_add(nb[i + 1], basis[i]);
            nb[i] = w4_sub(nb[i], w4_mul(basis[i], K.num[k - 1]));
        }
        for (int i = 0; i <= k; i++) basis[i] = w4_mul(nb[i], K.inv[k]);
        for (int i = 0; i <= k; i++) acc[i] = w4_add(acc[i], w4_mul(basis[i], delta[k]));
    }
    for (int i = 0; i <= d; i++) c[i] = acc[i];
}

W4 horner_w4(const W4 *c, int d, const W4 &x) {
    W4 acc = c[d];
    for (int i = d - 1; i >= 0; i--) acc = hfr::w4_add(hfr::w4_mul(acc, x), c[i]);
    return acc;
}

// Values of a degree-D polynomial at 0..D extended to 0..d (d >= D) by
// forward differences (the D-th difference is constant).
void extrapolate_w4(W4 *v, int D, int d) {
    W4 last[SC_MAX_NODES];  // last[k] = k-th difference ending at v[D]
    W4 row[SC_MAX_NODES];
    for (int i = 0; i <= D; i++) row[i] = v[i];
    for (int k = 0; k <= D; k++) {
        last[k] = row[D - k];
        for (int i = 0; i + k < D; i++) row[i] = hfr::w4_sub(row[i + 1], row[i]);
    }
    for (int x = D + 1; x <= d; x++) {
        for (int k = D - 1; k >= 0; k--) last[k] = hfr::w4_add(last[k], last[k + 1]);
        v[x] = last[0];
    }
}

// The one-variable factor of eq, eq(r, x) = (1 - r) + x (2r - 1): its base
// and slope.
struct EqLine {
    W4 base, slope;
    explicit EqLine(const Fe &r) {
        using namespace hfr;
        const W4 rw = w4_of(r), one = smallc().one;
        base = w4_sub(one, rw);
        slope = w4_sub(w4_add(rw, rw), one);
    }
    W4 at(const W4 &x) const { return hfr::w4_add(base, hfr::w4_mul(x, slope)); }
};

// g_j from T(X) = P S(X) at X = 0..D (T[1] may come from the claim): the
// values eq(r_j, X) T(X) at X = 0..d, then the coefficients. With from_claim,
// g(1) = claim - g(0).
void round_poly(W4 *T, int D, int d, const EqLine &eq, bool from_claim, const W4 &claim, W4 *coeffs) {
    using namespace hfr;
    extrapolate_w4(T, D, d);
    W4 vals[SC_MAX_NODES];
    W4 e = eq.base;  // eq at X = 0, then one slope per step
    for (int x = 0; x <= d; x++) {
        vals[x] = w4_mul(e, T[x]);
        e = w4_add(e, eq.slope);
    }
    if (from_claim) vals[1] = w4_sub(claim, vals[0]);
    interpolate_w4(vals, d, coeffs);
}
}  // namespace

void interpolate_host(const Fe *v, int d, Fe *c) {
    W4 vw[SC_MAX_NODES], cw[SC_MAX_NODES];
    for (int i = 0; i <= d; i++) vw[i] = hfr::w4_of(v[i]);
    interpolate_w4(vw, d, cw);
    for (int i = 0; i <= d; i++) c[i] = hfr::fe_of(cw[i]);
}

void combine_columns_device(Context &ctx, const Fe *cols, size_t n, int K, const Fe *w, const Fe *scale,
                            int outs, Fe *const *out) {
    if (K < 1 || K > SC_MAX_TABLES || outs < 1 || outs > 4)
        throw ApgError(-2, "combine: too many columns/outputs");
    CombineArgs a;
    for (int c = 0; c < K; c++) a.w[c] = f29_from_mont32<Fr>(w[c]);
    for (int o = 0; o < outs; o++) {
        a.scale[o] = f29_from_mont32<Fr>(scale[o]);
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

// Blocks for m pairs: one pair per thread up to per_cu blocks per CU, beyond
// that a grid-stride loop. The product rounds take two blocks per CU: with
// the separate finishing pass (no per-block fences) the 2^22 x 2 sumcheck
// measured 1.26, 1.16 and 1.18 ms at 1, 2 and 4 blocks per CU, 2^20 x 2
// alike at all three. The heavier gate rounds keep 4, as the GKR layer's
// rounds, which lose 16% at one.
constexpr int PRODUCT_BLOCKS_PER_CU = 2;
static unsigned round_blocks(const Context &ctx, size_t m, int per_cu) {
    const size_t cap = std::min<size_t>(MAX_RBLOCKS, (size_t)per_cu * ctx.cu_count);
    return (unsigned)std::min(cap, std::max<size_t>(1, (m + RTPB - 1) / RTPB));
}

template <int KT, bool LIFT>
static void launch_product(Context &ctx, bool full, const TablePtrs &tp, size_t m, bool fold,
                           const Fe &r, const EqArgs &eq, Fe *partials, Fe *out, unsigned *done) {
    const dim3 g(round_blocks(ctx, m, PRODUCT_BLOCKS_PER_CU)), b(RTPB);
    const F29 r29 = f29_from_mont32<Fr>(r);  // rho R' mod r
    if (full) {
        hipLaunchKernelGGL((k_round_product<KT, true, LIFT>), g, b, 0, ctx.stream, tp, m, (int)fold, r29, eq,
                           partials, out, done, (const Fe *)nullptr);
        launch_finish<KT + 1>(ctx.stream, g.x, KT + 1, partials, out, done);
    } else {
        hipLaunchKernelGGL((k_round_product<KT, false, LIFT>), g, b, 0, ctx.stream, tp, m, (int)fold, r29, eq,
                           partials, out, done, (const Fe *)nullptr);
        launch_finish<KT>(ctx.stream, g.x, KT, partials, out, done);
    }
}

template <int KT>
static void launch_product(Context &ctx, bool lift, bool full, const TablePtrs &tp, size_t m, bool fold,
                           const Fe &r, const EqArgs &eq, Fe *partials, Fe *out, unsigned *done) {
    if (lift)
        launch_product<KT, true>(ctx, full, tp, m, fold, r, eq, partials, out, done);
    else
        launch_product<KT, false>(ctx, full, tp, m, fold, r, eq, partials, out, done);
}

// A paired-round launch (k_round_pair) over m pairs of T^j, folding with the
// fd challenges in pw.
static void launch_pair(Context &ctx, int fd, const TablePtrs &tp, size_t m, const PairWeights &pw,
                        const EqArgs &eq, Fe *partials, Fe *out, unsigned *done) {
    const dim3 g(round_blocks(ctx, m, PRODUCT_BLOCKS_PER_CU)), b(RTPB);
    if (fd == 1)
        hipLaunchKernelGGL(k_round_pair<1>, g, b, 0, ctx.stream, tp, m, pw, eq, partials, out, done);
    else
        hipLaunchKernelGGL(k_round_pair<2>, g, b, 0, ctx.stream, tp, m, pw, eq, partials, out, done);
    launch_finish<PAIR_OUTS>(ctx.stream, g.x, PAIR_OUTS, partials, out, done);
    HIP_CHECK(hipGetLastError());
}

// lift: some source table is a narrow backing (the first two rounds); done:
// the completion flag the round sets after its node sums (or null)
static void launch_round(Context &ctx, const Gate &gate, int k, bool product, bool full, int D, bool lift,
                         const TablePtrs &tp, size_t m, bool fold, const Fe &r, const EqArgs &eq,
                         Fe *partials, Fe *out, unsigned *done) {
    if (product) {
        switch (k) {
            case 1: launch_product<1>(ctx, lift, full, tp, m, fold, r, eq, partials, out, done); return;
            case 2: launch_product<2>(ctx, lift, full, tp, m, fold, r, eq, partials, out, done); return;
            case 3: launch_product<3>(ctx, lift, full, tp, m, fold, r, eq, partials, out, done); return;
            default: break;
        }
    }
    GateArgs ga;
    ga.nterms = gate.nterms;
    for (int q = 0; q < gate.nterms; q++) {
        const Term &tm = gate.terms[q];
        Fe c = tm.coef;
        for (int f = 1; f < 3; f++)
            if (tm.f[f] >= 0)
                for (int d = 0; d < 5; d++) c = f_add<Fr>(c, c);  // times 2^5 per further factor
        ga.coef[q] = f29_from_mont32<Fr>(c);
        for (int f = 0; f < 3; f++) ga.f[q][f] = tm.f[f];
    }
    const int nn = full ? D + 1 : D;
    const dim3 g(round_blocks(ctx, m, 4)), b(RTPB);
#define APG_GATE(NN)                                                                                          \
    do {                                                                                                      \
        if (lift)                                                                                             \
            hipLaunchKernelGGL((k_round_gate<NN, true>), g, b, 0, ctx.stream, tp, k, ga, m, (int)fold, r, eq, \
                               (int)full, nn, partials, out, done);                                           \
        else                                                                                                  \
            hipLaunchKernelGGL((k_round_gate<NN, false>), g, b, 0, ctx.stream, tp, k, ga, m, (int)fold, r,    \
                               eq, (int)full, nn, partials, out, done);                                       \
        launch_finish<NN>(ctx.stream, g.x, nn, partials, out, done);                                          \
    } while (0)
    switch (D + 1) {
        case 2: APG_GATE(2); break;
        case 3: APG_GATE(3); break;
        default: APG_GATE(4); break;
    }
#undef APG_GATE
}

// Device scratch of one proof, carved from one region in a fixed order (each
// piece 256-byte aligned): the eq point, the suffix-eq tables, two ping-pong
// fold pools, the round kernels' per-block partial sums, and the node sums
// of a round (read back after each round).
struct SumcheckScratch {
    Fe *r, *eqA, *eqB, *poolA, *poolB, *partials, *nodes;
    // completion flag of a round (pinned, coherent; its device mapping), or null
    unsigned *done_host = nullptr, *done_dev = nullptr;
};

// Waits until the round kernel in flight has set its completion flag. The
// stream is queried now and then, so a kernel that failed (or ended without
// setting the flag) raises an error instead of a wait without end.
static void wait_round(Context &ctx, volatile unsigned *done) {
    for (uint32_t spin = 1;; spin++) {
        if (__atomic_load_n(done, __ATOMIC_ACQUIRE) != 0) return;
        if ((spin & 255) == 0) {
            const hipError_t e = hipStreamQuery(ctx.stream);
            if (e == hipSuccess) {
                if (__atomic_load_n(done, __ATOMIC_ACQUIRE) != 0) return;
                throw ApgError(-5, "sumcheck: a round kernel ended without its completion flag");
            }
            if (e != hipErrorNotReady) HIP_CHECK(e);
        }
#if defined(__x86_64__) || defined(__i386__)
        __builtin_ia32_pause();
#endif
    }
}

struct ScratchLayout {
    size_t off[7];
    size_t total;
};

static ScratchLayout sumcheck_layout(int k, int n) {
    const size_t len0 = size_t(1) << n;
    const int h0 = std::max(1, std::min(n, (n + 1) / 2));
    const size_t astride = size_t(1) << (h0 - 1), bstride = size_t(1) << (n - h0);
    const size_t asz = (size_t)(h0 - 1) * astride, bsz = (size_t)(n - h0 + 1) * bstride;
    const size_t bytes[7] = {
        (size_t)std::max(n, 1) * sizeof(Fe),
        std::max<size_t>(asz, 1) * sizeof(Fe),
        std::max<size_t>(bsz, 1) * sizeof(Fe),
        (size_t)k * std::max<size_t>(len0 / 2, 1) * sizeof(Fe),
        (size_t)k * std::max<size_t>(len0 / 4, 1) * sizeof(Fe),
        (size_t)MAX_RBLOCKS * NODE_SLOTS * sizeof(Fe),
        NODE_SLOTS * sizeof(Fe),
    };
    ScratchLayout l;
    size_t at = 0;
    for (int i = 0; i < 7; i++) {
        l.off[i] = at;
        at += (bytes[i] + 255) & ~size_t(255);
    }
    l.total = at;
    return l;
}

static SumcheckScratch sumcheck_carve(uint8_t *base, const ScratchLayout &l) {
    SumcheckScratch s;
    s.r = reinterpret_cast<Fe *>(base + l.off[0]);
    s.eqA = reinterpret_cast<Fe *>(base + l.off[1]);
    s.eqB = reinterpret_cast<Fe *>(base + l.off[2]);
    s.poolA = reinterpret_cast<Fe *>(base + l.off[3]);
    s.poolB = reinterpret_cast<Fe *>(base + l.off[4]);
    s.partials = reinterpret_cast<Fe *>(base + l.off[5]);
    s.nodes = reinterpret_cast<Fe *>(base + l.off[6]);
    return s;
}

static void check_sumcheck_args(const Gate &gate, int k) {
    if (k < 1 || k > SC_MAX_TABLES) throw ApgError(-2, "sumcheck: table count out of range");
    if (gate.nterms < 0 || gate.nterms > SC_MAX_TERMS) throw ApgError(-2, "sumcheck: too many terms");
    for (int q = 0; q < gate.nterms; q++)
        for (int f = 0; f < 3; f++)
            if (gate.terms[q].f[f] >= k || (f == 0 && gate.terms[q].f[0] < 0))
                throw ApgError(-2, "sumcheck: term factor out of range");
    if (gate.degree() + 1 > SC_MAX_NODES || gate.x_degree() > 3) throw ApgError(-2, "sumcheck: degree too high");
}

// The round loop over given scratch. mapped_nodes: the round kernels write
// their node sums to `nodes_host` through its device mapping `sc.nodes`
// (pinned memory, visible after the stream sync); otherwise they write
// sc.nodes in device memory and each round copies them to nodes_host.
static void sumcheck_rounds(Context &ctx, const SumcheckScratch &sc, Fe *nodes_host, bool mapped_nodes,
                            bool timing, const Gate &gate, Fe *const *tables, const uint8_t *kinds, int k, int n,
                            const Fe *eq_point, const RoundFn &round_fn, Fe *rounds_out, Fe *point_out,
                            Fe *finals_out);

// The single entry of each table (no variables to bind), lifted.
static void sumcheck_finals_n0(Context &ctx, Fe *const *tables, const uint8_t *kinds, int k, Fe *finals_out);

void sumcheck_prove_device(Context &ctx, const Gate &gate, Fe *const *tables, int k, int n,
                           const Fe *eq_point, const RoundFn &round_fn, Fe *rounds_out,
                           Fe *point_out, Fe *finals_out, const uint8_t *kinds) {
    check_sumcheck_args(gate, k);
    if (n == 0) {
        sumcheck_finals_n0(ctx, tables, kinds, k, finals_out);
        return;
    }
    const ScratchLayout l = sumcheck_layout(k, n);
    SumcheckScratch sc = sumcheck_carve(ctx.d<uint8_t>("sc.scratch", l.total), l);
    // node sums and the round's completion flag in coherent pinned memory:
    // the host waits on the flag, not on the stream
    Fe *hvals = ctx.hc<Fe>("sc.hvals", NODE_SLOTS + 1);
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&sc.nodes), hvals, 0));
    sc.done_host = reinterpret_cast<unsigned *>(hvals + NODE_SLOTS);
    sc.done_dev = reinterpret_cast<unsigned *>(sc.nodes + NODE_SLOTS);
    sumcheck_rounds(ctx, sc, hvals, true, ctx.profiling, gate, tables, kinds, k, n, eq_point, round_fn, rounds_out,
                    point_out, finals_out);
}

namespace {
// Scratch of apg_sumcheck_prove_dev_noalloc: one array per device, held
// until the process ends; its device address is taken once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu[NOALLOC_MAX_DEVICES];
void *g_noalloc_ptr[NOALLOC_MAX_DEVICES];
}  // namespace

size_t sumcheck_noalloc_limit() { return NOALLOC_SCRATCH_BYTES; }

size_t sumcheck_noalloc_bytes(int k, int n) { return n == 0 ? 0 : sumcheck_layout(k, n).total; }

void sumcheck_noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw ApgError(-2, "no-allocation scratch: device index out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu[device]);
    if (g_noalloc_ptr[device]) return;
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_noalloc_ptr[device] = p;
}

void sumcheck_prove_device_noalloc(Context &ctx, const Gate &gate, Fe *const *tables, int k, int n,
                                   const Fe *eq_point, const RoundFn &round_fn, Fe *rounds_out, Fe *point_out,
                                   Fe *finals_out, const uint8_t *kinds) {
    check_sumcheck_args(gate, k);
    if (ctx.device < 0 || ctx.device >= NOALLOC_MAX_DEVICES || !g_noalloc_ptr[ctx.device])
        throw ApgError(-2, "no-allocation scratch: not set up on this device");
    if (n == 0) {
        sumcheck_finals_n0(ctx, tables, kinds, k, finals_out);
        return;
    }
    const ScratchLayout l = sumcheck_layout(k, n);
    if (l.total > NOALLOC_SCRATCH_BYTES)
        throw ApgError(-2, "sumcheck (no allocation): " + std::to_string(k) + " tables of 2^" + std::to_string(n) +
                               " entries need " + std::to_string(l.total) + " bytes of scratch, above the " +
                               std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte limit");
    // one call at a time per device: the calls share the device's array
    std::lock_guard<std::mutex> lk(g_noalloc_mu[ctx.device]);
    // the array keeps the last call's contents; every round writes what it
    // reads (partial rows, node sums) before reading it
    const SumcheckScratch sc = sumcheck_carve(static_cast<uint8_t *>(g_noalloc_ptr[ctx.device]), l);
    Fe nodes_host[NODE_SLOTS];
    sumcheck_rounds(ctx, sc, nodes_host, false, false, gate, tables, kinds, k, n, eq_point, round_fn, rounds_out,
                    point_out, finals_out);
}

// Rounds j0.. on the host. A device round over small tables is bound by
// latency (one block, a launch and a synchronisation: about 20 us), while the
// host evaluates a round of m pairs with about ten multiplies per pair at
// ~30 ns each, so once the k tables hold at most TAIL_ELEMS entries in all
// they are read back once and the remaining rounds run here: the node sums
// S(X) at X = 0..D directly (no node from the claim), then the same
// extrapolation, interpolation and transcript step as a device round, then
// the fold. The tables read are T^(j0-1) (len = 2^(n-j0+1) entries, folded
// here with rho_(j0-1) first) or, when j0 == 0, the inputs.
constexpr size_t TAIL_ELEMS = 256;

// Bytes a table of `len` entries of backing `kind` occupies.
static size_t backing_bytes(uint8_t kind, size_t len) {
    switch (kind) {
        case SC_U1: return (len + 63) / 64 * 8;
        case SC_U8: return len;
        case SC_U16: return 2 * len;
        case SC_U32: return 4 * len;
        default: return len * sizeof(Fe);
    }
}

// Entry i of a host copy of a table of backing `kind`, lifted (as ld_entry).
static Fe host_entry(const uint8_t *b, uint8_t kind, size_t i) {
    switch (kind) {
        case SC_U1: {
            uint64_t w;
            std::memcpy(&w, b + 8 * (i >> 6), 8);
            return ((w >> (i & 63)) & 1) ? f_one<Fr>() : f_zero<Fr>();
        }
        case SC_U8: return f_from_u64<Fr>(b[i]);
        case SC_U16: {
            uint16_t v;
            std::memcpy(&v, b + 2 * i, 2);
            return f_from_u64<Fr>(v);
        }
        case SC_U32: {
            uint32_t v;
            std::memcpy(&v, b + 4 * i, 4);
            return f_from_u64<Fr>(v);
        }
        default: {
            Fe v;
            std::memcpy(&v, b + sizeof(Fe) * i, sizeof(Fe));
            return v;
        }
    }
}

static void sumcheck_finals_n0(Context &ctx, Fe *const *tables, const uint8_t *kinds, int k, Fe *finals_out) {
    uint8_t one[SC_MAX_TABLES][sizeof(Fe)];
    for (int t = 0; t < k; t++)
        HIP_CHECK(hipMemcpyAsync(one[t], tables[t], backing_bytes(kinds ? kinds[t] : SC_FR, 1), hipMemcpyDeviceToHost,
                                 ctx.stream));
    ctx.sync();
    for (int t = 0; t < k; t++) finals_out[t] = host_entry(one[t], kinds ? kinds[t] : SC_FR, 0);
}

static void sumcheck_host_tail(Context &ctx, const Gate &gate, int k, int n, int j0, const Fe *eq_point,
                               const Fe *const *dev_tables, const uint8_t *kinds, size_t len, const Fe *pending,
                               int npending, const Fe &P_in, const RoundFn &round_fn, Fe *rounds_out,
                               Fe *point_out, Fe *finals_out) {
    using namespace hfr;
    W4 P = w4_of(P_in);
    const int D = gate.x_degree(), d = gate.degree();
    std::vector<W4> tab((size_t)k * len);
    std::vector<std::vector<uint8_t>> narrow(k);
    for (int t = 0; t < k; t++) {
        const uint8_t kind = kinds ? kinds[t] : SC_FR;
        if (kind == SC_FR) {
            HIP_CHECK(hipMemcpyAsync(&tab[(size_t)t * len], dev_tables[t], len * sizeof(Fe), hipMemcpyDeviceToHost,
                                     ctx.stream));
        } else {
            narrow[t].resize(backing_bytes(kind, len));
            HIP_CHECK(hipMemcpyAsync(narrow[t].data(), dev_tables[t], narrow[t].size(), hipMemcpyDeviceToHost,
                                     ctx.stream));
        }
    }
    ctx.sync();
    for (int t = 0; t < k; t++)  // narrow tables lifted on the host, as ld_entry lifts them
        if (!narrow[t].empty())
            for (size_t i = 0; i < len; i++) tab[(size_t)t * len + i] = w4_of(host_entry(narrow[t].data(), kinds[t], i));
    size_t L = len;
    auto fold = [&](const W4 &r) {  // f'(i) = f(2i) + r (f(2i+1) - f(2i)), crates/poly/src/lib.rs:156-185
        for (int t = 0; t < k; t++) {
            W4 *f = &tab[(size_t)t * len];
            for (size_t i = 0; i < L / 2; i++) f[i] = w4_add(f[2 * i], w4_mul(r, w4_sub(f[2 * i + 1], f[2 * i])));
        }
        L /= 2;
    };
    for (int q = 0; q < npending; q++) fold(w4_of(pending[q]));  // challenges drawn since the tables were stored
    std::vector<W4> coef(gate.nterms);
    for (int q = 0; q < gate.nterms; q++) coef[q] = w4_of(gate.terms[q].coef);
    const W4 one = w4_of(f_one<Fr>()), zero = w4_of(f_zero<Fr>());
    std::vector<W4> E;
    for (int j = j0; j < n; j++) {
        const size_t m = L / 2;
        // E_(j+1)(i) = eq(r_(j+1).., i) by doubling (crates/poly/src/lib.rs:233-248)
        E.assign(m, zero);
        E[0] = one;
        for (size_t h = 1, b = 0; h < m; h *= 2, b++) {
            const W4 r = w4_of(eq_point[j + 1 + b]), omr = w4_sub(one, r);
            for (size_t i = 0; i < h; i++) {
                E[i + h] = w4_mul(E[i], r);
                E[i] = w4_mul(E[i], omr);
            }
        }
        W4 S[SC_MAX_NODES];
        for (int x = 0; x <= D; x++) S[x] = zero;
        for (size_t i = 0; i < m; i++) {
            W4 cur[SC_MAX_TABLES], step[SC_MAX_TABLES];
            for (int t = 0; t < k; t++) {
                const W4 *f = &tab[(size_t)t * len];
                cur[t] = f[2 * i];
                step[t] = w4_sub(f[2 * i + 1], f[2 * i]);
            }
            for (int x = 0; x <= D; x++) {
                W4 g = zero;
                for (int q = 0; q < gate.nterms; q++) {
                    const Term &tm = gate.terms[q];
                    W4 v = w4_mul(coef[q], cur[tm.f[0]]);
                    if (tm.f[1] >= 0) v = w4_mul(v, cur[tm.f[1]]);
                    if (tm.f[2] >= 0) v = w4_mul(v, cur[tm.f[2]]);
                    g = w4_add(g, v);
                }
                S[x] = w4_add(S[x], w4_mul(E[i], g));
                if (x < D)
                    for (int t = 0; t < k; t++) cur[t] = w4_add(cur[t], step[t]);
            }
        }
        const EqLine eqj(eq_point[j]);
        W4 T[SC_MAX_NODES], cw[SC_MAX_NODES];
        for (int x = 0; x <= D; x++) T[x] = w4_mul(P, S[x]);
        round_poly(T, D, d, eqj, false, zero, cw);
        Fe *coeffs = rounds_out + (size_t)j * (d + 1);
        for (int x = 0; x <= d; x++) coeffs[x] = fe_of(cw[x]);
        const Fe rho = round_fn(coeffs, d + 1);
        point_out[j] = rho;
        const W4 rw = w4_of(rho);
        P = w4_mul(P, eqj.at(rw));
        fold(rw);
    }
    for (int t = 0; t < k; t++) finals_out[t] = fe_of(tab[(size_t)t * len]);
}

static void sumcheck_rounds(Context &ctx, const SumcheckScratch &sc, Fe *hvals, bool mapped_nodes, bool timing,
                            const Gate &gate, Fe *const *tables, const uint8_t *kinds, int k, int n,
                            const Fe *eq_point, const RoundFn &round_fn, Fe *rounds_out, Fe *point_out,
                            Fe *finals_out) {
    bool narrow = false;  // a narrow backing: the first two rounds read through the lift
    for (int t = 0; t < k && kinds; t++) narrow = narrow || kinds[t] != SC_FR;
    const int D = gate.x_degree();  // degree of G along one variable
    const int d = gate.degree();    // round polynomial degree
    const bool product = gate.is_plain_product(k);
    hipStream_t st = ctx.stream;
    const size_t len0 = size_t(1) << n;

    if ((size_t)k * len0 <= TAIL_ELEMS) {  // small enough for the host from the start
        sumcheck_host_tail(ctx, gate, k, n, 0, eq_point, tables, kinds, len0, nullptr, 0, f_one<Fr>(),
                           round_fn, rounds_out, point_out, finals_out);
        return;
    }

    // suffix-eq tables (k_eq_suffix), split at h0
    const int h0 = std::max(1, std::min(n, (n + 1) / 2));
    const size_t astride = size_t(1) << (h0 - 1), bstride = size_t(1) << (n - h0);
    const size_t asz = (size_t)(h0 - 1) * astride, bsz = (size_t)(n - h0 + 1) * bstride;
    Fe *rd = sc.r, *eqA = sc.eqA, *eqB = sc.eqB;
    HIP_CHECK(hipMemcpyAsync(rd, eq_point, n * sizeof(Fe), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_eq_suffix, dim3(grid_for(asz + bsz, 256)), dim3(256), 0, st, rd, n, h0, eqA, eqB);
    HIP_CHECK(hipGetLastError());

    // 1 / r_j by one batched inversion (zeros skipped: such rounds use all nodes)
    std::vector<Fe> inv_r(n), pre(n + 1);
    pre[0] = f_one<Fr>();
    for (int j = 0; j < n; j++)
        pre[j + 1] = f_is_zero(eq_point[j]) ? pre[j] : f_mul<Fr>(pre[j], eq_point[j]);
    {
        Fe inv = f_inv<Fr>(pre[n]);
        for (int j = n - 1; j >= 0; j--) {
            if (f_is_zero(eq_point[j])) {
                inv_r[j] = f_zero<Fr>();
                continue;
            }
            inv_r[j] = f_mul<Fr>(inv, pre[j]);
            inv = f_mul<Fr>(inv, eq_point[j]);
        }
    }

    // two ping-pong buffers per table (the inputs are read in round 0 and
    // folded into scratch from round 1 on)
    std::vector<Fe *> bufA(k), bufB(k);
    for (int t = 0; t < k; t++) {
        bufA[t] = sc.poolA + (size_t)t * std::max<size_t>(len0 / 2, 1);
        bufB[t] = sc.poolB + (size_t)t * std::max<size_t>(len0 / 4, 1);
    }
    std::vector<const Fe *> cur(k);
    for (int t = 0; t < k; t++) cur[t] = tables[t];

    Fe *partials = sc.partials;
    Fe *hvals_dev = sc.nodes;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (timing) {
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

    // Round 0 reads the inputs T^0 (2^n entries). A single-round launch at
    // round j >= 1 folds T^(j-1) with rho_(j-1) into T^j (2^(n-j) entries:
    // odd j into bufA, even j into bufB, so source and destination never
    // alias) and evaluates T^j's pairs. A paired launch (k_round_pair) at
    // round j folds the stored table with the one or two challenges drawn
    // since it was written into T^j, in the buffer that does not hold its
    // source, and evaluates rounds j and j+1.
    // the product kernel's node sums carry 2^(-5 k), and 2^-5 more when E = lo hi
    const W4 prod_scale = hfr::w4_of(product ? f_from_u64<Fr>(uint64_t(1) << (5 * k)) : f_one<Fr>());  // k <= 3
    const W4 prod_scale_lo = hfr::w4_of(product ? f_from_u64<Fr>(uint64_t(1) << (5 * k + 5)) : f_one<Fr>());
    W4 claim = smallc().zero;
    W4 P = smallc().one;  // prod_{l<j} eq(r_l, rho_l)

    // E_(s-1), the suffix-eq factor of round s-1's pairs (tables at s)
    auto eq_args = [&](int s) {
        EqArgs eq;
        if (s < h0) {
            eq.lo = eqA + (size_t)(s - 1) * astride;
            eq.lo_bits = (uint32_t)(h0 - s);
            eq.hi = eqB;
        } else {
            eq.lo = nullptr;
            eq.lo_bits = 0;
            eq.hi = eqB + (size_t)(s - h0) * bstride;
        }
        return eq;
    };

    // Round j on the host from the node sums S (X = 0..D, or X = 0, 2..D with
    // node 1 from the claim): T(X) = P S(X) (rescaled for the product kernel)
    // at X = 0..D, then X = D+1..d by extrapolation; returns rho_j.
    auto host_round = [&](int j, const W4 *S, bool full, bool eq_lo) -> Fe {
        using namespace hfr;
        const EqLine eqj(eq_point[j]);
        const W4 Ps = product ? w4_mul(P, eq_lo ? prod_scale_lo : prod_scale) : P;
        W4 T[SC_MAX_NODES], cw[SC_MAX_NODES];
        if (full) {
            for (int x = 0; x <= D; x++) T[x] = w4_mul(Ps, S[x]);
        } else {
            T[0] = w4_mul(Ps, S[0]);
            for (int x = 2; x <= D; x++) T[x] = w4_mul(Ps, S[x - 1]);
            const W4 g0 = w4_mul(eqj.base, T[0]);
            T[1] = w4_mul(w4_sub(claim, g0), w4_of(inv_r[j]));  // g(1) = r_j T(1)
        }
        round_poly(T, D, d, eqj, !full, claim, cw);
        Fe *coeffs = rounds_out + (size_t)j * (d + 1);
        for (int x = 0; x <= d; x++) coeffs[x] = fe_of(cw[x]);
        const Fe rho = round_fn(coeffs, d + 1);
        point_out[j] = rho;
        const W4 rw = w4_of(rho);
        claim = horner_w4(cw, d, rw);
        P = w4_mul(P, eqj.at(rw));
        return rho;
    };

    // G: the first round left to the host. The device rounds from first_pair
    // to G-1 run in pairs (an even count of them) when the summand is the
    // product of two tables and every r_j of those rounds is nonzero.
    int G = 1;
    while (G < n && (size_t)k * (len0 >> G) > TAIL_ELEMS) G++;
    int first_pair = G;
    if (product && k == 2) {
        bool nonzero = true;
        for (int j = 1; j < G; j++) nonzero = nonzero && !f_is_zero(eq_point[j]);
        if (nonzero) {
            first_pair = (G - 1) % 2 == 0 ? 1 : 2;
            if (narrow && first_pair < 2) first_pair += 2;  // rounds 0 and 1 lift
        }
    }

    Fe pend[2];  // challenges drawn since cur was written, oldest first
    int npend = 0;
    int j = 0;
    while (j < n) {
        // between launches (the last one's flag was seen); round j starts
        // after 1 - 2^-j of the rounds' work
        if (j > 0) ctx.yield_point(1.0 - std::ldexp(1.0, -j));
        if (j > 0 && (size_t)k * (len0 >> j) <= TAIL_ELEMS) {  // T^j is small: the rest on the host
            std::vector<Fe *> src(k);
            for (int t = 0; t < k; t++) src[t] = const_cast<Fe *>(cur[t]);
            // cur is T^(j - npend), still the caller's table when j == npend
            sumcheck_host_tail(ctx, gate, k, n, j, eq_point, src.data(), j == npend ? kinds : nullptr,
                               len0 >> (j - npend), pend, npend, hfr::fe_of(P), round_fn, rounds_out, point_out,
                               finals_out);
            return;
        }
        const size_t m = len0 >> (j + 1);  // pairs of T^j
        std::vector<Fe *> nxt(k);
        TablePtrs tp;
        if (j >= first_pair) {
            // rounds j and j+1 in one launch; T^j into the buffer cur is not in
            // (bufB, len0 / 4 entries per table, only from j = 2 on: at j = 1
            // cur is still the caller's table)
            const bool into_a = cur[0] != bufA[0];
            for (int t = 0; t < k; t++) {
                nxt[t] = into_a ? bufA[t] : bufB[t];
                tp.src[t] = cur[t];
                tp.kind[t] = (uint8_t)SC_FR;
                tp.dst[t] = nxt[t];
            }
            const EqArgs eq = eq_args(j + 2);
            PairWeights pw;
            pw.w[0] = f29_from_mont32<Fr>(pend[0]);
            pw.w[1] = npend == 2 ? f29_from_mont32<Fr>(pend[1]) : F29{};
            pw.w[2] = npend == 2 ? f29_from_mont32<Fr>(f_mul<Fr>(pend[0], pend[1])) : F29{};
            pw.c0 = f29_from_mont32<Fr>(f_sub<Fr>(f_one<Fr>(), eq_point[j + 1]));
            pw.c1 = f29_from_mont32<Fr>(eq_point[j + 1]);
            if (timing) HIP_CHECK(hipEventRecord(ev0, st));
            if (sc.done_host) __atomic_store_n(sc.done_host, 0u, __ATOMIC_RELEASE);
            launch_pair(ctx, npend, tp, m, pw, eq, partials, hvals_dev, sc.done_dev);
            if (timing) HIP_CHECK(hipEventRecord(ev1, st));
            if (!mapped_nodes)
                HIP_CHECK(hipMemcpyAsync(hvals, hvals_dev, PAIR_OUTS * sizeof(Fe), hipMemcpyDeviceToHost, st));
            round_fn.begin(d + 1);
            if (sc.done_host)
                wait_round(ctx, sc.done_host);
            else
                ctx.sync();
            if (timing) {
                float ms = 0;
                HIP_CHECK(hipEventSynchronize(ev1));
                HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
                ctx.stats.sumcheck_kernel_ms += ms;
                ctx.stats.sumcheck_launches += 1;
                // reads the stored table once, writes T^j
                ctx.stats.sumcheck_bytes += (uint64_t)k * 32 * ((npend == 2 ? 8 : 4) * m + 2 * m);
            }
            using namespace hfr;
            W4 h[PAIR_OUTS];
            for (int x = 0; x < PAIR_OUTS; x++) h[x] = w4_of(hvals[x]);
            const Fe rj = host_round(j, h, false, eq.lo != nullptr);
            // round j+1's node sums at rho = rho_j: Lagrange weights over y = 0, 1, 2
            const SmallConsts &K = smallc();
            const W4 r = w4_of(rj), r1 = w4_sub(r, K.one), r2 = w4_sub(r, K.num[2]);
            const W4 L0 = w4_mul(w4_mul(r1, r2), K.inv2), L1 = w4_sub(K.zero, w4_mul(r, r2)),
                     L2 = w4_mul(w4_mul(r, r1), K.inv2);
            W4 S[2];
            for (int x = 0; x < 2; x++)
                S[x] = w4_add(w4_add(w4_mul(L0, h[2 + x]), w4_mul(L1, h[4 + x])), w4_mul(L2, h[6 + x]));
            const Fe rj1 = host_round(j + 1, S, false, eq.lo != nullptr);
            for (int t = 0; t < k; t++) cur[t] = nxt[t];
            pend[0] = rj;
            pend[1] = rj1;
            npend = 2;
            j += 2;
            continue;
        }
        const bool fold = j > 0;  // before any pair: at most one challenge pending
        const bool full = j == 0 || f_is_zero(eq_point[j]);
        for (int t = 0; t < k; t++) {
            nxt[t] = (j % 2 == 1) ? bufA[t] : bufB[t];
            tp.src[t] = cur[t];
            tp.kind[t] = (j <= 1 && kinds) ? kinds[t] : (uint8_t)SC_FR;
            tp.dst[t] = nxt[t];
        }
        const EqArgs eq = eq_args(j + 1);
        if (timing) HIP_CHECK(hipEventRecord(ev0, st));
        // cleared before the launch: the previous round's kernel has set it and writes it no more
        if (sc.done_host) __atomic_store_n(sc.done_host, 0u, __ATOMIC_RELEASE);
        launch_round(ctx, gate, k, product, full, D, narrow && j <= 1, tp, m, fold, fold ? pend[0] : f_zero<Fr>(),
                     eq, partials, hvals_dev, sc.done_dev);
        HIP_CHECK(hipGetLastError());
        if (timing) HIP_CHECK(hipEventRecord(ev1, st));
        const int nn = full ? D + 1 : D;
        if (!mapped_nodes)
            HIP_CHECK(hipMemcpyAsync(hvals, hvals_dev, nn * sizeof(Fe), hipMemcpyDeviceToHost, st));
        round_fn.begin(d + 1);  // the transcript's message prefix, while the kernel runs
        if (sc.done_host)
            wait_round(ctx, sc.done_host);
        else
            ctx.sync();
        if (timing) {
            float ms = 0;
            HIP_CHECK(hipEventSynchronize(ev1));
            HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
            ctx.stats.sumcheck_kernel_ms += ms;
            ctx.stats.sumcheck_launches += 1;
            // reads every table once; a folding round also writes the
            // half-size fold
            ctx.stats.sumcheck_bytes += (uint64_t)k * 32 * (fold ? 6 * m : 2 * m);
        }
        W4 S[SC_MAX_NODES];
        for (int x = 0; x < nn; x++) S[x] = hfr::w4_of(hvals[x]);
        pend[0] = host_round(j, S, full, eq.lo != nullptr);
        npend = 1;
        if (fold)
            for (int t = 0; t < k; t++) cur[t] = nxt[t];
        j += 1;
    }
    throw ApgError(-5, "sumcheck: the host tail was not reached");  // k * 2 <= TAIL_ELEMS always holds
}


// ---- the whole sumcheck as one captured graph, with the device transcript ----
//
// For callers that want no host in the loop (one graph launch per proof):
// the product summand eq * t_0 * ... * t_(k-1), every round evaluating all
// nodes (no node from the claim), and after each round kernel a one-thread
// kernel that turns the node sums into the round polynomial, runs the
// round's transcript step on the device (transcript_dev.cuh) and leaves the
// challenge in device memory, where the next round kernel folds with it. The
// bytes equal the host-driven path's; the price is the device transcript's
// latency (a permutation on one thread is ~70x the host's).

namespace {

struct GraphConsts {
    Fe num[SC_MAX_NODES + 1], inv[SC_MAX_NODES + 1], inv6;
};

struct GraphState {
    DevTranscript tr;
    Fe P;    // prod_{l<j} eq(r_l, rho_l)
    Fe rho;  // the last challenge
    GraphConsts K;
};

// the device forms of interpolate_w4 and extrapolate_w4
__device__ void dev_interpolate(const Fe *v, int d, Fe *c, const GraphConsts &K) {
    if (d == 3) {  // crates/sumcheck/src/lib.rs:294-305
        const Fe d1 = f_sub<Fr>(v[1], v[0]);
        const Fe t = f_sub<Fr>(v[2], v[1]);
        const Fe d2 = f_sub<Fr>(t, d1);
        const Fe d3 = f_sub<Fr>(f_sub<Fr>(f_sub<Fr>(v[3], v[2]), t), d2);
        c[0] = v[0];
        c[1] = f_add<Fr>(f_sub<Fr>(d1, f_mul<Fr>(d2, K.inv[2])), f_mul<Fr>(d3, K.inv[3]));
        c[2] = f_mul<Fr>(f_sub<Fr>(d2, d3), K.inv[2]);
        c[3] = f_mul<Fr>(d3, K.inv6);
        return;
    }
    Fe diff[SC_MAX_NODES], delta[SC_MAX_NODES];
    for (int i = 0; i <= d; i++) diff[i] = v[i];
    for (int k = 0; k <= d; k++) {
        delta[k] = diff[0];
        for (int i = 0; i + k < d; i++) diff[i] = f_sub<Fr>(diff[i + 1], diff[i]);
    }
    Fe basis[SC_MAX_NODES], acc[SC_MAX_NODES];
    for (int i = 0; i <= d; i++) basis[i] = acc[i] = f_zero<Fr>();
    basis[0] = K.num[1];
    acc[0] = delta[0];
    for (int k = 1; k <= d; k++) {
        Fe nb[SC_MAX_NODES];
        for (int i = 0; i <= d; i++) nb[i] = f_zero<Fr>();
        for (int i = 0; i < k; i++) {
            nb[i + 1] = f_add<Fr>(nb[i + 1], basis[i]);
            nb[i] = f_sub<Fr>(nb[i], f_mul<Fr>(basis[i], K.num[k - 1]));
        }
        for (int i = 0; i <= k; i++) basis[i] = f_mul<Fr>(nb[i], K.inv[k]);
        for (int i = 0; i <= k; i++) acc[i] = f_add<Fr>(acc[i], f_mul<Fr>(basis[i], delta[k]));
    }
    for (int i = 0; i <= d; i++) c[i] = acc[i];
}

__device__ void dev_extrapolate(Fe *v, int D, int d) {
    Fe last[SC_MAX_NODES], row[SC_MAX_NODES];
    for (int i = 0; i <= D; i++) row[i] = v[i];
    for (int k = 0; k <= D; k++) {
        last[k] = row[D - k];
        for (int i = 0; i + k < D; i++) row[i] = f_sub<Fr>(row[i + 1], row[i]);
    }
    for (int x = D + 1; x <= d; x++) {
        for (int k = D - 1; k >= 0; k--) last[k] = f_add<Fr>(last[k], last[k + 1]);
        v[x] = last[0];
    }
}

// One thread: round j's polynomial from the node sums S(0..D) (times
// `scale`, which undoes the product kernel's 2^-5 factors), its transcript
// step, the challenge and the running eq product.
__global__ void k_graph_round(GraphState *st, const Fe *__restrict__ nodes, int D, int d, Fe scale,
                              const Fe *__restrict__ r, int j, Fe *__restrict__ rounds, Fe *__restrict__ point,
                              const Fe *__restrict__ rc) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const Fe Ps = f_mul<Fr>(st->P, scale);
    Fe T[SC_MAX_NODES], c[SC_MAX_NODES];
    for (int x = 0; x <= D; x++) T[x] = f_mul<Fr>(Ps, nodes[x]);
    dev_extrapolate(T, D, d);
    const Fe one = st->K.num[1];
    const Fe base = f_sub<Fr>(one, r[j]), slope = f_sub<Fr>(f_add<Fr>(r[j], r[j]), one);
    Fe e = base;
    for (int x = 0; x <= d; x++) {
        T[x] = f_mul<Fr>(e, T[x]);
        e = f_add<Fr>(e, slope);
    }
    dev_interpolate(T, d, c, st->K);
    for (int x = 0; x <= d; x++) rounds[(size_t)j * (d + 1) + x] = c[x];
    DevTranscript t = st->tr;
    dt_append_scalars(t, TAG_SUMCHECK_ROUND, c, d + 1, rc);
    const Fe rho = dt_challenge(t, TAG_SUMCHECK_CHALLENGE, rc);
    st->tr = t;
    st->rho = rho;
    point[j] = rho;
    st->P = f_mul<Fr>(st->P, f_add<Fr>(base, f_mul<Fr>(rho, slope)));
}

// The final evaluations: each table's last pair folded with the last challenge.
__global__ void k_graph_finals(const GraphState *st, TablePtrs tp, int k, Fe *__restrict__ finals) {
    const int t = threadIdx.x;
    if (blockIdx.x != 0 || t >= k) return;
    const Fe lo = ldfe(tp.src[t], 0), hi = ldfe(tp.src[t], 1);
    finals[t] = f_add<Fr>(lo, f_mul<Fr>(st->rho, f_sub<Fr>(hi, lo)));
}

template <int KT>
void launch_graph_product(Context &ctx, bool full, const TablePtrs &tp, size_t m, bool fold, const EqArgs &eq,
                          Fe *partials, Fe *out, const Fe *rho_dev) {
    const dim3 g(round_blocks(ctx, m, PRODUCT_BLOCKS_PER_CU)), b(RTPB);
    const F29 none{};
    unsigned *no_flag = nullptr;
    if (full) {
        hipLaunchKernelGGL((k_round_product<KT, true, false>), g, b, 0, ctx.stream, tp, m, (int)fold, none, eq,
                           partials, out, no_flag, rho_dev);
        launch_finish<KT + 1>(ctx.stream, g.x, KT + 1, partials, out, no_flag);
    } else {
        hipLaunchKernelGGL((k_round_product<KT, false, false>), g, b, 0, ctx.stream, tp, m, (int)fold, none, eq,
                           partials, out, no_flag, rho_dev);
        launch_finish<KT>(ctx.stream, g.x, KT, partials, out, no_flag);
    }
}

}  // namespace

void sumcheck_prove_graph(Context &ctx, Fe *const *tables, int k, int n, const Fe *eq_point, Transcript &tr,
                          Fe *rounds_out, Fe *point_out, Fe *finals_out) {
    if (k < 1 || k > 3) throw ApgError(-2, "graph sumcheck: the product of 1 to 3 tables");
    if (n < 1 || n > 30) throw ApgError(-2, "graph sumcheck: num_vars out of range");
    // S has degree k in the round variable and eq adds one; rounds are at
    // least cubic (four coefficients, the reference's [Fr; 4]), as Gate::degree
    const int D = k, d = std::max(3, k + 1);
    const size_t len0 = size_t(1) << n;
    const ScratchLayout l = sumcheck_layout(k, n);
    SumcheckScratch sc = sumcheck_carve(ctx.d<uint8_t>("sg.scratch", l.total), l);
    GraphState *st = ctx.d<GraphState>("sg.state", 1);
    Fe *rounds = ctx.d<Fe>("sg.rounds", (size_t)n * (d + 1));
    Fe *point = ctx.d<Fe>("sg.point", n);
    Fe *finals = ctx.d<Fe>("sg.finals", k);
    const Fe *rc = poseidon2_rc_device(ctx);

    // inputs of the graph: the eq point and the starting state
    GraphState h{};
    std::memcpy(&h.tr, &tr, sizeof(DevTranscript));
    h.P = f_one<Fr>();
    h.rho = f_zero<Fr>();
    for (int x = 0; x <= SC_MAX_NODES; x++) {
        h.K.num[x] = f_from_u64<Fr>((uint64_t)x);
        h.K.inv[x] = x ? f_inv<Fr>(h.K.num[x]) : f_zero<Fr>();
    }
    h.K.inv6 = f_inv<Fr>(f_from_u64<Fr>(6));
    hipStream_t s = ctx.stream;
    HIP_CHECK(hipMemcpyAsync(sc.r, eq_point, n * sizeof(Fe), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    ctx.sync();  // h lives on this frame; nothing else is in flight when the capture starts

    const int h0 = std::max(1, std::min(n, (n + 1) / 2));
    const size_t astride = size_t(1) << (h0 - 1), bstride = size_t(1) << (n - h0);
    const size_t asz = (size_t)(h0 - 1) * astride, bsz = (size_t)(n - h0 + 1) * bstride;
    std::vector<Fe *> bufA(k), bufB(k);
    for (int t = 0; t < k; t++) {
        bufA[t] = sc.poolA + (size_t)t * std::max<size_t>(len0 / 2, 1);
        bufB[t] = sc.poolB + (size_t)t * std::max<size_t>(len0 / 4, 1);
    }
    std::vector<const Fe *> cur(k);
    for (int t = 0; t < k; t++) cur[t] = tables[t];
    // the product kernel's node sums carry 2^(-5 k), and 2^-5 more with two eq factor tables
    const Fe scale = f_from_u64<Fr>(uint64_t(1) << (5 * k)), scale_lo = f_from_u64<Fr>(uint64_t(1) << (5 * k + 5));

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    struct GraphGuard {
        hipGraph_t &g;
        hipGraphExec_t &e;
        ~GraphGuard() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
        }
    } guard{graph, exec};
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        hipLaunchKernelGGL(k_eq_suffix, dim3(grid_for(asz + bsz, 256)), dim3(256), 0, s, sc.r, n, h0, sc.eqA, sc.eqB);
        for (int j = 0; j < n; j++) {
            const bool fold = j > 0;
            const size_t m = len0 >> (j + 1);
            TablePtrs tp;
            for (int t = 0; t < k; t++) {
                Fe *nxt = (j % 2 == 1) ? bufA[t] : bufB[t];
                tp.src[t] = cur[t];
                tp.kind[t] = (uint8_t)SC_FR;
                tp.dst[t] = nxt;
            }
            const int sidx = j + 1;
            EqArgs eq;
            if (sidx < h0) {
                eq.lo = sc.eqA + (size_t)(sidx - 1) * astride;
                eq.lo_bits = (uint32_t)(h0 - sidx);
                eq.hi = sc.eqB;
            } else {
                eq.lo = nullptr;
                eq.lo_bits = 0;
                eq.hi = sc.eqB + (size_t)(sidx - h0) * bstride;
            }
            switch (k) {
                case 1: launch_graph_product<1>(ctx, true, tp, m, fold, eq, sc.partials, sc.nodes, &st->rho); break;
                case 2: launch_graph_product<2>(ctx, true, tp, m, fold, eq, sc.partials, sc.nodes, &st->rho); break;
                default: launch_graph_product<3>(ctx, true, tp, m, fold, eq, sc.partials, sc.nodes, &st->rho); break;
            }
            hipLaunchKernelGGL(k_graph_round, dim3(1), dim3(64), 0, s, st, sc.nodes, D, d, eq.lo ? scale_lo : scale,
                               sc.r, j, rounds, point, rc);
            if (fold)
                for (int t = 0; t < k; t++) cur[t] = tp.dst[t];
        }
        TablePtrs last;
        for (int t = 0; t < k; t++) {
            last.src[t] = cur[t];
            last.dst[t] = nullptr;
            last.kind[t] = (uint8_t)SC_FR;
        }
        hipLaunchKernelGGL(k_graph_finals, dim3(1), dim3(64), 0, s, st, last, k, finals);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIP_CHECK(hipStreamEndCapture(s, &graph));
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphLaunch(exec, s));
    HIP_CHECK(hipMemcpyAsync(rounds_out, rounds, (size_t)n * (d + 1) * sizeof(Fe), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(point_out, point, n * sizeof(Fe), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(finals_out, finals, k * sizeof(Fe), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
    ctx.sync();
    std::memcpy(&tr, &h.tr, sizeof(DevTranscript));
}

}  // namespace apg
