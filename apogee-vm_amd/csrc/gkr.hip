// The general GKR layer for gfx950: one resolved gate list (cached entries,
// producing and enforcing gates of any of the seven shapes, the virtual
// tables, both children of a halving list) evaluated by the device.
//
// Replaces the pair loop of gkr::prove_sumcheck (crates/gkr/src/lib.rs:
// 792-899) with the summand of ResolvedList::summand (crates/gkr-verify/src/
// lib.rs:430-460) and the row loop of gkr::produce (crates/gkr/src/lib.rs:
// 300-336):
//
//  * The program (gates, sources, coefficients, weights) sits in device
//    memory and every thread walks it in the same order, so its loads are
//    wave-uniform. Every shape is a stream of operand steps over three
//    accumulators (one multiply site for all shapes).
//  * A validated list has every gate of degree <= 2 after substitution (the
//    artifact's law, checked on the host), so the summand S(X) of a round is
//    a quadratic in the round variable: the device sums E(i) S_i(X) at
//    X = 0, 1, 2, and the host takes S's coefficients and multiplies by the
//    linear eq factor. The cubic is unique, so these are the coefficients the
//    reference interpolates from its four nodes; no node comes from the claim.
//  * eq is factored as in the gate sumcheck (sumcheck_dev.cuh): a host
//    scalar times the one-variable factor times the suffix table E_{j+1}(i).
//  * Virtual tables and cached entries are evaluated once per point into the
//    thread's slots (global memory, slot-major so a wave's accesses are
//    contiguous) before the weighted gates read them.
//  * The fold of round j - 1 runs at the top of round j's pair loop: each
//    thread folds its pair of every table and then reads the folded values.
//  * Virtual tables are never materialised: their closed forms are evaluated
//    at (rho_0..rho_{j-1}, X, bits of i) (crates/gkr/src/lib.rs:760-782;
//    each form is affine in every coordinate, so X = 2 gives the line's
//    value) and at each row in the forward pass.
#include <algorithm>
#include <cstring>

#include "field29.cuh"
#include "gkr.h"
#include "sumcheck_dev.cuh"

namespace apg {

namespace {

// The multiplies of a gate are the lazy 29-bit ones of field29.cuh on 32-bit
// images: a value times a multiplier held in the R' = 2^261 domain is the
// product's exact image, and a value times a value is the image of their
// product over 2^5. So the host stores each coefficient twice (coefm as a
// multiplier, coefc as a constant), times 2^5 where its product is later
// multiplied by another value (a Product's coefficient, an AffineProduct's
// left side, a Quadratic's product coefficients), and an operand that is
// multiplied later without a coefficient is doubled five times: every gate
// value comes out exact.
struct DProg {
    const GkrGate *gates;
    const uint32_t *src;
    const F29 *coefm;
    const Fe *coefc;
    const uint32_t *kinds;
    const F29 *w;  // weights of the producing then enforcing gates, R' domain
    uint32_t ncached, ngates, nvirt;
};

// --- virtual tables (crates/gkr-verify/src/lib.rs:204-288) ---

__device__ __forceinline__ Fe fe_u64(uint64_t x) { return f_from_u64<Fr>(x); }

__device__ Fe virtual_at_row(uint32_t kind, uint64_t row) {
    switch (kind) {
        case VK_ROW_INDEX: return fe_u64(row);
        case VK_RAM_LIVE: return fe_u64(row >= (1ull << RAM_LIVE_BIT) ? 1 : 0);
        case VK_RANGE19: return fe_u64(row & ((1ull << 19) - 1));
        case VK_RANGE16: return fe_u64(row & ((1ull << 16) - 1));
        case VK_XOR8_A: return fe_u64(row & 0xff);
        case VK_XOR8_B: return fe_u64((row >> 8) & 0xff);
        default: return fe_u64((row & 0xff) ^ ((row >> 8) & 0xff));
    }
}

// Coordinate j of the round's point (rho_0..rho_{b-1}, X, bits of i), zero
// past the n variables. Every closed form is affine in each coordinate, so
// X = 2 gives the line's value there.
struct RoundPoint {
    const Fe *rho;
    int n, b;
    size_t i;
    int x;
    __device__ __forceinline__ Fe at(int j) const {
        if (j >= n) return f_zero<Fr>();
        if (j < b) return ldfe(rho, j);
        if (j == b) return fe_u64((uint64_t)x);
        return ((i >> (j - b - 1)) & 1) ? f_one<Fr>() : f_zero<Fr>();
    }
};

__device__ __noinline__ Fe virtual_at_point(uint32_t kind, const RoundPoint &p) {
    Fe acc = f_zero<Fr>();
    switch (kind) {
        case VK_ROW_INDEX:  // sum_j 2^j y_j, Horner from the highest variable
            for (int j = p.n - 1; j >= 0; j--) acc = f_add<Fr>(f_add<Fr>(acc, acc), p.at(j));
            return acc;
        case VK_RAM_LIVE: {  // 1 - prod_{j >= RAM_LIVE_BIT} (1 - y_j)
            Fe prod = f_one<Fr>();
            for (int j = min(p.n, RAM_LIVE_BIT); j < p.n; j++) prod = f_mul<Fr>(prod, f_sub<Fr>(f_one<Fr>(), p.at(j)));
            return f_sub<Fr>(f_one<Fr>(), prod);
        }
        case VK_RANGE19:
        case VK_RANGE16: {
            const int bits = kind == VK_RANGE19 ? 19 : 16;
            for (int j = min(p.n, bits) - 1; j >= 0; j--) acc = f_add<Fr>(f_add<Fr>(acc, acc), p.at(j));
            return acc;
        }
        case VK_XOR8_A:
        case VK_XOR8_B: {
            const int first = kind == VK_XOR8_A ? 0 : 8;
            for (int j = 7; j >= 0; j--) acc = f_add<Fr>(f_add<Fr>(acc, acc), p.at(first + j));
            return acc;
        }
        default: {  // sum_j 2^j (y_j + y_{j+8} - 2 y_j y_{j+8})
            for (int j = 7; j >= 0; j--) {
                const Fe y = p.at(j), z = p.at(j + 8), yz = f_mul<Fr>(y, z);
                acc = f_sub<Fr>(f_add<Fr>(f_add<Fr>(f_add<Fr>(acc, acc), y), z), f_add<Fr>(yz, yz));
            }
            return acc;
        }
    }
}

// Operands at one pair and node X in {0, 1, 2} (the round) or at one row
// (the forward pass: X unused, child 1 at row y + rows). Virtual tables and
// cached entries are evaluated once per point into this thread's slots
// (slot k of thread tid at slots[k * nthr + tid]) before the gates read them.
struct Reader {
    const Fe *const *tab;
    uint32_t nlower;
    size_t i, rows;
    int x;
    bool row_mode;
    Fe *slots;
    size_t nthr, tid;
    // set: virtual tables are evaluated where a gate reads them (the round's
    // point from rho, n, b) instead of from the slots
    const Fe *rho;
    int n, b;
    bool lazy_virtual;
    __device__ __forceinline__ Fe line(uint32_t t) const {
        const Fe *p = tab[t];
        if (x == 0) return ldfe(p, 2 * i);
        const Fe hi = ldfe(p, 2 * i + 1);
        if (x == 1) return hi;
        return f_sub<Fr>(f_add<Fr>(hi, hi), ldfe(p, 2 * i));
    }
    __device__ __forceinline__ Fe lower(uint32_t c) const { return row_mode ? ldfe(tab[c], i) : line(c); }
    __device__ __forceinline__ Fe upper(uint32_t c) const {
        return row_mode ? ldfe(tab[c], i + rows) : line(nlower + c);
    }
    __device__ __forceinline__ Fe slot(uint32_t k) const { return ldfe(slots, (size_t)k * nthr + tid); }
    __device__ __forceinline__ void set_slot(uint32_t k, const Fe &v) const { stfe(slots, (size_t)k * nthr + tid, v); }
};

// What operand q of a gate does (in eval_gate's value order): t = k v, v, or
// A v / B v (a product with an earlier operand); then acc += t, A = t,
// A += t, B = t or B += t.
enum : int { TO_ACC = 0, SET_A = 1, ADD_A = 2, SET_B = 3, ADD_B = 4 };
enum : int { BY_ONE = 0, BY_K = 1, BY_A = 2, BY_B = 3 };

// gkr_verify::eval_gate (crates/gkr-verify/src/lib.rs:142-202) with its
// operands gathered as ResolvedList::eval gathers them (:462-489): every
// shape as a stream of operand steps over three accumulators, so one
// multiply site serves all of them.
__device__ Fe eval_gate(const DProg &p, uint32_t g, const Reader &rd) {
    const GkrGate G = p.gates[g];
    const uint32_t nv = G.src_end - G.src_begin, t = G.split;
    const F29 *cm = p.coefm + G.coef_begin;
    const Fe *cc = p.coefc + G.coef_begin;
    Fe acc = f_zero<Fr>(), A = acc, B = acc;
    switch (G.shape) {
        case GS_LINEAR: acc = cc[nv]; break;       // c_1..c_t, c_0
        case GS_MASK: acc = f_one<Fr>(); break;    // x m + (1 - m)
        case GS_AFFINE_PRODUCT:                    // a_1..a_t, a_0, b_1..b_u, b_0
            A = cc[t];
            B = cc[nv + 1];
            break;
        case GS_QUADRATIC: acc = cc[0]; break;     // c_0, a_1..a_t, b_1..b_u
        default: break;
    }
    for (uint32_t q = 0; q < nv; q++) {
        int to = TO_ACC, by = BY_K;
        uint32_t kq = 0;
        bool minus_v = false;
        switch (G.shape) {
            case GS_LINEAR: kq = q; break;
            case GS_PRODUCT:  // c x y
                if (q == 0) to = SET_A; else by = BY_A;
                break;
            case GS_MASK:
                if (q == 0) { to = SET_A; by = BY_ONE; } else { by = BY_A; minus_v = true; }
                break;
            case GS_AFFINE_PRODUCT:
                if (q < t) { to = ADD_A; kq = q; } else { to = ADD_B; kq = q + 1; }
                break;
            case GS_TREE_PRODUCT:
                if (q == 0) { to = SET_A; by = BY_ONE; } else { by = BY_A; }
                break;
            case GS_QUADRATIC:  // the products read y then z
                if (q < t) kq = 1 + q;
                else if (((q - t) & 1) == 0) { to = SET_A; kq = 1 + t + (q - t) / 2; }
                else by = BY_A;
                break;
            default:  // GS_TREE_CROSS: p0, p1, q0, q1 -> p0 q1 + p1 q0
                if (q == 0) { to = SET_A; by = BY_ONE; }
                else if (q == 1) { to = SET_B; by = BY_ONE; }
                else by = q == 2 ? BY_B : BY_A;
                break;
        }
        const uint32_t s = p.src[G.src_begin + q], idx = s & SRC_MASK;
        Fe v;
        switch (s >> SRC_SHIFT) {
            case SRC_LOWER: v = rd.lower(idx); break;
            case SRC_UPPER: v = rd.upper(idx); break;
            case SRC_VIRTUAL:
                v = rd.lazy_virtual ? virtual_at_point(p.kinds[idx], RoundPoint{rd.rho, rd.n, rd.b, rd.i, rd.x})
                                    : rd.slot(idx);
                break;
            default: v = rd.slot(p.nvirt + idx); break;  // cached entry
        }
        Fe x;
        if (by == BY_ONE) {  // multiplied later without a coefficient: times 2^5
            x = v;
#pragma unroll
            for (int d = 0; d < 5; d++) x = f_add<Fr>(x, x);
        } else if (by == BY_K) {
            x = f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(v), cm[kq]));
        } else {
            x = f_mul_r261<Fr>(by == BY_A ? A : B, v);
        }
        if (minus_v) x = f_sub<Fr>(x, v);
        switch (to) {
            case TO_ACC: acc = f_add<Fr>(acc, x); break;
            case SET_A: A = x; break;
            case ADD_A: A = f_add<Fr>(A, x); break;
            case SET_B: B = x; break;
            default: B = f_add<Fr>(B, x); break;
        }
    }
    if (G.shape == GS_AFFINE_PRODUCT) acc = f_mul_r261<Fr>(A, B);
    return acc;
}

// the cached entries at the reader's point, after its virtual tables
__device__ __forceinline__ void fill_cached(const DProg &p, const Reader &rd) {
    for (uint32_t c = 0; c < p.ncached; c++) rd.set_slot(p.nvirt + c, eval_gate(p, c, rd));
}

// One round: for each pair i, (fold, then) S_i(X) = sum_j w_j gate_j at
// X = 0, 1, 2, and sum_i E(i) S_i(X) into out[0..3).
template <bool FOLD>
__global__ void __launch_bounds__(RTPB) k_gkr_round(DProg p, const Fe *const *src, Fe *const *dst, uint32_t ntab,
                                                  uint32_t nlower, size_t m, F29 r, EqArgs eq, const Fe *rho, int n,
                                                  int b, Fe *slots, Fe *partials, Fe *out) {
    Fe acc[3];
#pragma unroll
    for (int x = 0; x < 3; x++) acc[x] = f_zero<Fr>();
    const size_t stride = (size_t)gridDim.x * blockDim.x, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = tid; i < m; i += stride) {
        if (FOLD) {
            for (uint32_t t = 0; t < ntab; t++) {
                const Fe *s = src[t];
                const Fe a0 = ldfe(s, 4 * i), a1 = ldfe(s, 4 * i + 1), a2 = ldfe(s, 4 * i + 2),
                         a3 = ldfe(s, 4 * i + 3);
                stfe(dst[t], 2 * i, f_add<Fr>(a0, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(f_sub<Fr>(a1, a0)), r))));
                stfe(dst[t], 2 * i + 1, f_add<Fr>(a2, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(f_sub<Fr>(a3, a2)), r))));
            }
        }
        const Fe e = eq_factor(eq, i);
#pragma unroll 1
        for (int x = 0; x < 3; x++) {
            const Reader rd{FOLD ? dst : src, nlower, i, 0, x, false, slots, stride, tid, rho, n, b, false};
            for (uint32_t v = 0; v < p.nvirt; v++)
                rd.set_slot(v, virtual_at_point(p.kinds[v], RoundPoint{rho, n, b, i, x}));
            fill_cached(p, rd);
            Fe s = f_zero<Fr>();
            for (uint32_t g = 0; g < p.ngates; g++)
                s = f_add<Fr>(s, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(eval_gate(p, p.ncached + g, rd)), p.w[g])));
            const Fe es = f_mul<Fr>(e, s);
            if (x == 0) acc[0] = f_add<Fr>(acc[0], es);
            else if (x == 1) acc[1] = f_add<Fr>(acc[1], es);
            else acc[2] = f_add<Fr>(acc[2], es);
        }
    }
    reduce_nodes<3>(acc, 3, partials, out);
}

// The fold of a round's tables alone: work item w folds pair w % m of table
// w / m (rounds too small to fill the device with pairs).
__global__ void __launch_bounds__(RTPB) k_gkr_fold(const Fe *const *src, Fe *const *dst, uint32_t ntab, size_t m,
                                                 F29 r) {
    const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= m * ntab) return;
    const size_t i = w % m;
    const Fe *s = src[w / m];
    Fe *d = dst[w / m];
    const Fe a0 = ldfe(s, 4 * i), a1 = ldfe(s, 4 * i + 1), a2 = ldfe(s, 4 * i + 2), a3 = ldfe(s, 4 * i + 3);
    stfe(d, 2 * i, f_add<Fr>(a0, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(f_sub<Fr>(a1, a0)), r))));
    stfe(d, 2 * i + 1, f_add<Fr>(a2, f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(f_sub<Fr>(a3, a2)), r))));
}

// A round split over (pair, gate) work items: item w evaluates weighted gate
// w / m at pair w % m (after that pair's virtual values and cached entries),
// so a round with few pairs still spreads its gates over the device; the
// sums are the same, the node sums being linear in the gates.
__global__ void __launch_bounds__(RTPB) k_gkr_round_split(DProg p, const Fe *const *tab, uint32_t nlower, size_t m,
                                                        EqArgs eq, const Fe *rho, int n, int b, Fe *slots,
                                                        Fe *partials, Fe *out) {
    Fe acc[3];
#pragma unroll
    for (int x = 0; x < 3; x++) acc[x] = f_zero<Fr>();
    const size_t stride = (size_t)gridDim.x * blockDim.x, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t items = m * p.ngates;
    for (size_t w = tid; w < items; w += stride) {
        const size_t i = w % m;
        const uint32_t g = (uint32_t)(w / m);
        const Fe ew = f29_pack_small<Fr>(f29_mul<Fr>(f29_from32(eq_factor(eq, i)), p.w[g]));
#pragma unroll 1
        for (int x = 0; x < 3; x++) {
            // one gate reads few virtual tables: evaluate them where read
            const Reader rd{tab, nlower, i, 0, x, false, slots, stride, tid, rho, n, b, true};
            fill_cached(p, rd);
            const Fe es = f_mul<Fr>(ew, eval_gate(p, p.ncached + g, rd));
            if (x == 0) acc[0] = f_add<Fr>(acc[0], es);
            else if (x == 1) acc[1] = f_add<Fr>(acc[1], es);
            else acc[2] = f_add<Fr>(acc[2], es);
        }
    }
    reduce_nodes<3>(acc, 3, partials, out);
}

// gkr::produce: the producing gates at every row y
__global__ void __launch_bounds__(RTPB) k_gkr_forward(DProg p, const Fe *const *cols, size_t rows, Fe *slots,
                                                    Fe *const *out, uint32_t nproducing) {
    const size_t stride = (size_t)gridDim.x * blockDim.x, tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t y = tid; y < rows; y += stride) {
        const Reader rd{cols, 0, y, rows, 0, true, slots, stride, tid, nullptr, 0, 0, false};
        for (uint32_t v = 0; v < p.nvirt; v++) rd.set_slot(v, virtual_at_row(p.kinds[v], y));
        fill_cached(p, rd);
        for (uint32_t g = 0; g < nproducing; g++) stfe(out[g], y, eval_gate(p, p.ncached + g, rd));
    }
}

// The program in one device buffer: gates, sources, kinds, then (aligned)
// coefficients and weights.
struct Staged {
    DProg prog;
};

// 2^5 where a coefficient's product is multiplied by another value later
// (the DProg comment): a Product's coefficient, an AffineProduct's left side
// and constant, a Quadratic's product coefficients.
bool coef_scaled(const GkrGate &g, uint32_t j) {
    const uint32_t nv = g.src_end - g.src_begin, t = g.split;
    switch (g.shape) {
        case GS_PRODUCT: return true;
        case GS_AFFINE_PRODUCT: return j <= t;        // a_1..a_t, a_0
        case GS_QUADRATIC: return j >= 1 + t;         // b_1..b_u after c_0, a_1..a_t
        default: (void)nv; return false;
    }
}

Staged stage_program(Context &ctx, const GkrList &l, const Fe *weights, uint32_t nweights) {
    // coefficients as multipliers (R' domain) and as constants, role-scaled
    std::vector<F29> cm(l.coeffs.size());
    std::vector<Fe> cc(l.coeffs.size());
    std::vector<bool> done(l.coeffs.size(), false);
    for (const GkrGate &g : l.gates) {
        const uint32_t nv = g.src_end - g.src_begin;
        uint32_t ncoef = 0;
        switch (g.shape) {
            case GS_LINEAR: ncoef = nv + 1; break;
            case GS_PRODUCT: ncoef = 1; break;
            case GS_AFFINE_PRODUCT: ncoef = nv + 2; break;
            case GS_QUADRATIC: ncoef = 1 + g.split + (nv - g.split) / 2; break;
            default: break;
        }
        for (uint32_t j = 0; j < ncoef; j++) {
            const size_t at = (size_t)g.coef_begin + j;
            Fe c = l.coeffs[at];
            if (coef_scaled(g, j))
                for (int d = 0; d < 5; d++) c = f_add<Fr>(c, c);
            // a slot shared by two gates must have one meaning for both
            if (done[at] && !f_eq(cc[at], c)) throw ApgError(-2, "gkr: a coefficient slot shared across roles");
            cc[at] = c;
            cm[at] = f29_from_mont32<Fr>(c);
            done[at] = true;
        }
    }
    std::vector<F29> w29(nweights);
    for (uint32_t j = 0; j < nweights; j++) w29[j] = f29_from_mont32<Fr>(weights[j]);

    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_gates = 0;
    const size_t o_src = al(o_gates + l.gates.size() * sizeof(GkrGate));
    const size_t o_kinds = al(o_src + l.sources.size() * 4);
    const size_t o_cm = al(o_kinds + l.virtual_kinds.size() * 4);
    const size_t o_cc = al(o_cm + cm.size() * sizeof(F29));
    const size_t o_w = al(o_cc + cc.size() * sizeof(Fe));
    const size_t total = al(o_w + (size_t)nweights * sizeof(F29)) + 256;
    uint8_t *h = ctx.h<uint8_t>("gkr.prog", total);
    uint8_t *d = ctx.d<uint8_t>("gkr.prog", total);
    if (!l.gates.empty()) std::memcpy(h + o_gates, l.gates.data(), l.gates.size() * sizeof(GkrGate));
    if (!l.sources.empty()) std::memcpy(h + o_src, l.sources.data(), l.sources.size() * 4);
    if (!l.virtual_kinds.empty()) std::memcpy(h + o_kinds, l.virtual_kinds.data(), l.virtual_kinds.size() * 4);
    if (!cm.empty()) std::memcpy(h + o_cm, cm.data(), cm.size() * sizeof(F29));
    if (!cc.empty()) std::memcpy(h + o_cc, cc.data(), cc.size() * sizeof(Fe));
    if (nweights) std::memcpy(h + o_w, w29.data(), (size_t)nweights * sizeof(F29));
    HIP_CHECK(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, ctx.stream));
    Staged s;
    s.prog.gates = reinterpret_cast<const GkrGate *>(d + o_gates);
    s.prog.src = reinterpret_cast<const uint32_t *>(d + o_src);
    s.prog.kinds = reinterpret_cast<const uint32_t *>(d + o_kinds);
    s.prog.coefm = reinterpret_cast<const F29 *>(d + o_cm);
    s.prog.coefc = reinterpret_cast<const Fe *>(d + o_cc);
    s.prog.w = reinterpret_cast<const F29 *>(d + o_w);
    s.prog.ncached = l.ncached;
    s.prog.ngates = l.nproducing + l.nenforcing;
    s.prog.nvirt = l.halving ? 0 : (uint32_t)l.virtual_kinds.size();  // a halving list reads none
    return s;
}

Fe eq_lin(const Fe &r, const Fe &x) {  // (1 - r) + x (2r - 1)
    const Fe one = f_one<Fr>();
    return f_add<Fr>(f_sub<Fr>(one, r), f_mul<Fr>(x, f_sub<Fr>(f_add<Fr>(r, r), one)));
}

}  // namespace

int gkr_check_list(const GkrList &l, size_t ncols) {
    const size_t all = (size_t)l.ncached + l.nproducing + l.nenforcing;
    if (l.gates.size() != all) throw ApgError(-2, "gkr: gate count differs from cached + producing + enforcing");
    if (l.halving && (l.ncached || l.nenforcing))
        throw ApgError(-2, "gkr: a halving list has no cached or enforcing entries");
    if (ncols > SRC_MASK || l.virtual_kinds.size() > SRC_MASK || l.ncached > SRC_MASK)
        throw ApgError(-2, "gkr: too many columns");
    for (uint32_t k : l.virtual_kinds)
        if (k > VK_XOR8_OUT) throw ApgError(-2, "gkr: unknown virtual kind");
    std::vector<int> deg(all, 0);
    int maxdeg = 0;
    for (size_t i = 0; i < all; i++) {
        const GkrGate &g = l.gates[i];
        if (g.src_begin > g.src_end || g.src_end > l.sources.size())
            throw ApgError(-2, "gkr: gate operands out of range");
        const uint32_t nv = g.src_end - g.src_begin;
        std::vector<int> od(nv);
        for (uint32_t q = 0; q < nv; q++) {
            const uint32_t s = l.sources[g.src_begin + q], idx = s & SRC_MASK;
            switch (s >> SRC_SHIFT) {
                case SRC_LOWER:
                    if (idx >= ncols) throw ApgError(-2, "gkr: column source out of range");
                    od[q] = 1;
                    break;
                case SRC_UPPER:
                    if (!l.halving || idx >= ncols) throw ApgError(-2, "gkr: child-1 source out of range");
                    od[q] = 1;
                    break;
                case SRC_VIRTUAL:
                    if (l.halving || idx >= l.virtual_kinds.size())
                        throw ApgError(-2, "gkr: virtual source out of range");
                    od[q] = 1;
                    break;
                case SRC_CACHED:
                    if (i < l.ncached || idx >= l.ncached)
                        throw ApgError(-2, "gkr: cached source out of range (or read by a cached entry)");
                    od[q] = deg[idx];
                    break;
                default:
                    throw ApgError(-2, "gkr: unknown source kind");
            }
        }
        auto mx = [&](uint32_t a, uint32_t b) {
            int d = 0;
            for (uint32_t q = a; q < b; q++) d = std::max(d, od[q]);
            return d;
        };
        uint32_t ncoef = 0;
        int d = 0;
        switch (g.shape) {
            case GS_LINEAR: ncoef = nv + 1; d = mx(0, nv); break;
            case GS_PRODUCT:
            case GS_MASK:
            case GS_TREE_PRODUCT:
                if (nv != 2) throw ApgError(-2, "gkr: a two-operand shape with another count");
                ncoef = g.shape == GS_PRODUCT ? 1 : 0;
                d = od[0] + od[1];
                break;
            case GS_AFFINE_PRODUCT:
                if (g.split > nv) throw ApgError(-2, "gkr: AffineProduct split out of range");
                ncoef = nv + 2;
                d = mx(0, g.split) + mx(g.split, nv);
                break;
            case GS_QUADRATIC:
                if (g.split > nv || (nv - g.split) % 2) throw ApgError(-2, "gkr: Quadratic operand count");
                ncoef = 1 + g.split + (nv - g.split) / 2;
                d = mx(0, g.split);
                for (uint32_t q = g.split; q < nv; q += 2) d = std::max(d, od[q] + od[q + 1]);
                break;
            case GS_TREE_CROSS:
                if (nv != 4) throw ApgError(-2, "gkr: TreeCross reads four values");
                d = std::max(od[0] + od[3], od[1] + od[2]);
                break;
            default:
                throw ApgError(-2, "gkr: unknown gate shape");
        }
        if ((size_t)g.coef_begin + ncoef > l.coeffs.size()) throw ApgError(-2, "gkr: coefficients out of range");
        deg[i] = d;
        if (i >= l.ncached) maxdeg = std::max(maxdeg, d);
    }
    return maxdeg;
}

void gkr_forward_device(Context &ctx, const GkrList &l, const Fe *const *cols, size_t ncols, int num_vars,
                        Fe *const *out) {
    (void)gkr_check_list(l, ncols);
    if (num_vars < 0 || num_vars > 40) throw ApgError(-2, "gkr: num_vars out of range");
    if (l.halving && num_vars < 1) throw ApgError(-2, "gkr: a halving list needs num_vars >= 1");
    const size_t rows = l.halving ? size_t(1) << (num_vars - 1) : size_t(1) << num_vars;
    if (!l.nproducing) return;
    Staged s = stage_program(ctx, l, nullptr, 0);
    const size_t np = std::max<size_t>(ncols, 1) + l.nproducing;
    const Fe **hp = ctx.h<const Fe *>("gkr.fptr", np);
    const Fe **dp = ctx.d<const Fe *>("gkr.fptr", np);
    for (size_t c = 0; c < ncols; c++) hp[c] = cols[c];
    const size_t oo = std::max<size_t>(ncols, 1);
    for (uint32_t g = 0; g < l.nproducing; g++) hp[oo + g] = out[g];
    HIP_CHECK(hipMemcpyAsync(dp, hp, np * sizeof(Fe *), hipMemcpyHostToDevice, ctx.stream));
    const unsigned grid = (unsigned)std::min<size_t>(grid_for(rows, RTPB), (size_t)MAX_RBLOCKS);
    const size_t nslots = (size_t)s.prog.nvirt + l.ncached;
    Fe *slots = ctx.d<Fe>("gkr.slots", std::max<size_t>(nslots * grid * RTPB, 1));
    hipLaunchKernelGGL(k_gkr_forward, dim3(grid), dim3(RTPB), 0, ctx.stream, s.prog, dp, rows, slots,
                       reinterpret_cast<Fe *const *>(const_cast<Fe **>(dp + oo)), l.nproducing);
    HIP_CHECK(hipGetLastError());
    ctx.sync();
}

void gkr_prove_sumcheck_device(Context &ctx, const GkrList &l, const Fe *weights, Fe *const *lower,
                               Fe *const *upper, size_t ncols, int n, const Fe *eq_point,
                               const RoundFn &round_fn, Fe *rounds_out, Fe *point_out, Fe *finals_out) {
    const int deg = gkr_check_list(l, ncols);
    if (deg > 2) throw ApgError(-2, "gkr: a gate of degree above 2 (the artifact's degree law)");
    if (n < 0 || n > 40) throw ApgError(-2, "gkr: num_vars out of range");
    hipStream_t st = ctx.stream;
    const size_t nup = l.halving ? ncols : 0, nt = ncols + nup;
    if (nt == 0) throw ApgError(-2, "gkr: no tables");
    std::vector<Fe *> T(nt);
    for (size_t c = 0; c < ncols; c++) T[c] = lower[c];
    for (size_t c = 0; c < nup; c++) T[ncols + c] = upper[c];
    if (n == 0) {
        for (size_t t = 0; t < nt; t++)
            HIP_CHECK(hipMemcpyAsync(&finals_out[t], T[t], sizeof(Fe), hipMemcpyDeviceToHost, st));
        ctx.sync();
        return;
    }
    const uint32_t nw = l.nproducing + l.nenforcing;
    Staged s = stage_program(ctx, l, weights, nw);
    const size_t len0 = size_t(1) << n;

    // suffix-eq tables split at h0 (as the gate sumcheck)
    const int h0 = std::max(1, std::min(n, (n + 1) / 2));
    const size_t astride = size_t(1) << (h0 - 1), bstride = size_t(1) << (n - h0);
    const size_t asz = (size_t)(h0 - 1) * astride, bsz = (size_t)(n - h0 + 1) * bstride;
    Fe *rd = ctx.d<Fe>("gkr.r", n);
    Fe *eqA = ctx.d<Fe>("gkr.eqA", std::max<size_t>(asz, 1));
    Fe *eqB = ctx.d<Fe>("gkr.eqB", bsz);
    Fe *hr = ctx.h<Fe>("gkr.r", n);
    std::memcpy(hr, eq_point, n * sizeof(Fe));
    HIP_CHECK(hipMemcpyAsync(rd, hr, n * sizeof(Fe), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_eq_suffix, dim3(grid_for(asz + bsz, 256)), dim3(256), 0, st, rd, n, h0, eqA, eqB);
    HIP_CHECK(hipGetLastError());

    // ping-pong pools and the three pointer arrays (inputs, A, B)
    const size_t ha = std::max<size_t>(len0 / 2, 1), hb = std::max<size_t>(len0 / 4, 1);
    Fe *poolA = ctx.d<Fe>("gkr.poolA", nt * ha);
    Fe *poolB = ctx.d<Fe>("gkr.poolB", nt * hb);
    Fe **hp = ctx.h<Fe *>("gkr.ptr", 3 * nt);
    Fe **dp = ctx.d<Fe *>("gkr.ptr", 3 * nt);
    for (size_t t = 0; t < nt; t++) {
        hp[t] = T[t];
        hp[nt + t] = poolA + t * ha;
        hp[2 * nt + t] = poolB + t * hb;
    }
    HIP_CHECK(hipMemcpyAsync(dp, hp, 3 * nt * sizeof(Fe *), hipMemcpyHostToDevice, st));
    Fe *rho_d = ctx.d<Fe>("gkr.rho", n);
    Fe *rho_h = ctx.h<Fe>("gkr.rho", n);

    // per-thread slots for the virtual tables and cached entries at a point
    const size_t nslots = (size_t)s.prog.nvirt + l.ncached;
    Fe *slots = ctx.d<Fe>("gkr.slots", std::max<size_t>(nslots * MAX_RBLOCKS * RTPB, 1));
    Fe *partials = ctx.d<Fe>("gkr.partials", (size_t)MAX_RBLOCKS * 3);
    // the round kernel (one block) or its finishing pass writes the node sums
    // straight into pinned host memory; the stream sync after each round
    // makes them visible
    Fe *hvals = ctx.h<Fe>("gkr.hvals", 3);
    Fe *hvals_dev = nullptr;
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void **>(&hvals_dev), hvals, 0));

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

    const Fe inv2 = f_inv<Fr>(f_from_u64<Fr>(2));
    Fe rho = f_zero<Fr>(), P = f_one<Fr>();
    Fe *const *cur = dp;  // round j's tables before the fold
    for (int j = 0; j < n; j++) {
        const bool fold = j > 0;
        const size_t m = len0 >> (j + 1);
        // round 0 reads the inputs; round j folds into A (odd j) or B (even j)
        Fe *const *src = j == 0 ? dp : (j % 2 == 1 ? (j == 1 ? dp : dp + 2 * nt) : dp + nt);
        Fe *const *dst = j % 2 == 1 ? dp + nt : dp + 2 * nt;
        if (fold && j >= 1) {
            rho_h[j - 1] = rho;
            HIP_CHECK(hipMemcpyAsync(rho_d + (j - 1), rho_h + (j - 1), sizeof(Fe), hipMemcpyHostToDevice, st));
        }
        const int sfx = j + 1;
        EqArgs eq;
        if (sfx < h0) {
            eq.lo = eqA + (size_t)(sfx - 1) * astride;
            eq.lo_bits = (uint32_t)(h0 - sfx);
            eq.hi = eqB;
        } else {
            eq.lo = nullptr;
            eq.lo_bits = 0;
            eq.hi = eqB + (size_t)(sfx - h0) * bstride;
        }
        const size_t cap = std::min<size_t>(MAX_RBLOCKS, (size_t)4 * ctx.cu_count);
        const dim3 g((unsigned)std::min(cap, std::max<size_t>(1, (m + RTPB - 1) / RTPB))), b(RTPB);
        if (ctx.profiling) HIP_CHECK(hipEventRecord(ev0, st));
        const F29 rho29 = f29_from_mont32<Fr>(rho);  // the fold's challenge in the R' domain
        if (m < cap * RTPB / 8 && s.prog.ngates > 1) {
            // fewer pairs than an eighth of the resident threads: fold, then
            // spread the (pair, gate) items over the device
            if (fold)
                hipLaunchKernelGGL(k_gkr_fold, dim3(grid_for(m * nt, RTPB)), b, 0, st, src, dst, (uint32_t)nt, m,
                                   rho29);
            const size_t items = m * s.prog.ngates;
            const dim3 gs((unsigned)std::min(cap, std::max<size_t>(1, (items + RTPB - 1) / RTPB)));
            hipLaunchKernelGGL(k_gkr_round_split, gs, b, 0, st, s.prog, fold ? dst : src, (uint32_t)ncols, m, eq,
                               rho_d, n, j, slots, partials, hvals_dev);
            launch_finish<3>(st, gs.x, 3, partials, hvals_dev, nullptr);
        } else if (fold) {
            hipLaunchKernelGGL(k_gkr_round<true>, g, b, 0, st, s.prog, src, dst, (uint32_t)nt, (uint32_t)ncols, m,
                               rho29, eq, rho_d, n, j, slots, partials, hvals_dev);
            launch_finish<3>(st, g.x, 3, partials, hvals_dev, nullptr);
        } else {
            hipLaunchKernelGGL(k_gkr_round<false>, g, b, 0, st, s.prog, src, dst, (uint32_t)nt, (uint32_t)ncols, m,
                               rho29, eq, rho_d, n, j, slots, partials, hvals_dev);
            launch_finish<3>(st, g.x, 3, partials, hvals_dev, nullptr);
        }
        HIP_CHECK(hipGetLastError());
        if (ctx.profiling) HIP_CHECK(hipEventRecord(ev1, st));
        round_fn.begin(4);  // the transcript's message prefix, while the kernel runs
        ctx.sync();
        if (ctx.profiling) {
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
            ctx.stats.sumcheck_kernel_ms += ms;
            ctx.stats.sumcheck_launches += 1;
            // every table read once (a pair's 2 entries, or 4 of the
            // previous round's table and the 2 folded entries written)
            ctx.stats.sumcheck_bytes += (uint64_t)nt * 32 * (fold ? 6 * m : 2 * m);
        }
        // S(X) = s0 + s1 X + s2 X^2 from S(0), S(1), S(2); then
        // g(X) = P eq(r_j, X) S(X) with eq(r_j, X) = (1 - r_j) + X (2 r_j - 1)
        const Fe s0 = hvals[0];
        const Fe s2 = f_mul<Fr>(f_add<Fr>(f_sub<Fr>(hvals[2], f_add<Fr>(hvals[1], hvals[1])), hvals[0]), inv2);
        const Fe s1 = f_sub<Fr>(f_sub<Fr>(hvals[1], hvals[0]), s2);
        const Fe &rj = eq_point[j];
        const Fe one = f_one<Fr>();
        const Fe a = f_mul<Fr>(P, f_sub<Fr>(one, rj));
        const Fe bb = f_mul<Fr>(P, f_sub<Fr>(f_add<Fr>(rj, rj), one));
        Fe *c = rounds_out + (size_t)j * 4;
        c[0] = f_mul<Fr>(a, s0);
        c[1] = f_add<Fr>(f_mul<Fr>(a, s1), f_mul<Fr>(bb, s0));
        c[2] = f_add<Fr>(f_mul<Fr>(a, s2), f_mul<Fr>(bb, s1));
        c[3] = f_mul<Fr>(bb, s2);
        rho = round_fn(c, 4);
        point_out[j] = rho;
        P = f_mul<Fr>(P, eq_lin(rj, rho));
        cur = fold ? dst : dp;
    }
    // the last round's two-entry tables, bound on the host
    std::vector<Fe *> last(nt);
    std::vector<Fe> two(2 * nt);
    for (size_t t = 0; t < nt; t++) last[t] = n == 1 ? T[t] : hp[(cur == dp + nt ? nt : 2 * nt) + t];
    for (size_t t = 0; t < nt; t++)
        HIP_CHECK(hipMemcpyAsync(&two[2 * t], last[t], 2 * sizeof(Fe), hipMemcpyDeviceToHost, st));
    ctx.sync();
    for (size_t t = 0; t < nt; t++) {
        const Fe lo = two[2 * t], hi = two[2 * t + 1];
        finals_out[t] = f_add<Fr>(lo, f_mul<Fr>(rho, f_sub<Fr>(hi, lo)));
    }
}

}  // namespace apg
