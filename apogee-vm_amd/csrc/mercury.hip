// The Mercury opening, pcs::open (crates/pcs/src/lib.rs:313-492), for gfx950.
//
// The size-n work runs in HBM: the restriction h (one block per row dot with
// eq(u1, .)), the fold by X^b - alpha, the numerator f - (z^b - alpha) q - g(z)
// and its synthetic division by X - z, and the MSMs of every commitment (q and
// the fold's opening have n - b and n - 1 coefficients). Both the fold and the
// division are Horner recurrences A_j = c_j + x A_{j+1}; they run as a
// segmented scan: each thread runs one segment of SEG elements from the top,
// the segments' bottom values form the same recurrence with multiplier x^SEG
// (solved recursively), and a last pass adds x^(top - j) times the value that
// enters each segment from above. The size-b algebra (the BDFG20 batch of
// crates/pcs-verify/src/bdfg.rs, the univariate helpers of
// crates/pcs-verify/src/uni.rs, the six evaluations) stays on the host; the
// symmetric witness product g rev(P_u1) + gamma h rev(P_u2)
// (crates/pcs/src/lib.rs:812-866) is one block per output coefficient. The
// product of two polynomials of degree < b is unique, so this direct form
// returns the coefficients the reference's size-2b FFT does.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/apogee_gpu.h"
#include "curve.cuh"
#include "mercury.h"

namespace apg {

namespace {

constexpr int MTPB = 256;
constexpr uint32_t SEG = 16;

struct Powers {
    Fe p[SEG + 1];  // x^0 .. x^SEG
};

__device__ __forceinline__ Fe block_sum_fr(Fe v, Fe *lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        Fe o;
#pragma unroll
        for (int q = 0; q < 8; q++) o.l[q] = __shfl_xor(v.l[q], off, 64);
        v = f_add<Fr>(v, o);
    }
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) lds[tid >> 6] = v;
    __syncthreads();
    Fe s = lds[0];
    for (int w = 1; w < MTPB / 64; w++) s = f_add<Fr>(s, lds[w]);
    return s;
}

// h[j] = sum_i eq1[i] f[j b + i] (crates/pcs/src/lib.rs:353-357), one block per row
__global__ void __launch_bounds__(MTPB) k_row_dot(const Fe *__restrict__ f, const Fe *__restrict__ eq1,
                                                uint32_t b, Fe *__restrict__ h) {
    __shared__ Fe lds[MTPB / 64];
    const size_t row = (size_t)blockIdx.x * b;
    Fe acc = f_zero<Fr>();
    for (uint32_t i = threadIdx.x; i < b; i += MTPB) acc = f_add<Fr>(acc, f_mul<Fr>(eq1[i], f[row + i]));
    const Fe s = block_sum_fr(acc, lds);
    if (threadIdx.x == 0) h[blockIdx.x] = s;
}

// R[k] = sum_i g[i] r1[k - i] + gamma h[i] r2[k - i] for k < 2b - 1, one block per k
__global__ void __launch_bounds__(MTPB) k_conv2(const Fe *__restrict__ g, const Fe *__restrict__ r1,
                                              const Fe *__restrict__ h, const Fe *__restrict__ r2, Fe gamma,
                                              uint32_t b, Fe *__restrict__ R) {
    __shared__ Fe lds[MTPB / 64];
    const uint32_t k = blockIdx.x;
    const uint32_t lo = k >= b ? k - b + 1 : 0, hi = k < b - 1 ? k : b - 1;
    Fe acc = f_zero<Fr>();
    for (uint32_t i = lo + threadIdx.x; i <= hi; i += MTPB) {
        const Fe a = f_mul<Fr>(g[i], r1[k - i]);
        const Fe c = f_mul<Fr>(h[i], r2[k - i]);
        acc = f_add<Fr>(acc, f_add<Fr>(a, f_mul<Fr>(gamma, c)));
    }
    const Fe s = block_sum_fr(acc, lds);
    if (threadIdx.x == 0) R[k] = s;
}

// One thread per (segment s, column i): the segment's rows [s SEG, hi) from
// the top with nothing entering from above; its bottom value to carry.
__global__ void k_seg_local(Fe *__restrict__ v, size_t count, size_t stride, size_t ncols, Fe x,
                            Fe *__restrict__ carry) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nseg = (count + SEG - 1) / SEG;
    if (t >= nseg * ncols) return;
    const size_t s = t / ncols, i = t - s * ncols;
    const size_t lo = s * SEG, hi = lo + SEG < count ? lo + SEG : count;
    Fe acc = f_zero<Fr>();
    for (size_t j = hi; j-- > lo;) {
        Fe *p = v + j * stride + i;
        acc = f_add<Fr>(*p, f_mul<Fr>(x, acc));
        *p = acc;
    }
    carry[s * ncols + i] = acc;
}

// A_j = L_j + x^(top(s) - j) B_{s+1}, B the solved carries; the top segment
// has nothing above it and is final already.
__global__ void k_seg_fix(Fe *__restrict__ v, size_t count, size_t stride, size_t ncols, Powers xp,
                          const Fe *__restrict__ B) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * ncols) return;
    const size_t j = t / ncols, i = t - j * ncols;
    const size_t s = j / SEG, nseg = (count + SEG - 1) / SEG;
    if (s + 1 >= nseg) return;
    const size_t top = (s + 1) * SEG;
    Fe *p = v + j * stride + i;
    *p = f_add<Fr>(*p, f_mul<Fr>(xp.p[top - j], B[(s + 1) * ncols + i]));
}

// num = f - (z^b - alpha) q - g(z) [X^0] over the fold's quotient rows q
// (the first (b - 1) b coefficients), f beyond them (crates/pcs/src/lib.rs:411-417)
__global__ void k_numerator(const Fe *__restrict__ f, const Fe *__restrict__ q, size_t n, size_t nq,
                            Fe zb_a, Fe g_z, Fe *__restrict__ num) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fe val = f[i];
    if (i < nq) val = f_sub<Fr>(val, f_mul<Fr>(zb_a, q[i]));
    if (i == 0) val = f_sub<Fr>(val, g_z);
    num[i] = val;
}

// eq(u, i) = prod_j (bit j of i ? u_j : 1 - u_j), the entries of eq_table(u)
// (crates/poly/src/lib.rs:233-248) in closed form
__global__ void k_eq_point(const Fe *__restrict__ u, int width, size_t len, Fe *__restrict__ eq) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const Fe one = f_one<Fr>();
    Fe acc = one;
    for (int j = 0; j < width; j++) acc = f_mul<Fr>(acc, ((i >> j) & 1) ? u[j] : f_sub<Fr>(one, u[j]));
    eq[i] = acc;
}

// partial sums of x . y, one per block (grid-stride)
__global__ void __launch_bounds__(MTPB) k_dot_partial(const Fe *__restrict__ x, const Fe *__restrict__ y, size_t len,
                                                    Fe *__restrict__ partials) {
    __shared__ Fe lds[MTPB / 64];
    Fe acc = f_zero<Fr>();
    for (size_t i = (size_t)blockIdx.x * MTPB + threadIdx.x; i < len; i += (size_t)gridDim.x * MTPB)
        acc = f_add<Fr>(acc, f_mul<Fr>(x[i], y[i]));
    const Fe s = block_sum_fr(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// f*(i) = sum_s w_s stack_s[slot](row), slot = i / column, row = i % column
// (crates/pcs/src/lib.rs:769-780); stack s's columns start at column firsts[s]
__global__ void k_stack_combine(const Fe *__restrict__ cols, size_t column, size_t n,
                                const uint32_t *__restrict__ firsts, const uint32_t *__restrict__ sizes,
                                const Fe *__restrict__ w, uint32_t nstacks, Fe *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t slot = i / column, row = i - slot * column;
    Fe acc = f_zero<Fr>();
    for (uint32_t s = 0; s < nstacks; s++)
        if (slot < sizes[s]) acc = f_add<Fr>(acc, f_mul<Fr>(w[s], cols[((size_t)firsts[s] + slot) * column + row]));
    out[i] = acc;
}

// ---- host algebra over Fr (Montgomery Fe, the 32-bit-word host build) ----
using Poly = std::vector<Fe>;

Fe one() { return f_one<Fr>(); }
Fe zero() { return f_zero<Fr>(); }
Fe add(const Fe &a, const Fe &b) { return f_add<Fr>(a, b); }
Fe sub(const Fe &a, const Fe &b) { return f_sub<Fr>(a, b); }
Fe mul(const Fe &a, const Fe &b) { return f_mul<Fr>(a, b); }
Fe neg(const Fe &a) { return f_neg<Fr>(a); }
Fe inv(const Fe &a) { return f_inv<Fr>(a); }

// uni.rs:18-24
Fe eval(const Poly &c, const Fe &x) {
    Fe acc = zero();
    for (size_t i = c.size(); i-- > 0;) acc = add(mul(acc, x), c[i]);
    return acc;
}

// uni.rs:32-44
Poly div_by_linear(const Poly &c, const Fe &r, Fe &rem) {
    if (c.empty()) {
        rem = zero();
        return {};
    }
    Poly q(c.size() - 1);
    Fe acc = c.back();
    for (size_t i = c.size() - 1; i-- > 0;) {
        q[i] = acc;
        acc = add(c[i], mul(acc, r));
    }
    rem = acc;
    return q;
}

// uni.rs:47-58
Poly pmul(const Poly &a, const Poly &b) {
    if (a.empty() || b.empty()) return {};
    Poly o(a.size() + b.size() - 1, zero());
    for (size_t i = 0; i < a.size(); i++)
        for (size_t j = 0; j < b.size(); j++) o[i + j] = add(o[i + j], mul(a[i], b[j]));
    return o;
}

// uni.rs:61-68
void add_scaled(Poly &acc, const Poly &src, const Fe &c) {
    if (src.size() > acc.size()) acc.resize(src.size(), zero());
    for (size_t i = 0; i < src.size(); i++) acc[i] = add(acc[i], mul(c, src[i]));
}

// uni.rs:71-77
Poly vanishing(const std::vector<Fe> &roots) {
    Poly out{one()};
    for (const Fe &r : roots) out = pmul(out, Poly{neg(r), one()});
    return out;
}

// uni.rs:84-102
Poly interpolate(const std::vector<Fe> &xs, const std::vector<Fe> &ys) {
    Poly out(xs.size(), zero());
    for (size_t i = 0; i < xs.size(); i++) {
        Poly num{one()};
        Fe den = one();
        for (size_t j = 0; j < xs.size(); j++) {
            if (i == j) continue;
            num = pmul(num, Poly{neg(xs[j]), one()});
            den = mul(den, sub(xs[i], xs[j]));
        }
        add_scaled(out, num, mul(ys[i], inv(den)));
    }
    return out;
}

// uni.rs:105-118
Fe pow_usize(Fe x, size_t e) {
    Fe acc = one();
    while (e) {
        if (e & 1) acc = mul(acc, x);
        x = mul(x, x);
        e >>= 1;
    }
    return acc;
}

// crates/poly/src/lib.rs:233-248 eq_table by doubling, variable j = bit j
Poly eq_table(const Fe *r, size_t t) {
    Poly e((size_t)1 << t);
    e[0] = one();
    for (size_t j = 0; j < t; j++) {
        const size_t half = (size_t)1 << j;
        for (size_t i = 0; i < half; i++) {
            const Fe w = mul(e[i], r[j]);
            e[i + half] = w;
            e[i] = sub(e[i], w);
        }
    }
    return e;
}

// pcs-verify/src/lib.rs:186-196
Fe tensor_eval(const Fe *u, size_t t, const Fe &x) {
    Fe acc = one(), power = x;
    for (size_t k = 0; k < t; k++) {
        acc = mul(acc, add(mul(u[k], power), sub(one(), u[k])));
        power = mul(power, power);
    }
    return acc;
}

// pcs-verify/src/lib.rs:205-227
Fe derive_h_alpha(const Fe *u1, const Fe *u2, size_t t, const Fe &z, const Fe &z_inv, const Fe &gamma,
                  const Fe &v, const Fe ev[6]) {
    const Fe two_inv = inv(f_from_u64<Fr>(2));
    Fe inner = add(mul(ev[0], tensor_eval(u1, t, z_inv)), mul(ev[1], tensor_eval(u1, t, z)));
    Fe hp = add(mul(ev[2], tensor_eval(u2, t, z_inv)), mul(ev[3], tensor_eval(u2, t, z)));
    hp = sub(sub(hp, v), v);
    inner = add(inner, mul(gamma, hp));
    inner = sub(sub(inner, mul(z, ev[4])), mul(z_inv, ev[5]));
    return mul(inner, two_inv);
}

struct Item {
    Poly zc, r;
};

// bdfg.rs:94-118
Poly bdfg_quotient(const Poly *polys[4], const Item it[4], const Fe t_set[3], const Fe &delta) {
    Poly f;
    Fe power = one();
    for (int i = 0; i < 4; i++) {
        Poly shifted = *polys[i];
        add_scaled(shifted, it[i].r, neg(one()));
        add_scaled(f, pmul(it[i].zc, shifted), power);
        power = mul(power, delta);
    }
    for (int k = 0; k < 3; k++) {
        Fe rem;
        f = div_by_linear(f, t_set[k], rem);
        if (!f_is_zero(rem)) throw ApgError(APG_ERR_INTERNAL, "mercury: F is not divisible by Z_T");
    }
    return f;
}

// bdfg.rs:122-142
Poly bdfg_linearization(const Poly *polys[4], const Item it[4], const Fe t_set[3], const Poly &w,
                        const Fe &delta, const Fe &zp) {
    Poly l;
    Fe constant = zero();
    add_scaled(l, w, neg(eval(vanishing({t_set[0], t_set[1], t_set[2]}), zp)));
    for (int i = 0; i < 4; i++) {
        const Fe c = mul(pow_usize(delta, (size_t)i), eval(it[i].zc, zp));
        add_scaled(l, *polys[i], c);
        constant = add(constant, mul(c, eval(it[i].r, zp)));
    }
    if (l.empty()) l.push_back(zero());
    l[0] = sub(l[0], constant);
    return l;
}

unsigned grid(size_t n) { return grid_for(n, MTPB); }

void upload(Context &ctx, Fe *dst, const Poly &src) {
    if (!src.empty())
        HIP_CHECK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(Fe), hipMemcpyHostToDevice, ctx.stream));
}

void download(Context &ctx, Poly &dst, const Fe *src, size_t n) {
    dst.resize(n);
    if (n) HIP_CHECK(hipMemcpyAsync(dst.data(), src, n * sizeof(Fe), hipMemcpyDeviceToHost, ctx.stream));
    ctx.sync();
}

}  // namespace

void horner_scan(Context &ctx, Fe *v, size_t count, size_t stride, size_t ncols, const Fe &x, int level) {
    if (count == 0 || ncols == 0) return;
    const size_t nseg = (count + SEG - 1) / SEG;
    const std::string name = "horner.carry" + std::to_string(level);
    Fe *carry = ctx.d<Fe>(name.c_str(), nseg * ncols);
    hipLaunchKernelGGL(k_seg_local, dim3(grid(nseg * ncols)), dim3(MTPB), 0, ctx.stream, v, count, stride,
                       ncols, x, carry);
    HIP_CHECK(hipGetLastError());
    if (nseg == 1) return;
    Powers xp;
    xp.p[0] = one();
    for (uint32_t k = 1; k <= SEG; k++) xp.p[k] = mul(xp.p[k - 1], x);
    // the segments' bottom values: B_s = C_s + x^SEG B_{s+1}
    horner_scan(ctx, carry, nseg, ncols, ncols, xp.p[SEG], level + 1);
    hipLaunchKernelGGL(k_seg_fix, dim3(grid(count * ncols)), dim3(MTPB), 0, ctx.stream, v, count, stride, ncols,
                       xp, carry);
    HIP_CHECK(hipGetLastError());
}

void mercury_open(Context &ctx, const CommitFn &commit, const Fe *f_dev, int num_vars, const uint8_t cm_wire[64],
                  const Fe *u, Transcript &tr, Fe &v, uint8_t proof[MERCURY_PROOF_BYTES]) {
    // pcs-verify/src/lib.rs:97-103 check_num_vars
    if (num_vars < 2 || num_vars > 30 || (num_vars & 1))
        throw ApgError(APG_ERR_ARG, "mercury: num_vars must be even and in [2, 30]");
    const size_t n = (size_t)1 << num_vars, t = (size_t)num_vars / 2, b = (size_t)1 << t;
    uint8_t pts[8][64];

    // step 1: h, v, [h]
    const Poly eq1 = eq_table(u, t), eq2 = eq_table(u + t, t);
    Fe *eq_dev = ctx.d<Fe>("mercury.eq", 4 * b);  // eq1, eq2, rev(eq1), rev(eq2)
    Poly rev1(eq1.rbegin(), eq1.rend()), rev2(eq2.rbegin(), eq2.rend());
    upload(ctx, eq_dev, eq1);
    upload(ctx, eq_dev + b, eq2);
    upload(ctx, eq_dev + 2 * b, rev1);
    upload(ctx, eq_dev + 3 * b, rev2);
    Fe *h_dev = ctx.d<Fe>("mercury.h", b);
    hipLaunchKernelGGL(k_row_dot, dim3((unsigned)b), dim3(MTPB), 0, ctx.stream, f_dev, eq_dev, (uint32_t)b, h_dev);
    HIP_CHECK(hipGetLastError());
    Poly h;
    download(ctx, h, h_dev, b);
    v = zero();
    for (size_t i = 0; i < b; i++) v = add(v, mul(eq2[i], h[i]));
    commit(h_dev, b, pts[0]);

    const Fe nf = f_from_u64<Fr>((uint64_t)n);
    tr.append_scalars(TAG_MERCURY_INSTANCE, &nf, 1);
    tr.append_g1_wire(TAG_COMMITMENT, cm_wire, 1);
    std::vector<Fe> claim(u, u + num_vars);
    claim.push_back(v);
    tr.append_scalars(TAG_EVALUATION_CLAIM, claim.data(), claim.size());
    tr.append_g1_wire(TAG_PCS_OPENING, pts[0], 1);
    const Fe alpha = tr.challenge(TAG_MERCURY_ALPHA);

    // step 2: the fold; A holds g (row 0) then q (rows 1..b-1) in place
    Fe *A = ctx.d<Fe>("mercury.fold", n);
    HIP_CHECK(hipMemcpyAsync(A, f_dev, n * sizeof(Fe), hipMemcpyDeviceToDevice, ctx.stream));
    horner_scan(ctx, A, b, b, b, alpha);
    const Fe *g_dev = A, *q_dev = A + b;
    commit(q_dev, (b - 1) * b, pts[1]);
    commit(g_dev, b, pts[2]);
    tr.append_g1_wire(TAG_PCS_OPENING, pts[1], 2);
    const Fe gamma = tr.challenge(TAG_MERCURY_GAMMA);

    // step 3: S and D
    Poly g;
    download(ctx, g, g_dev, b);
    const Fe h_alpha = eval(h, alpha);
    Fe *R_dev = ctx.d<Fe>("mercury.R", 2 * b);
    hipLaunchKernelGGL(k_conv2, dim3((unsigned)(2 * b - 1)), dim3(MTPB), 0, ctx.stream, g_dev, eq_dev + 2 * b,
                       h_dev, eq_dev + 3 * b, gamma, (uint32_t)b, R_dev);
    HIP_CHECK(hipGetLastError());
    Poly R;
    download(ctx, R, R_dev, 2 * b - 1);
    if (!f_eq(R[b - 1], add(h_alpha, mul(gamma, v))))
        throw ApgError(APG_ERR_INTERNAL, "mercury: the symmetric witness's constant term is off");
    Poly s(b - 1), d(b);
    for (size_t k = b; k < 2 * b - 1; k++) s[k - b] = add(R[k], R[2 * b - 2 - k]);
    for (size_t i = 0; i < b; i++) d[i] = g[b - 1 - i];
    Fe *sd_dev = ctx.d<Fe>("mercury.sd", 2 * b);
    upload(ctx, sd_dev, s);
    upload(ctx, sd_dev + b, d);
    commit(sd_dev, b - 1, pts[3]);
    commit(sd_dev + b, b, pts[4]);
    tr.append_g1_wire(TAG_PCS_OPENING, pts[3], 2);
    Fe z;
    do {
        z = tr.challenge(TAG_MERCURY_Z);  // pcs-verify/src/lib.rs:156-169
    } while (f_is_zero(z));
    // pcs-verify/src/lib.rs:176-178
    if (f_is_zero(z) || f_eq(mul(z, z), one()) || f_eq(z, alpha) || f_eq(mul(z, alpha), one()))
        throw ApgError(APG_ERR_DEGENERATE, "mercury: degenerate challenge");
    const Fe z_inv = inv(z);

    // step 4: the six values, then the fold's opening at z
    Fe ev[6] = {eval(g, z), eval(g, z_inv), eval(h, z), eval(h, z_inv), eval(s, z), eval(s, z_inv)};
    tr.append_scalars(TAG_PCS_OPENING, ev, 6);
    Fe *num = ctx.d<Fe>("mercury.num", n);
    hipLaunchKernelGGL(k_numerator, dim3(grid(n)), dim3(MTPB), 0, ctx.stream, f_dev, q_dev, n, (b - 1) * b,
                       sub(pow_usize(z, b), alpha), ev[0], num);
    HIP_CHECK(hipGetLastError());
    horner_scan(ctx, num, n, 1, 1, z);  // num[0] = remainder, num[1..] = quotient
    Poly rem;
    download(ctx, rem, num, 1);
    if (!f_is_zero(rem[0])) throw ApgError(APG_ERR_INTERNAL, "mercury: the fold identity does not vanish at z");
    commit(num + 1, n - 1, pts[5]);
    tr.append_g1_wire(TAG_PCS_OPENING, pts[5], 1);

    // step 4(e): BDFG20 over T = {z, 1/z, alpha} (bdfg.rs:63-86)
    const Fe h_alpha2 = derive_h_alpha(u, u + t, t, z, z_inv, gamma, v, ev);
    const Fe d_z = mul(pow_usize(z, b - 1), ev[1]);
    if (!f_eq(h_alpha2, h_alpha)) throw ApgError(APG_ERR_INTERNAL, "mercury: h(alpha) does not re-derive");
    Item it[4];
    it[0] = {vanishing({alpha}), interpolate({z, z_inv}, {ev[0], ev[1]})};
    it[1] = {Poly{one()}, interpolate({z, z_inv, alpha}, {ev[2], ev[3], h_alpha2})};
    it[2] = {vanishing({alpha}), interpolate({z, z_inv}, {ev[4], ev[5]})};
    it[3] = {vanishing({z_inv, alpha}), Poly{d_z}};
    const Fe t_set[3] = {z, z_inv, alpha};
    const Poly *polys[4] = {&g, &h, &s, &d};
    const Fe delta = tr.challenge(TAG_BDFG_BATCH);
    const Poly w = bdfg_quotient(polys, it, t_set, delta);
    Fe *w_dev = ctx.d<Fe>("mercury.w", std::max<size_t>(w.size(), 1));
    upload(ctx, w_dev, w);
    commit(w_dev, w.size(), pts[6]);
    tr.append_g1_wire(TAG_PCS_OPENING, pts[6], 1);
    const Fe zp = tr.challenge(TAG_BDFG_POINT);
    const Poly l = bdfg_linearization(polys, it, t_set, w, delta, zp);
    Fe lrem;
    const Poly wp = div_by_linear(l, zp, lrem);
    if (!f_is_zero(lrem)) throw ApgError(APG_ERR_INTERNAL, "mercury: the linearization does not vanish at z'");
    Fe *wp_dev = ctx.d<Fe>("mercury.wp", std::max<size_t>(wp.size(), 1));
    upload(ctx, wp_dev, wp);
    commit(wp_dev, wp.size(), pts[7]);
    tr.append_g1_wire(TAG_PCS_OPENING, pts[7], 1);
    (void)tr.challenge(TAG_PAIRING_MERGE);

    // MercuryProof::to_bytes (crates/pcs/src/lib.rs:152-161)
    std::memcpy(proof, pts, sizeof(pts));
    for (int i = 0; i < 6; i++) fe_to_bytes_fr(ev[i], proof + 8 * 64 + 32 * i);
}

void mercury_batch_open_stacked(Context &ctx, const CommitFn &commit, const Fe *cols_dev,
                                const uint32_t *stack_sizes, size_t nstacks, int width, const uint8_t *cms_wire,
                                const Fe *u, const Fe *r, int nr, Transcript &tr, Fe *vs_out,
                                uint8_t proof[MERCURY_PROOF_BYTES]) {
    if (nstacks == 0) throw ApgError(APG_ERR_ARG, "EmptyBatch");
    for (size_t s = 0; s < nstacks; s++)
        if (stack_sizes[s] == 0) throw ApgError(APG_ERR_ARG, "EmptyBatch");
    const int nv = width + nr;
    if (width < 0 || nr < 0 || nv < 2 || nv > 30 || (nv & 1))
        throw ApgError(APG_ERR_ARG, "UnsupportedNumVars");
    const size_t slots = (size_t)1 << nr, column = (size_t)1 << width, n = (size_t)1 << nv;
    std::vector<uint32_t> firsts(nstacks);
    uint32_t ncols = 0;
    for (size_t s = 0; s < nstacks; s++) {
        if (stack_sizes[s] > slots) throw ApgError(APG_ERR_ARG, "BatchLengthMismatch");
        firsts[s] = ncols;
        ncols += stack_sizes[s];
    }
    std::vector<Fe> point(u, u + width);
    point.insert(point.end(), r, r + nr);

    // every column's value at u: dot with eq(u, .) (MultilinearPoly::evaluate)
    Fe *u_dev = ctx.d<Fe>("mercury.bu", std::max(width, 1));
    Fe *eq_u = ctx.d<Fe>("mercury.beq", column);
    const unsigned G = (unsigned)std::min<size_t>(1024, (column + MTPB - 1) / MTPB);
    Fe *partials = ctx.d<Fe>("mercury.bpart", (size_t)G * ncols);
    if (width) HIP_CHECK(hipMemcpyAsync(u_dev, u, width * sizeof(Fe), hipMemcpyHostToDevice, ctx.stream));
    hipLaunchKernelGGL(k_eq_point, dim3(grid(column)), dim3(MTPB), 0, ctx.stream, u_dev, width, column, eq_u);
    for (uint32_t c = 0; c < ncols; c++)
        hipLaunchKernelGGL(k_dot_partial, dim3(G), dim3(MTPB), 0, ctx.stream, cols_dev + (size_t)c * column, eq_u,
                           column, partials + (size_t)c * G);
    HIP_CHECK(hipGetLastError());
    Poly part;
    download(ctx, part, partials, (size_t)G * ncols);
    std::vector<Fe> colval(ncols, zero());
    for (uint32_t c = 0; c < ncols; c++)
        for (unsigned q = 0; q < G; q++) colval[c] = add(colval[c], part[(size_t)c * G + q]);
    const Poly eq_r = eq_table(r, (size_t)nr);
    std::vector<Fe> vs(nstacks, zero());
    for (size_t s = 0; s < nstacks; s++)
        for (uint32_t j = 0; j < stack_sizes[s]; j++) vs[s] = add(vs[s], mul(eq_r[j], colval[firsts[s] + j]));

    // the preamble (crates/pcs-verify/src/lib.rs:333-345), then cm* = sum rho^i cm_i
    tr.append_g1_wire(TAG_COMMITMENT, cms_wire, nstacks);
    std::vector<Fe> claim = point;
    claim.insert(claim.end(), vs.begin(), vs.end());
    tr.append_scalars(TAG_EVALUATION_CLAIM, claim.data(), claim.size());
    const Fe rho = tr.challenge(TAG_MERCURY_BATCH);
    std::vector<Fe> w(nstacks);
    Fe v_star = zero(), pw = one();
    for (size_t s = 0; s < nstacks; s++) {
        w[s] = pw;
        v_star = add(v_star, mul(pw, vs[s]));
        pw = mul(pw, rho);
    }
    Xyzz cm = xyzz_identity();
    for (size_t s = 0; s < nstacks; s++) {
        Affine a;
        const uint8_t *wire = cms_wire + 64 * s;
        bool inf = true;
        for (int q = 0; q < 64; q++) inf = inf && wire[q] == 0;
        if (inf) continue;
        if (!fe_from_bytes_fq(wire, a.x) || !fe_from_bytes_fq(wire + 32, a.y))
            throw ApgError(APG_ERR_ARG, "batch open: a commitment is not a canonical point");
        const Fe y2 = f_sqr<Fq>(a.y), x3 = f_add<Fq>(f_mul<Fq>(f_sqr<Fq>(a.x), a.x), f_from_u64<Fq>(3));
        if (!f_eq(y2, x3)) throw ApgError(APG_ERR_ARG, "batch open: a commitment is off the curve");
        const Fe k = f_from_mont<Fr>(w[s]);
        Xyzz acc = xyzz_identity();
        for (int bit = 255; bit >= 0; bit--) {
            acc = xyzz_dbl(acc);
            if ((k.l[bit >> 5] >> (bit & 31)) & 1) acc = xyzz_add_affine(acc, a);
        }
        cm = xyzz_add(cm, acc);
    }
    uint8_t cm_star[64];
    const Affine cs = xyzz_to_affine(cm);
    if (aff_is_inf(cs)) {
        std::memset(cm_star, 0, 64);
    } else {
        fe_to_bytes_fq(cs.x, cm_star);
        fe_to_bytes_fq(cs.y, cm_star + 32);
    }

    // the combined column, then one opening at u || r
    uint32_t *meta = ctx.d<uint32_t>("mercury.bmeta", 2 * nstacks);
    Fe *w_dev = ctx.d<Fe>("mercury.bw", nstacks);
    HIP_CHECK(hipMemcpyAsync(meta, firsts.data(), nstacks * 4, hipMemcpyHostToDevice, ctx.stream));
    HIP_CHECK(hipMemcpyAsync(meta + nstacks, stack_sizes, nstacks * 4, hipMemcpyHostToDevice, ctx.stream));
    HIP_CHECK(hipMemcpyAsync(w_dev, w.data(), nstacks * sizeof(Fe), hipMemcpyHostToDevice, ctx.stream));
    Fe *comb = ctx.d<Fe>("mercury.bcomb", n);
    hipLaunchKernelGGL(k_stack_combine, dim3(grid(n)), dim3(MTPB), 0, ctx.stream, cols_dev, column, n, meta,
                       meta + nstacks, w_dev, (uint32_t)nstacks, comb);
    HIP_CHECK(hipGetLastError());
    Fe v;
    mercury_open(ctx, commit, comb, nv, cm_star, point.data(), tr, v, proof);
    if (!f_eq(v, v_star)) throw ApgError(APG_ERR_INTERNAL, "batch open: the combined value is off");
    for (size_t s = 0; s < nstacks; s++) vs_out[s] = vs[s];
}

}  // namespace apg
