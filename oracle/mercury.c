/*
 * mercury.c — CPU restatement of the Mercury opening, pcs::open
 * (crates/pcs/src/lib.rs:313-492), with the univariate helpers of
 * crates/pcs-verify/src/uni.rs, the BDFG20 batch of
 * crates/pcs-verify/src/bdfg.rs and the challenge rules of
 * crates/pcs-verify/src/lib.rs:156-240.
 *
 * TEST INFRASTRUCTURE ONLY, like oracle.c: the checker for the GPU opening,
 * pinned by the reference's committed full proof at n = 2^4
 * (crates/pcs/tests/vectors/mercury_proof.txt: proof bytes and the terminal
 * transcript sample).
 *
 * The symmetric witness is computed here as the plain polynomial product the
 * reference's size-2b FFT evaluates (crates/pcs/src/lib.rs:812-866): the
 * product of two polynomials of degree < b is unique, so any exact method gives
 * the same coefficients.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef uint64_t u64;
typedef struct {
    u64 v[4];
} fe;

/* crates/constants/src/lib.rs:579-660 */
enum {
    TAG_COMMITMENT = 3,
    TAG_EVALUATION_CLAIM = 6,
    TAG_PCS_OPENING = 7,
    TAG_MERCURY_INSTANCE = 10,
    TAG_MERCURY_ALPHA = 11,
    TAG_MERCURY_GAMMA = 12,
    TAG_MERCURY_Z = 13,
    TAG_BDFG_BATCH = 14,
    TAG_BDFG_POINT = 15,
    TAG_PAIRING_MERGE = 16,
    TAG_MERCURY_BATCH = 17,
};

static fe F0(void) {
    fe r = {{0, 0, 0, 0}};
    return r;
}
static fe Fu(u64 x) {
    fe r;
    orc_from_u64(0, x, r.v);
    return r;
}
static fe fadd(fe a, fe b) {
    fe r;
    orc_add(0, a.v, b.v, r.v);
    return r;
}
static fe fsub(fe a, fe b) {
    fe r;
    orc_sub(0, a.v, b.v, r.v);
    return r;
}
static fe fmul(fe a, fe b) {
    fe r;
    orc_mul(0, a.v, b.v, r.v);
    return r;
}
static fe fneg(fe a) {
    fe r;
    orc_neg(0, a.v, r.v);
    return r;
}
static int fzero(fe a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }
static int feq(fe a, fe b) { return memcmp(a.v, b.v, 32) == 0; }
static fe finv(fe a) {
    fe r;
    orc_inverse(0, a.v, r.v);
    return r;
}

/* A coefficient vector; len is the coefficient count (uni.rs:1-10). */
typedef struct {
    fe *c;
    size_t len;
} poly;

static poly pnew(size_t len) {
    poly p;
    p.len = len;
    p.c = (fe *)calloc(len ? len : 1, sizeof(fe));
    return p;
}
static void pfree(poly *p) {
    free(p->c);
    p->c = NULL;
    p->len = 0;
}

/* uni.rs:18-24 eval, Horner */
static fe peval(const fe *c, size_t len, fe x) {
    fe acc = F0();
    for (size_t i = len; i-- > 0;) acc = fadd(fmul(acc, x), c[i]);
    return acc;
}

/* uni.rs:32-44 div_by_linear: quotient of len - 1 coefficients, remainder */
static poly div_by_linear(const fe *c, size_t len, fe r, fe *rem) {
    if (len == 0) {
        *rem = F0();
        return pnew(0);
    }
    poly q = pnew(len - 1);
    fe acc = c[len - 1];
    for (size_t i = len - 1; i-- > 0;) {
        q.c[i] = acc;
        acc = fadd(c[i], fmul(acc, r));
    }
    *rem = acc;
    return q;
}

/* uni.rs:47-58 mul, schoolbook */
static poly pmul(const fe *a, size_t la, const fe *b, size_t lb) {
    if (!la || !lb) return pnew(0);
    poly o = pnew(la + lb - 1);
    for (size_t i = 0; i < la; i++)
        for (size_t j = 0; j < lb; j++) o.c[i + j] = fadd(o.c[i + j], fmul(a[i], b[j]));
    return o;
}

/* uni.rs:61-68 add_scaled: acc += c * src, growing acc */
static void add_scaled(poly *acc, const fe *src, size_t ls, fe c) {
    if (ls > acc->len) {
        acc->c = (fe *)realloc(acc->c, ls * sizeof(fe));
        for (size_t i = acc->len; i < ls; i++) acc->c[i] = F0();
        acc->len = ls;
    }
    for (size_t i = 0; i < ls; i++) acc->c[i] = fadd(acc->c[i], fmul(c, src[i]));
}

/* uni.rs:71-77 vanishing: prod (X - r_i) */
static poly vanishing(const fe *roots, size_t k) {
    poly out = pnew(1);
    out.c[0] = Fu(1);
    for (size_t i = 0; i < k; i++) {
        fe lin[2] = {fneg(roots[i]), Fu(1)};
        poly t = pmul(out.c, out.len, lin, 2);
        pfree(&out);
        out = t;
    }
    return out;
}

/* uni.rs:84-102 interpolate, Lagrange */
static poly interpolate(const fe *xs, const fe *ys, size_t k) {
    poly out = pnew(k);
    for (size_t i = 0; i < k; i++) {
        poly num = pnew(1);
        num.c[0] = Fu(1);
        fe den = Fu(1);
        for (size_t j = 0; j < k; j++) {
            if (i == j) continue;
            fe lin[2] = {fneg(xs[j]), Fu(1)};
            poly t = pmul(num.c, num.len, lin, 2);
            pfree(&num);
            num = t;
            den = fmul(den, fsub(xs[i], xs[j]));
        }
        add_scaled(&out, num.c, num.len, fmul(ys[i], finv(den)));
        pfree(&num);
    }
    return out;
}

/* uni.rs:105-118 pow_usize */
static fe pow_usize(fe x, size_t e) {
    fe acc = Fu(1), base = x;
    while (e) {
        if (e & 1) acc = fmul(acc, base);
        base = fmul(base, base);
        e >>= 1;
    }
    return acc;
}

/* pcs-verify/src/lib.rs:125-131 dot */
static fe dot(const fe *a, const fe *b, size_t n) {
    fe acc = F0();
    for (size_t i = 0; i < n; i++) acc = fadd(acc, fmul(a[i], b[i]));
    return acc;
}

/* pcs-verify/src/lib.rs:186-196 tensor_eval: prod_k (u_k x^(2^k) + 1 - u_k) */
static fe tensor_eval(const fe *u, size_t t, fe x) {
    fe acc = Fu(1), power = x, one = Fu(1);
    for (size_t k = 0; k < t; k++) {
        acc = fmul(acc, fadd(fmul(u[k], power), fsub(one, u[k])));
        power = fmul(power, power);
    }
    return acc;
}

/* pcs-verify/src/lib.rs:205-227 derive_h_alpha */
static fe derive_h_alpha(const fe *u1, const fe *u2, size_t t, fe z, fe z_inv, fe gamma, fe v,
                         const fe ev[6]) {
    fe two_inv = finv(Fu(2));
    fe inner = fadd(fmul(ev[0], tensor_eval(u1, t, z_inv)), fmul(ev[1], tensor_eval(u1, t, z)));
    fe hpart = fadd(fmul(ev[2], tensor_eval(u2, t, z_inv)), fmul(ev[3], tensor_eval(u2, t, z)));
    hpart = fsub(fsub(hpart, v), v);
    inner = fadd(inner, fmul(gamma, hpart));
    inner = fsub(fsub(inner, fmul(z, ev[4])), fmul(z_inv, ev[5]));
    return fmul(inner, two_inv);
}

/* pcs/src/lib.rs:812-866 symmetric_witness: R = g rev(eq1) + gamma h rev(eq2),
 * S[k - b] = R[k] + R[2b - 2 - k] for k in [b, 2b - 1) */
static poly symmetric_witness(const fe *g, const fe *h, const fe *eq1, const fe *eq2, size_t b,
                              fe gamma) {
    fe *r1 = (fe *)malloc(b * sizeof(fe)), *r2 = (fe *)malloc(b * sizeof(fe));
    for (size_t i = 0; i < b; i++) {
        r1[i] = eq1[b - 1 - i];
        r2[i] = eq2[b - 1 - i];
    }
    poly a = pmul(g, b, r1, b);
    poly c = pmul(h, b, r2, b);
    /* R has 2b - 1 coefficients (the reference's R[2b - 1] is zero) */
    for (size_t k = 0; k < a.len; k++) a.c[k] = fadd(a.c[k], fmul(gamma, c.c[k]));
    poly s = pnew(b - 1);
    for (size_t k = b; k < 2 * b - 1; k++) s.c[k - b] = fadd(a.c[k], a.c[2 * b - 2 - k]);
    pfree(&a);
    pfree(&c);
    free(r1);
    free(r2);
    return s;
}

/* commit_to (pcs/src/lib.rs:342-346): msm(powers[..len], c).to_affine() in wire form */
static void commit_to(const u64 *srs, const fe *c, size_t len, size_t workers, uint8_t out[64]) {
    u64 jac[12], aff[8];
    orc_msm(srs, (const u64 *)c, len, workers, workers, jac);
    orc_g1_to_affine(jac, aff);
    orc_g1_affine_to_wire(aff, out);
}

/* pcs-verify/src/lib.rs:156-169 challenge_z: squeeze under MERCURY_Z until nonzero */
static fe challenge_z(orc_transcript *tr) {
    for (;;) {
        fe z;
        orc_tr_challenge(tr, TAG_MERCURY_Z, z.v);
        if (!fzero(z)) return z;
    }
}

/* pcs-verify/src/lib.rs:176-178 degenerate */
static int degenerate(fe alpha, fe z) {
    fe one = Fu(1);
    return fzero(z) || feq(fmul(z, z), one) || feq(z, alpha) || feq(fmul(z, alpha), one);
}

typedef struct {
    poly zc, r;
} item;

/* bdfg.rs:63-86 items over T = {z, 1/z, alpha}, in the order g, h, S, D */
static void bdfg_items(fe alpha, fe z, fe z_inv, const fe cl[8], item it[4]) {
    /* cl: g_z, g_inv_z, h_z, h_inv_z, s_z, s_inv_z, h_alpha, d_z */
    fe xs2[2] = {z, z_inv}, xs3[3] = {z, z_inv, alpha};
    fe ys_g[2] = {cl[0], cl[1]}, ys_h[3] = {cl[2], cl[3], cl[6]}, ys_s[2] = {cl[4], cl[5]};
    it[0].zc = vanishing(&alpha, 1);
    it[0].r = interpolate(xs2, ys_g, 2);
    it[1].zc = pnew(1);
    it[1].zc.c[0] = Fu(1);
    it[1].r = interpolate(xs3, ys_h, 3);
    it[2].zc = vanishing(&alpha, 1);
    it[2].r = interpolate(xs2, ys_s, 2);
    fe za[2] = {z_inv, alpha};
    it[3].zc = vanishing(za, 2);
    it[3].r = pnew(1);
    it[3].r.c[0] = cl[7];
}

/* bdfg.rs:94-118 quotient: F / Z_T, F = sum_i delta^i Z_{T\S_i} (f_i - r_i) */
static int bdfg_quotient(const poly polys[4], const item it[4], const fe t_set[3], fe delta,
                         poly *out) {
    poly f = pnew(0);
    fe power = Fu(1);
    for (int i = 0; i < 4; i++) {
        poly shifted = pnew(polys[i].len);
        memcpy(shifted.c, polys[i].c, polys[i].len * sizeof(fe));
        add_scaled(&shifted, it[i].r.c, it[i].r.len, fneg(Fu(1)));
        poly m = pmul(it[i].zc.c, it[i].zc.len, shifted.c, shifted.len);
        add_scaled(&f, m.c, m.len, power);
        pfree(&m);
        pfree(&shifted);
        power = fmul(power, delta);
    }
    for (int k = 0; k < 3; k++) {
        fe rem;
        poly q = div_by_linear(f.c, f.len, t_set[k], &rem);
        pfree(&f);
        f = q;
        if (!fzero(rem)) {
            pfree(&f);
            return -1;
        }
    }
    *out = f;
    return 0;
}

/* bdfg.rs:122-142 linearization */
static poly bdfg_linearization(const poly polys[4], const item it[4], const fe t_set[3],
                               const poly *w, fe delta, fe zp) {
    poly l = pnew(0);
    fe constant = F0();
    poly zt = vanishing(t_set, 3);
    add_scaled(&l, w->c, w->len, fneg(peval(zt.c, zt.len, zp)));
    pfree(&zt);
    for (int i = 0; i < 4; i++) {
        fe c = fmul(pow_usize(delta, (size_t)i), peval(it[i].zc.c, it[i].zc.len, zp));
        add_scaled(&l, polys[i].c, polys[i].len, c);
        constant = fadd(constant, fmul(c, peval(it[i].r.c, it[i].r.len, zp)));
    }
    if (l.len == 0) {
        pfree(&l);
        l = pnew(1);
    }
    l.c[0] = fsub(l.c[0], constant);
    return l;
}

static void absorb_points(orc_transcript *tr, u64 tag, const uint8_t *pts, size_t k) {
    orc_tr_append_g1(tr, tag, pts, k);
}

int orc_mercury_open(const uint64_t *srs_aff, size_t srs_len, const uint64_t *table, size_t num_vars,
                     const uint8_t *cm_wire, const uint64_t *u_in, orc_transcript *tr, size_t workers,
                     uint64_t *v_out, uint8_t *proof_out) {
    /* pcs-verify/src/lib.rs:97-103 check_num_vars */
    if (num_vars < 2 || num_vars > 54 || (num_vars & 1)) return -2;
    const size_t n = (size_t)1 << num_vars;
    if (srs_len < n) return -3;
    const size_t t = num_vars / 2, b = (size_t)1 << t;
    const fe *coeffs = (const fe *)table;
    const fe *u = (const fe *)u_in;
    int rc = 0;

    fe *eq1 = (fe *)malloc(b * sizeof(fe)), *eq2 = (fe *)malloc(b * sizeof(fe));
    orc_eq_table(u[0].v, t, (u64 *)eq1);
    orc_eq_table(u[t].v, t, (u64 *)eq2);

    /* step 1: h_j = dot(eq1, row j), v = dot(eq2, h) */
    fe *h = (fe *)malloc(b * sizeof(fe));
    for (size_t j = 0; j < b; j++) h[j] = dot(eq1, coeffs + j * b, b);
    const fe v = dot(eq2, h, b);
    uint8_t pts[8][64];
    commit_to(srs_aff, h, b, workers, pts[0]);

    fe nf = Fu((u64)n);
    orc_tr_append_scalars(tr, TAG_MERCURY_INSTANCE, nf.v, 1);
    absorb_points(tr, TAG_COMMITMENT, cm_wire, 1);
    fe *claim = (fe *)malloc((num_vars + 1) * sizeof(fe));
    memcpy(claim, u, num_vars * sizeof(fe));
    claim[num_vars] = v;
    orc_tr_append_scalars(tr, TAG_EVALUATION_CLAIM, (const u64 *)claim, num_vars + 1);
    free(claim);
    absorb_points(tr, TAG_PCS_OPENING, pts[0], 1);
    fe alpha;
    orc_tr_challenge(tr, TAG_MERCURY_ALPHA, alpha.v);

    /* step 2: the fold, b Horner divisions by X - alpha interleaved over rows */
    fe *acc = (fe *)malloc(b * sizeof(fe));
    memcpy(acc, coeffs + (b - 1) * b, b * sizeof(fe));
    fe *q = (fe *)calloc(b > 1 ? (b - 1) * b : 1, sizeof(fe));
    for (size_t j = b - 1; j-- > 0;) {
        memcpy(q + j * b, acc, b * sizeof(fe));
        for (size_t i = 0; i < b; i++) acc[i] = fadd(coeffs[j * b + i], fmul(acc[i], alpha));
    }
    fe *g = acc;
    commit_to(srs_aff, q, (b - 1) * b, workers, pts[1]);
    commit_to(srs_aff, g, b, workers, pts[2]);
    absorb_points(tr, TAG_PCS_OPENING, pts[1], 2);
    fe gamma;
    orc_tr_challenge(tr, TAG_MERCURY_GAMMA, gamma.v);

    /* step 3: S and D */
    const fe h_alpha = peval(h, b, alpha);
    poly s = symmetric_witness(g, h, eq1, eq2, b, gamma);
    fe *d = (fe *)malloc(b * sizeof(fe));
    for (size_t i = 0; i < b; i++) d[i] = g[b - 1 - i];
    commit_to(srs_aff, s.c, s.len, workers, pts[3]);
    commit_to(srs_aff, d, b, workers, pts[4]);
    absorb_points(tr, TAG_PCS_OPENING, pts[3], 2);
    const fe z = challenge_z(tr);
    if (degenerate(alpha, z)) {
        rc = -4;
        goto out;
    }
    const fe z_inv = finv(z);

    /* step 4: the six values, then the fold's opening at z */
    fe ev[6];
    ev[0] = peval(g, b, z);
    ev[1] = peval(g, b, z_inv);
    ev[2] = peval(h, b, z);
    ev[3] = peval(h, b, z_inv);
    ev[4] = peval(s.c, s.len, z);
    ev[5] = peval(s.c, s.len, z_inv);
    orc_tr_append_scalars(tr, TAG_PCS_OPENING, (const u64 *)ev, 6);

    {
        const fe zb_a = fsub(pow_usize(z, b), alpha);
        fe *num = (fe *)malloc(n * sizeof(fe));
        memcpy(num, coeffs, n * sizeof(fe));
        for (size_t i = 0; i < (b - 1) * b; i++) num[i] = fsub(num[i], fmul(zb_a, q[i]));
        num[0] = fsub(num[0], ev[0]);
        fe rem;
        poly hq = div_by_linear(num, n, z, &rem);
        free(num);
        if (!fzero(rem)) {
            pfree(&hq);
            rc = -5;
            goto out;
        }
        commit_to(srs_aff, hq.c, hq.len, workers, pts[5]);
        pfree(&hq);
    }
    absorb_points(tr, TAG_PCS_OPENING, pts[5], 1);

    /* step 4(e): BDFG20 */
    {
        fe cl[8];
        memcpy(cl, ev, sizeof(ev));
        cl[6] = derive_h_alpha(u, u + t, t, z, z_inv, gamma, v, ev);
        cl[7] = fmul(pow_usize(z, b - 1), ev[1]);
        if (!feq(cl[6], h_alpha)) {
            rc = -6;
            goto out;
        }
        const fe t_set[3] = {z, z_inv, alpha};
        item it[4];
        bdfg_items(alpha, z, z_inv, cl, it);
        poly polys[4] = {{g, b}, {h, b}, s, {d, b}};
        fe delta;
        orc_tr_challenge(tr, TAG_BDFG_BATCH, delta.v);
        poly w;
        if (bdfg_quotient(polys, it, t_set, delta, &w) != 0) {
            rc = -7;
        } else {
            commit_to(srs_aff, w.c, w.len, workers, pts[6]);
            absorb_points(tr, TAG_PCS_OPENING, pts[6], 1);
            fe zp;
            orc_tr_challenge(tr, TAG_BDFG_POINT, zp.v);
            poly l = bdfg_linearization(polys, it, t_set, &w, delta, zp);
            fe rem;
            poly wp = div_by_linear(l.c, l.len, zp, &rem);
            if (!fzero(rem)) {
                rc = -8;
            } else {
                commit_to(srs_aff, wp.c, wp.len, workers, pts[7]);
                absorb_points(tr, TAG_PCS_OPENING, pts[7], 1);
                fe merge;
                orc_tr_challenge(tr, TAG_PAIRING_MERGE, merge.v);
            }
            pfree(&wp);
            pfree(&l);
            pfree(&w);
        }
        for (int i = 0; i < 4; i++) {
            pfree(&it[i].zc);
            pfree(&it[i].r);
        }
        if (rc) goto out;
    }

    /* MercuryProof::to_bytes (pcs/src/lib.rs:152-161) */
    memcpy(proof_out, pts, 8 * 64);
    for (int i = 0; i < 6; i++) orc_to_bytes(0, ev[i].v, proof_out + 8 * 64 + 32 * i);
    memcpy(v_out, v.v, 32);

out:
    pfree(&s);
    free(d);
    free(q);
    free(acc);
    free(h);
    free(eq1);
    free(eq2);
    return rc;
}

/* pcs/src/lib.rs:701-789 batch_open_stacked (batch_open is r = [], one column
 * per stack, :689-698): each column's value at u, the stacks' values
 * sum_j eq(r, j) col_j(u), the batch preamble (pcs-verify/src/lib.rs:333-345:
 * the commitments as one message, u || r then the values, rho), cm* = sum
 * rho^i cm_i by an MSM, the combined column, then one opening at u || r.
 * cols: the stacks' columns back to back in slot order, 2^width each.
 * Returns 0, -9 EmptyBatch, -10 BatchLengthMismatch, -11 bad commitment,
 * -2 / -3 as open, or open's own errors. */
int orc_mercury_batch_open_stacked(const uint64_t *srs_aff, size_t srs_len, const uint64_t *cols,
                                   const uint32_t *stack_sizes, size_t nstacks, size_t width,
                                   const uint8_t *cms_wire, const uint64_t *u_in, const uint64_t *r_in,
                                   size_t nr, orc_transcript *tr, size_t workers, uint64_t *vs_out,
                                   uint8_t *proof_out) {
    if (nstacks == 0) return -9;
    for (size_t s = 0; s < nstacks; s++)
        if (stack_sizes[s] == 0) return -9;
    const size_t nv = width + nr;
    if (nv < 2 || nv > 54 || (nv & 1)) return -2;
    const size_t slots = (size_t)1 << nr;
    for (size_t s = 0; s < nstacks; s++)
        if (stack_sizes[s] > slots) return -10;
    const size_t n = (size_t)1 << nv, column = (size_t)1 << width;
    if (srs_len < n) return -3;
    const fe *c = (const fe *)cols;

    fe *point = (fe *)malloc((nv ? nv : 1) * sizeof(fe));
    memcpy(point, u_in, width * sizeof(fe));
    memcpy(point + width, r_in, nr * sizeof(fe));
    fe *eq_r = (fe *)malloc(slots * sizeof(fe));
    orc_eq_table(r_in, nr, (u64 *)eq_r);
    fe *vs = (fe *)malloc(nstacks * sizeof(fe));
    size_t first = 0;
    for (size_t s = 0; s < nstacks; s++) {
        fe acc = F0();
        for (size_t j = 0; j < stack_sizes[s]; j++) {
            fe val;
            orc_evaluate((const u64 *)(c + (first + j) * column), width, u_in, val.v);
            acc = fadd(acc, fmul(eq_r[j], val));
        }
        vs[s] = acc;
        first += stack_sizes[s];
    }

    /* the preamble */
    absorb_points(tr, TAG_COMMITMENT, cms_wire, nstacks);
    fe *claim = (fe *)malloc((nv + nstacks) * sizeof(fe));
    memcpy(claim, point, nv * sizeof(fe));
    memcpy(claim + nv, vs, nstacks * sizeof(fe));
    orc_tr_append_scalars(tr, TAG_EVALUATION_CLAIM, (const u64 *)claim, nv + nstacks);
    free(claim);
    fe rho;
    orc_tr_challenge(tr, TAG_MERCURY_BATCH, rho.v);
    fe *w = (fe *)malloc(nstacks * sizeof(fe));
    fe v_star = F0(), pw = Fu(1);
    for (size_t s = 0; s < nstacks; s++) {
        w[s] = pw;
        v_star = fadd(v_star, fmul(pw, vs[s]));
        pw = fmul(pw, rho);
    }
    int rc = 0;
    uint8_t cm_star[64];
    {
        u64 *bases = (u64 *)malloc(nstacks * 64);
        for (size_t s = 0; s < nstacks && !rc; s++)
            if (!orc_g1_affine_from_wire(cms_wire + 64 * s, bases + 8 * s)) rc = -11;
        if (!rc) {
            u64 jac[12], aff[8];
            orc_msm(bases, (const u64 *)w, nstacks, workers, workers, jac);
            orc_g1_to_affine(jac, aff);
            orc_g1_affine_to_wire(aff, cm_star);
        }
        free(bases);
    }
    fe *comb = NULL;
    if (!rc) {
        comb = (fe *)calloc(n, sizeof(fe));
        for (size_t i = 0; i < n; i++) {
            const size_t slot = i / column, row = i % column;
            fe acc = F0();
            size_t f0 = 0;
            for (size_t s = 0; s < nstacks; s++) {
                if (slot < stack_sizes[s]) acc = fadd(acc, fmul(w[s], c[(f0 + slot) * column + row]));
                f0 += stack_sizes[s];
            }
            comb[i] = acc;
        }
        fe v;
        rc = orc_mercury_open(srs_aff, srs_len, (const u64 *)comb, nv, cm_star, (const u64 *)point, tr, workers,
                              v.v, proof_out);
        if (!rc && !feq(v, v_star)) rc = -12;
        if (!rc) memcpy(vs_out, vs, nstacks * sizeof(fe));
    }
    free(comb);
    free(w);
    free(vs);
    free(eq_r);
    free(point);
    return rc;
}
