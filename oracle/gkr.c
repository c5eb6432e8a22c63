/*
 * gkr.c — CPU restatement of the general GKR layer: the gate kernel
 * (gkr_verify::eval_gate, crates/gkr-verify/src/lib.rs:142-202), the virtual
 * tables (virtual_at_row / virtual_at_point, :204-288), a resolved gate list's
 * summand and gate values (ResolvedList, :300-490), the forward pass of one
 * list (gkr::produce with RowReader, crates/gkr/src/lib.rs:193-336) and the
 * layer sumcheck prover (gkr::prove_sumcheck, crates/gkr/src/lib.rs:792-899).
 *
 * TEST INFRASTRUCTURE ONLY, like oracle.c: the checker for the GPU layer
 * kernels. Pinned by the gate kernel's known answers (crates/gkr/tests/
 * kernel.rs, restated as fixtures), by the tables' defining property (the
 * closed form at every cube point equals the row value), and by a brute-force
 * definition of the round polynomials on small layers (tests/).
 *
 * A gate list arrives flattened, as ResolvedList::new resolves it: gates in
 * the order cached entries, producing, enforcing; gate i reads
 * sources[src_begin, src_end) in eval_gate's value order; its coefficients are
 * GateDef::coefficients (crates/constraints/src/lib.rs:393-430) with every
 * challenge slot already replaced by its value.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef uint64_t u64;
typedef struct {
    u64 v[4];
} fe;

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
static fe ld(const u64 *p) {
    fe r;
    memcpy(r.v, p, 32);
    return r;
}

/* crates/constants/src/lib.rs memory::RAM_LIVE_BIT */
#define RAM_LIVE_BIT 14

/* --- eval_gate (crates/gkr-verify/src/lib.rs:142-202) --- */

/* c_0 + sum_i k_i v_i, folded from the constant as the reference folds */
static fe affine(const u64 *ks, const u64 *c0, const fe *vals, uint32_t t) {
    fe acc = ld(c0);
    for (uint32_t i = 0; i < t; i++) acc = fadd(acc, fmul(ld(ks + 4 * i), vals[i]));
    return acc;
}

static fe eval_gate(uint32_t shape, const u64 *co, uint32_t split, const fe *v, uint32_t nv) {
    switch (shape) {
        case ORC_GATE_LINEAR: /* terms' coefficients, then the constant */
            return affine(co, co + 4 * nv, v, nv);
        case ORC_GATE_PRODUCT:
            return fmul(fmul(ld(co), v[0]), v[1]);
        case ORC_GATE_MASK: /* x m + 1 - m */
            return fsub(fadd(fmul(v[0], v[1]), Fu(1)), v[1]);
        case ORC_GATE_AFFINE_PRODUCT: { /* a_1..a_t, a_0, b_1..b_u, b_0 */
            const uint32_t t = split, u = nv - split;
            const fe l = affine(co, co + 4 * t, v, t);
            const fe r = affine(co + 4 * (t + 1), co + 4 * (t + 1 + u), v + t, u);
            return fmul(l, r);
        }
        case ORC_GATE_TREE_PRODUCT:
            return fmul(v[0], v[1]);
        case ORC_GATE_QUADRATIC: { /* c_0, a_1..a_t, b_1..b_u */
            const uint32_t t = split, u = (nv - split) / 2;
            fe acc = affine(co + 4, co, v, t);
            for (uint32_t j = 0; j < u; j++)
                acc = fadd(acc, fmul(fmul(ld(co + 4 * (1 + t + j)), v[t + 2 * j]), v[t + 2 * j + 1]));
            return acc;
        }
        case ORC_GATE_TREE_CROSS:
            return fadd(fmul(v[0], v[3]), fmul(v[1], v[2]));
        default:
            abort();
    }
}

void orc_gkr_eval_gate(uint32_t shape, const uint64_t *coeffs, uint32_t split, const uint64_t *values,
                       uint32_t nvalues, uint64_t *out) {
    fe *v = malloc(sizeof(fe) * (nvalues ? nvalues : 1));
    for (uint32_t i = 0; i < nvalues; i++) v[i] = ld(values + 4 * i);
    fe r = eval_gate(shape, coeffs, split, v, nvalues);
    memcpy(out, r.v, 32);
    free(v);
}

/* --- virtual tables (crates/gkr-verify/src/lib.rs:204-288) --- */

static fe virtual_at_row(uint32_t kind, u64 row) {
    switch (kind) {
        case ORC_VIRT_ROW_INDEX:
            return Fu(row);
        case ORC_VIRT_RAM_LIVE:
            return Fu(row >= (1ull << RAM_LIVE_BIT) ? 1 : 0);
        case ORC_VIRT_RANGE19:
            return Fu(row & ((1ull << 19) - 1));
        case ORC_VIRT_RANGE16:
            return Fu(row & ((1ull << 16) - 1));
        case ORC_VIRT_XOR8_A:
            return Fu(row & 0xff);
        case ORC_VIRT_XOR8_B:
            return Fu((row >> 8) & 0xff);
        case ORC_VIRT_XOR8_OUT:
            return Fu((row & 0xff) ^ ((row >> 8) & 0xff));
        default:
            abort();
    }
}

static fe coord(const fe *point, size_t n, size_t j) { return j < n ? point[j] : F0(); }

/* sum_{j < len} 2^j y_j by Horner from the highest variable */
static fe horner2(const fe *y, size_t len) {
    fe acc = F0();
    for (size_t j = len; j-- > 0;) acc = fadd(fadd(acc, acc), y[j]);
    return acc;
}

static fe virtual_at_point(uint32_t kind, const fe *point, size_t n) {
    switch (kind) {
        case ORC_VIRT_ROW_INDEX:
            return horner2(point, n);
        case ORC_VIRT_RAM_LIVE: {
            const size_t lo = n < RAM_LIVE_BIT ? n : RAM_LIVE_BIT;
            fe p = Fu(1);
            for (size_t j = lo; j < n; j++) p = fmul(p, fsub(Fu(1), point[j]));
            return fsub(Fu(1), p);
        }
        case ORC_VIRT_RANGE19:
        case ORC_VIRT_RANGE16: {
            const size_t bits = kind == ORC_VIRT_RANGE19 ? 19 : 16;
            return horner2(point, n < bits ? n : bits);
        }
        case ORC_VIRT_XOR8_A:
        case ORC_VIRT_XOR8_B: {
            const size_t first = kind == ORC_VIRT_XOR8_A ? 0 : 8;
            fe acc = F0();
            for (size_t j = 8; j-- > 0;) acc = fadd(fadd(acc, acc), coord(point, n, first + j));
            return acc;
        }
        case ORC_VIRT_XOR8_OUT: {
            fe acc = F0();
            for (size_t j = 8; j-- > 0;) {
                const fe y = coord(point, n, j), z = coord(point, n, j + 8);
                const fe yz = fmul(y, z);
                acc = fsub(fadd(fadd(fadd(acc, acc), y), z), fadd(yz, yz));
            }
            return acc;
        }
        default:
            abort();
    }
}

void orc_gkr_virtual_at_row(uint32_t kind, uint64_t row, uint64_t *out) {
    fe r = virtual_at_row(kind, row);
    memcpy(out, r.v, 32);
}

void orc_gkr_virtual_at_point(uint32_t kind, const uint64_t *point, size_t n, uint64_t *out) {
    fe *p = malloc(sizeof(fe) * (n ? n : 1));
    for (size_t j = 0; j < n; j++) p[j] = ld(point + 4 * j);
    fe r = virtual_at_point(kind, p, n);
    memcpy(out, r.v, 32);
    free(p);
}

/* --- the resolved list (crates/gkr-verify/src/lib.rs:300-490) --- */

typedef struct {
    const orc_gkr_list *l;
    fe *cached; /* ncached */
    fe *ops;    /* the widest gate's operands */
} resolved;

static void resolved_init(resolved *r, const orc_gkr_list *l) {
    uint32_t widest = 1;
    const uint32_t all = l->ncached + l->nproducing + l->nenforcing;
    for (uint32_t i = 0; i < all; i++) {
        const uint32_t w = l->gates[i].src_end - l->gates[i].src_begin;
        if (w > widest) widest = w;
    }
    r->l = l;
    r->cached = malloc(sizeof(fe) * (l->ncached ? l->ncached : 1));
    r->ops = malloc(sizeof(fe) * widest);
}

static void resolved_free(resolved *r) {
    free(r->cached);
    free(r->ops);
}

/* gate i of the whole list, its operands gathered from their sources */
static fe resolved_eval(resolved *r, uint32_t i, const fe *lower, const fe *upper, const fe *virtuals) {
    const orc_gkr_gate *g = &r->l->gates[i];
    const uint32_t nv = g->src_end - g->src_begin;
    for (uint32_t q = 0; q < nv; q++) {
        const uint32_t s = r->l->sources[g->src_begin + q];
        const uint32_t idx = s & ORC_SRC_INDEX_MASK;
        switch (s >> ORC_SRC_KIND_SHIFT) {
            case ORC_SRC_LOWER: r->ops[q] = lower[idx]; break;
            case ORC_SRC_UPPER: r->ops[q] = upper[idx]; break;
            case ORC_SRC_VIRTUAL: r->ops[q] = virtuals[idx]; break;
            case ORC_SRC_CACHED: r->ops[q] = r->cached[idx]; break;
            default: abort();
        }
    }
    return eval_gate(g->shape, r->l->coeffs + 4 * (size_t)g->coef_begin, g->split, r->ops, nv);
}

static void resolved_cache(resolved *r, const fe *lower, const fe *upper, const fe *virtuals) {
    for (uint32_t i = 0; i < r->l->ncached; i++) r->cached[i] = resolved_eval(r, i, lower, upper, virtuals);
}

/* ResolvedList::summand: sum_j w_j gate_j, cached entries first */
static fe resolved_summand(resolved *r, const fe *w, const fe *lower, const fe *upper, const fe *virtuals) {
    resolved_cache(r, lower, upper, virtuals);
    fe acc = F0();
    const uint32_t gates = r->l->nproducing + r->l->nenforcing;
    for (uint32_t j = 0; j < gates; j++)
        acc = fadd(acc, fmul(resolved_eval(r, r->l->ncached + j, lower, upper, virtuals), w[j]));
    return acc;
}

static fe *lda(const uint64_t *p, size_t n) {
    fe *a = malloc(sizeof(fe) * (n ? n : 1));
    for (size_t i = 0; i < n; i++) a[i] = ld(p + 4 * i);
    return a;
}

void orc_gkr_gate_values(const orc_gkr_list *l, const uint64_t *lower, size_t nlower, const uint64_t *upper,
                         size_t nupper, const uint64_t *virtuals, uint64_t *out) {
    resolved r;
    resolved_init(&r, l);
    fe *lo = lda(lower, nlower), *up = lda(upper, nupper), *vi = lda(virtuals, l->nvirtual);
    resolved_cache(&r, lo, up, vi);
    for (uint32_t j = 0; j < l->nproducing + l->nenforcing; j++) {
        fe v = resolved_eval(&r, l->ncached + j, lo, up, vi);
        memcpy(out + 4 * j, v.v, 32);
    }
    free(lo);
    free(up);
    free(vi);
    resolved_free(&r);
}

void orc_gkr_summand(const orc_gkr_list *l, const uint64_t *weights, const uint64_t *lower, size_t nlower,
                     const uint64_t *upper, size_t nupper, const uint64_t *virtuals, uint64_t *out) {
    resolved r;
    resolved_init(&r, l);
    fe *lo = lda(lower, nlower), *up = lda(upper, nupper), *vi = lda(virtuals, l->nvirtual);
    fe *w = lda(weights, l->nproducing + l->nenforcing);
    fe s = resolved_summand(&r, w, lo, up, vi);
    memcpy(out, s.v, 32);
    free(lo);
    free(up);
    free(vi);
    free(w);
    resolved_free(&r);
}

/* --- the forward pass of one list (crates/gkr/src/lib.rs:193-264,300-336) ---
 * cols: layer k's columns of 2^num_vars rows. A row-wise list defines as many
 * rows and reads the virtual tables at each; a halving list defines half as
 * many, row y reading rows y and y + half. out: one column per producing
 * gate. */
void orc_gkr_forward_layer(const orc_gkr_list *l, uint64_t *const *cols, size_t ncols, size_t num_vars,
                           uint64_t **out) {
    const size_t rows_in = (size_t)1 << num_vars;
    const size_t rows = l->halving ? rows_in / 2 : rows_in;
    resolved r;
    resolved_init(&r, l);
    fe *lo = malloc(sizeof(fe) * (ncols ? ncols : 1));
    fe *up = malloc(sizeof(fe) * (ncols ? ncols : 1));
    fe *vi = malloc(sizeof(fe) * (l->nvirtual ? l->nvirtual : 1));
    for (size_t y = 0; y < rows; y++) {
        if (l->halving) {
            for (size_t c = 0; c < ncols; c++) {
                lo[c] = ld(cols[c] + 4 * y);
                up[c] = ld(cols[c] + 4 * (y + rows));
            }
        } else {
            for (size_t c = 0; c < ncols; c++) lo[c] = ld(cols[c] + 4 * y);
            for (uint32_t v = 0; v < l->nvirtual; v++) vi[v] = virtual_at_row(l->virtual_kinds[v], y);
        }
        resolved_cache(&r, lo, up, vi);
        for (uint32_t j = 0; j < l->nproducing; j++) {
            fe v = resolved_eval(&r, l->ncached + j, lo, up, vi);
            memcpy(out[j] + 4 * y, v.v, 32);
        }
    }
    free(lo);
    free(up);
    free(vi);
    resolved_free(&r);
}

/* --- the layer sumcheck (crates/gkr/src/lib.rs:705-899) --- */

static void interpolate_cubic(const fe v[4], fe c[4]) {
    const fe inv2 = ({ fe t; orc_inverse(0, Fu(2).v, t.v); t; });
    const fe inv3 = ({ fe t; orc_inverse(0, Fu(3).v, t.v); t; });
    const fe inv6 = ({ fe t; orc_inverse(0, Fu(6).v, t.v); t; });
    const fe d1 = fsub(v[1], v[0]);
    const fe d2 = fsub(fsub(v[2], v[1]), d1);
    const fe d3 = fsub(fsub(fsub(v[3], v[2]), fsub(v[2], v[1])), d2);
    c[0] = v[0];
    c[1] = fadd(fsub(d1, fmul(d2, inv2)), fmul(d3, inv3));
    c[2] = fmul(fsub(d2, d3), inv2);
    c[3] = fmul(d3, inv6);
}

static void bind(fe *t, size_t len, fe r) {
    for (size_t i = 0; i < len / 2; i++) t[i] = fadd(t[2 * i], fmul(r, fsub(t[2 * i + 1], t[2 * i])));
}

/* lower/upper: ncols tables each (upper only for a halving list) of 2^n
 * elements, bound in place to the point; rounds_out n x 4; point_out n. The
 * virtual tables are read only by a row-wise list. */
void orc_gkr_prove_sumcheck(const orc_gkr_list *l, const uint64_t *weights, const uint64_t *eq_point, size_t n,
                            uint64_t *const *lower, uint64_t *const *upper, size_t ncols, orc_transcript *tr,
                            uint64_t *rounds_out, uint64_t *point_out) {
    const size_t len0 = (size_t)1 << n;
    const size_t nup = l->halving ? ncols : 0;
    const uint32_t nk = l->halving ? 0 : l->nvirtual;
    fe *w = lda(weights, l->nproducing + l->nenforcing);
    fe *eq = malloc(sizeof(fe) * len0);
    orc_eq_table(eq_point, n, (u64 *)eq);
    fe **T = malloc(sizeof(fe *) * (ncols + nup + 1));
    for (size_t c = 0; c < ncols; c++) T[c] = (fe *)lower[c];
    for (size_t c = 0; c < nup; c++) T[ncols + c] = (fe *)upper[c];
    const size_t nt = ncols + nup;
    fe *cur = malloc(sizeof(fe) * (nt + 1)), *step = malloc(sizeof(fe) * (nt + 1));
    fe *vcur = malloc(sizeof(fe) * (nk + 1)), *vstep = malloc(sizeof(fe) * (nk + 1));
    fe *at = malloc(sizeof(fe) * (n + 1));
    resolved r;
    resolved_init(&r, l);
    size_t len = len0;
    for (size_t j = 0; j < n; j++) {
        const size_t half = len / 2, bound = j;
        fe acc[4] = {F0(), F0(), F0(), F0()};
        for (size_t i = 0; i < half; i++) {
            for (size_t t = 0; t < nt; t++) {
                cur[t] = T[t][2 * i];
                step[t] = fsub(T[t][2 * i + 1], cur[t]);
            }
            if (nk) { /* virtual_lines: X then the bits of i after the bound challenges */
                for (size_t q = bound + 1; q < n; q++) at[q] = Fu((i >> (q - bound - 1)) & 1);
                at[bound] = F0();
                for (uint32_t v = 0; v < nk; v++) vcur[v] = virtual_at_point(l->virtual_kinds[v], at, n);
                at[bound] = Fu(1);
                for (uint32_t v = 0; v < nk; v++)
                    vstep[v] = fsub(virtual_at_point(l->virtual_kinds[v], at, n), vcur[v]);
            }
            const fe eq_step = fsub(eq[2 * i + 1], eq[2 * i]);
            fe eq_at = eq[2 * i];
            for (int x = 0; x < 4; x++) {
                acc[x] = fadd(acc[x], fmul(eq_at, resolved_summand(&r, w, cur, cur + ncols, vcur)));
                eq_at = fadd(eq_at, eq_step);
                for (size_t t = 0; t < nt; t++) cur[t] = fadd(cur[t], step[t]);
                for (uint32_t v = 0; v < nk; v++) vcur[v] = fadd(vcur[v], vstep[v]);
            }
        }
        fe g[4];
        interpolate_cubic(acc, g);
        orc_tr_append_scalars(tr, ORC_TAG_SUMCHECK_ROUND, (const u64 *)g, 4);
        fe rho;
        orc_tr_challenge(tr, ORC_TAG_SUMCHECK_CHALLENGE, rho.v);
        for (size_t t = 0; t < nt; t++) bind(T[t], len, rho);
        bind(eq, len, rho);
        memcpy(rounds_out + 16 * j, g, sizeof g);
        memcpy(point_out + 4 * j, rho.v, 32);
        at[j] = rho;
        len = half;
    }
    resolved_free(&r);
    free(at);
    free(vcur);
    free(vstep);
    free(cur);
    free(step);
    free(T);
    free(eq);
    free(w);
}
