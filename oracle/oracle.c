/*
 * oracle.c — CPU restatement of apogee-vm's per-shard proving hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the tests, the smoke entry
 * and bench.py's cpu_baseline leg use; the product library never links or
 * calls it. Every function restates the reference file:line named above it,
 * with the reference's own limb width (4 x u64, unsigned __int128
 * intermediates), so a mismatch against the GPU path is a bug in one of the
 * two and the KAT replays in tests/ say which.
 *
 * Pins (tests/test_oracle_*.py): fr_kats, fq_kats, g1_kats, msm_kats,
 * poly_kats, poseidon2_perm, transcript_cases, g1_absorb_kats and the
 * mercury_proof commitment from the reference's committed vectors; the
 * sumcheck round polynomials against the brute-force definition
 * (crates/sumcheck/tests/oracle.rs:45-116) at n <= 5 plus the round-sum
 * invariants; the degree-4 interpolation against evaluation at its nodes;
 * ChaCha20 against the RFC 8439 block vector.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "poseidon2_rc3.h"

typedef uint64_t u64;
typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------
 * Field parameters: crates/constants/src/lib.rs:22-57 (Fr), :97-143 (Fq).
 * ---------------------------------------------------------------------- */
typedef struct {
    u64 m[4];
    u64 r[4];
    u64 r2[4];
    u64 inv;
} fparams;

static const fparams FR_P = {
    {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL},
    {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL},
    0xc2e1f593efffffffULL};
static const fparams FQ_P = {
    {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL},
    {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL},
    0x87d20782e4866389ULL};

static const fparams *fp(int which) { return which ? &FQ_P : &FR_P; }

/* crates/field/src/lib.rs:40-63: mac / adc / sbb on u128. */
static inline u64 mac(u64 acc, u64 a, u64 b, u64 carry, u64 *hi) {
    u128 t = (u128)acc + (u128)a * b + carry;
    *hi = (u64)(t >> 64);
    return (u64)t;
}
static inline u64 adc(u64 a, u64 b, u64 carry, u64 *c) {
    u128 t = (u128)a + b + carry;
    *c = (u64)(t >> 64);
    return (u64)t;
}
static inline u64 sbb(u64 a, u64 b, u64 borrow, u64 *bo) {
    u128 t = (u128)a - b - borrow;
    *bo = (u64)(t >> 64) & 1;
    return (u64)t;
}

/* crates/field/src/lib.rs:66-89: is_ge_modulus, reduce_once */
static int ge_mod(const u64 *a, const u64 *m) {
    for (int i = 3; i >= 0; i--)
        if (a[i] != m[i]) return a[i] > m[i];
    return 1;
}
static void reduce_once(u64 *a, const u64 *m) {
    if (ge_mod(a, m)) {
        u64 b = 0;
        for (int i = 0; i < 4; i++) a[i] = sbb(a[i], m[i], b, &b);
    }
}

static int is_zero4(const u64 *a) { return (a[0] | a[1] | a[2] | a[3]) == 0; }
static int eq4(const u64 *a, const u64 *b) {
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2] && a[3] == b[3];
}

/* crates/field/src/lib.rs:92-108 add_limbs */
static void f_add(const fparams *P, const u64 *a, const u64 *b, u64 *out) {
    u64 r[4], c = 0;
    for (int i = 0; i < 4; i++) r[i] = adc(a[i], b[i], c, &c);
    reduce_once(r, P->m);
    memcpy(out, r, 32);
}
/* crates/field/src/lib.rs:111-128 sub_limbs */
static void f_sub(const fparams *P, const u64 *a, const u64 *b, u64 *out) {
    u64 r[4], bo = 0;
    for (int i = 0; i < 4; i++) r[i] = sbb(a[i], b[i], bo, &bo);
    if (bo) {
        u64 c = 0;
        for (int i = 0; i < 4; i++) r[i] = adc(r[i], P->m[i], c, &c);
    }
    memcpy(out, r, 32);
}
/* crates/field/src/lib.rs:131-137 neg_limbs */
static void f_neg(const fparams *P, const u64 *a, u64 *out) {
    if (is_zero4(a)) {
        memset(out, 0, 32);
        return;
    }
    f_sub(P, P->m, a, out);
}
/* crates/field/src/lib.rs:148-184 mont_mul (CIOS over four u64 limbs) */
static void f_mul(const fparams *P, const u64 *a, const u64 *b, u64 *out) {
    u64 t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u64 bi = b[i], carry = 0, c;
        for (int j = 0; j < 4; j++) t[j] = mac(t[j], a[j], bi, carry, &carry);
        t[4] = adc(t[4], carry, 0, &c);
        t[5] = c;
        u64 m = t[0] * P->inv;
        (void)mac(t[0], m, P->m[0], 0, &carry);
        for (int j = 1; j < 4; j++) t[j - 1] = mac(t[j], m, P->m[j], carry, &carry);
        t[3] = adc(t[4], carry, 0, &c);
        t[4] = t[5] + c;
    }
    u64 r[4] = {t[0], t[1], t[2], t[3]};
    reduce_once(r, P->m);
    memcpy(out, r, 32);
}
static void f_sqr(const fparams *P, const u64 *a, u64 *out) { f_mul(P, a, a, out); }
/* crates/field/src/lib.rs:207-209 from_u64 */
static void f_from_u64(const fparams *P, u64 x, u64 *out) {
    u64 a[4] = {x, 0, 0, 0};
    f_mul(P, a, P->r2, out);
}
/* crates/field/src/lib.rs:217-228 pow (square-and-multiply, MSB first) */
static void f_pow(const fparams *P, const u64 *a, const u64 *e, u64 *out) {
    u64 acc[4], base[4];
    memcpy(base, a, 32);
    memcpy(acc, P->r, 32);
    for (int l = 3; l >= 0; l--)
        for (int bit = 63; bit >= 0; bit--) {
            f_sqr(P, acc, acc);
            if ((e[l] >> bit) & 1) f_mul(P, acc, base, acc);
        }
    memcpy(out, acc, 32);
}
/* crates/field/src/lib.rs:240-249 inverse (Fermat, p - 2) */
static int f_inv(const fparams *P, const u64 *a, u64 *out) {
    if (is_zero4(a)) {
        memset(out, 0, 32);
        return 0;
    }
    u64 e[4];
    memcpy(e, P->m, 32);
    e[0] -= 2; /* both moduli end in ...01 / ...47, no borrow */
    f_pow(P, a, e, out);
    return 1;
}
/* crates/field/src/lib.rs:252-260 to_bytes (canonical LE) */
static void f_to_bytes(const fparams *P, const u64 *a, uint8_t *o) {
    u64 one[4] = {1, 0, 0, 0}, c[4];
    f_mul(P, a, one, c);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 8; k++) o[8 * i + k] = (uint8_t)(c[i] >> (8 * k));
}
/* crates/field/src/lib.rs:296-307 from_bytes (rejects >= modulus) */
static int f_from_bytes(const fparams *P, const uint8_t *b, u64 *out) {
    u64 l[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 8; k++) l[i] |= (u64)b[8 * i + k] << (8 * k);
    if (ge_mod(l, P->m)) return 0;
    f_mul(P, l, P->r2, out);
    return 1;
}
/* crates/field/src/lib.rs:281-294 from_hex: 64 big-endian hex digits */
static int hexval(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    return -1;
}
static int f_from_hex(const fparams *P, const char *s, u64 *out) {
    uint8_t le[32];
    for (int i = 0; i < 32; i++) {
        int hi = hexval(s[2 * i]), lo = hexval(s[2 * i + 1]);
        if (hi < 0 || lo < 0) return 0;
        le[31 - i] = (uint8_t)((hi << 4) | lo);
    }
    return f_from_bytes(P, le, out);
}

void orc_mul(int w, const u64 *a, const u64 *b, u64 *o) { f_mul(fp(w), a, b, o); }
void orc_add(int w, const u64 *a, const u64 *b, u64 *o) { f_add(fp(w), a, b, o); }
void orc_sub(int w, const u64 *a, const u64 *b, u64 *o) { f_sub(fp(w), a, b, o); }
void orc_neg(int w, const u64 *a, u64 *o) { f_neg(fp(w), a, o); }
void orc_square(int w, const u64 *a, u64 *o) { f_sqr(fp(w), a, o); }
int orc_inverse(int w, const u64 *a, u64 *o) { return f_inv(fp(w), a, o); }
void orc_pow(int w, const u64 *a, const u64 *e, u64 *o) { f_pow(fp(w), a, e, o); }
void orc_from_u64(int w, u64 x, u64 *o) { f_from_u64(fp(w), x, o); }
void orc_to_bytes(int w, const u64 *a, uint8_t *o) { f_to_bytes(fp(w), a, o); }
int orc_from_bytes(int w, const uint8_t *b, u64 *o) { return f_from_bytes(fp(w), b, o); }

/* Fr / Fq shorthands */
#define FR (&FR_P)
#define FQ (&FQ_P)

/* ------------------------------------------------------------------------
 * G1: crates/curve/src/g1.rs. Jacobian point = X[4] Y[4] Z[4]; affine = x[4]
 * y[4] with all-zero meaning infinity ((0,0) is off the curve, g1.rs:92-127).
 * ---------------------------------------------------------------------- */
typedef struct {
    u64 x[4], y[4], z[4];
} jac;
typedef struct {
    u64 x[4], y[4];
} aff;

static int aff_is_inf(const aff *p) { return is_zero4(p->x) && is_zero4(p->y); }
static int jac_is_id(const jac *p) { return is_zero4(p->z); }
static void jac_identity(jac *p) {
    memcpy(p->x, FQ_P.r, 32);
    memcpy(p->y, FQ_P.r, 32);
    memset(p->z, 0, 32);
}
static void jac_from_aff(const aff *a, jac *p) {
    if (aff_is_inf(a)) {
        jac_identity(p);
        return;
    }
    memcpy(p->x, a->x, 32);
    memcpy(p->y, a->y, 32);
    memcpy(p->z, FQ_P.r, 32);
}

/* g1.rs:246-277 dbl-2009-l */
static void jac_double(const jac *p, jac *o) {
    if (jac_is_id(p)) {
        jac_identity(o);
        return;
    }
    u64 a[4], b[4], c[4], d[4], e[4], f[4], t[4], x3[4], y3[4], z3[4];
    f_sqr(FQ, p->x, a);
    f_sqr(FQ, p->y, b);
    f_sqr(FQ, b, c);
    f_add(FQ, p->x, b, t);
    f_sqr(FQ, t, t);
    f_sub(FQ, t, a, t);
    f_sub(FQ, t, c, t);
    f_add(FQ, t, t, d);
    f_add(FQ, a, a, e);
    f_add(FQ, e, a, e);
    f_sqr(FQ, e, f);
    f_sub(FQ, f, d, x3);
    f_sub(FQ, x3, d, x3);
    u64 c8[4];
    f_add(FQ, c, c, c8);
    f_add(FQ, c8, c8, c8);
    f_add(FQ, c8, c8, c8);
    f_sub(FQ, d, x3, t);
    f_mul(FQ, e, t, y3);
    f_sub(FQ, y3, c8, y3);
    f_mul(FQ, p->y, p->z, t);
    f_add(FQ, t, t, z3);
    memcpy(o->x, x3, 32);
    memcpy(o->y, y3, 32);
    memcpy(o->z, z3, 32);
}

/* g1.rs:155-197 add-2007-bl with the H == 0 branches */
static void jac_add(const jac *p, const jac *q, jac *o) {
    if (jac_is_id(p)) {
        *o = *q;
        return;
    }
    if (jac_is_id(q)) {
        *o = *p;
        return;
    }
    u64 z1z1[4], z2z2[4], u1[4], u2[4], s1[4], s2[4], h[4], r[4], i[4], j[4], v[4], t[4];
    f_sqr(FQ, p->z, z1z1);
    f_sqr(FQ, q->z, z2z2);
    f_mul(FQ, p->x, z2z2, u1);
    f_mul(FQ, q->x, z1z1, u2);
    f_mul(FQ, p->y, z2z2, s1);
    f_mul(FQ, s1, q->z, s1);
    f_mul(FQ, q->y, z1z1, s2);
    f_mul(FQ, s2, p->z, s2);
    f_sub(FQ, u2, u1, h);
    f_sub(FQ, s2, s1, t);
    f_add(FQ, t, t, r);
    if (is_zero4(h)) {
        if (is_zero4(r)) {
            jac_double(p, o);
            return;
        }
        jac_identity(o);
        return;
    }
    f_add(FQ, h, h, t);
    f_sqr(FQ, t, i);
    f_mul(FQ, h, i, j);
    f_mul(FQ, u1, i, v);
    u64 x3[4], y3[4], z3[4], s1j[4];
    f_sqr(FQ, r, x3);
    f_sub(FQ, x3, j, x3);
    f_sub(FQ, x3, v, x3);
    f_sub(FQ, x3, v, x3);
    f_mul(FQ, s1, j, s1j);
    f_sub(FQ, v, x3, t);
    f_mul(FQ, r, t, y3);
    f_add(FQ, s1j, s1j, t);
    f_sub(FQ, y3, t, y3);
    f_add(FQ, p->z, q->z, t);
    f_sqr(FQ, t, t);
    f_sub(FQ, t, z1z1, t);
    f_sub(FQ, t, z2z2, t);
    f_mul(FQ, t, h, z3);
    memcpy(o->x, x3, 32);
    memcpy(o->y, y3, 32);
    memcpy(o->z, z3, 32);
}

/* g1.rs:203-243 madd-2007-bl */
static void jac_add_aff(const jac *p, const aff *q, jac *o) {
    if (aff_is_inf(q)) {
        *o = *p;
        return;
    }
    if (jac_is_id(p)) {
        jac_from_aff(q, o);
        return;
    }
    u64 z1z1[4], u2[4], s2[4], h[4], r[4], t[4];
    f_sqr(FQ, p->z, z1z1);
    f_mul(FQ, q->x, z1z1, u2);
    f_mul(FQ, q->y, z1z1, s2);
    f_mul(FQ, s2, p->z, s2);
    f_sub(FQ, u2, p->x, h);
    f_sub(FQ, s2, p->y, t);
    f_add(FQ, t, t, r);
    if (is_zero4(h)) {
        if (is_zero4(r)) {
            jac_double(p, o);
            return;
        }
        jac_identity(o);
        return;
    }
    u64 hh[4], i[4], j[4], v[4], x3[4], y3[4], z3[4], y1j[4];
    f_sqr(FQ, h, hh);
    f_add(FQ, hh, hh, i);
    f_add(FQ, i, i, i);
    f_mul(FQ, h, i, j);
    f_mul(FQ, p->x, i, v);
    f_sqr(FQ, r, x3);
    f_sub(FQ, x3, j, x3);
    f_sub(FQ, x3, v, x3);
    f_sub(FQ, x3, v, x3);
    f_mul(FQ, p->y, j, y1j);
    f_sub(FQ, v, x3, t);
    f_mul(FQ, r, t, y3);
    f_add(FQ, y1j, y1j, t);
    f_sub(FQ, y3, t, y3);
    f_add(FQ, p->z, h, t);
    f_sqr(FQ, t, t);
    f_sub(FQ, t, z1z1, t);
    f_sub(FQ, t, hh, z3);
    memcpy(o->x, x3, 32);
    memcpy(o->y, y3, 32);
    memcpy(o->z, z3, 32);
}

/* g1.rs:291-307 mul_le_bytes: fixed 4-bit window over sixteen multiples */
static void jac_mul_le(const jac *p, const uint8_t *k, jac *o) {
    jac table[16];
    jac_identity(&table[0]);
    table[1] = *p;
    for (int i = 2; i < 16; i++) jac_add(&table[i - 1], p, &table[i]);
    jac acc;
    jac_identity(&acc);
    for (int bi = 31; bi >= 0; bi--) {
        uint8_t nib[2] = {(uint8_t)(k[bi] >> 4), (uint8_t)(k[bi] & 0x0f)};
        for (int q = 0; q < 2; q++) {
            for (int d = 0; d < 4; d++) jac_double(&acc, &acc);
            jac_add(&acc, &table[nib[q]], &acc);
        }
    }
    *o = acc;
}

/* g1.rs:310-323 to_affine */
static void jac_to_aff(const jac *p, aff *o) {
    u64 zi[4];
    if (!f_inv(FQ, p->z, zi)) {
        memset(o, 0, sizeof(*o));
        return;
    }
    u64 zi2[4], zi3[4];
    f_sqr(FQ, zi, zi2);
    f_mul(FQ, zi2, zi, zi3);
    f_mul(FQ, p->x, zi2, o->x);
    f_mul(FQ, p->y, zi3, o->y);
}

static void aff_neg(const aff *a, aff *o) {
    if (aff_is_inf(a)) {
        *o = *a;
        return;
    }
    memcpy(o->x, a->x, 32);
    f_neg(FQ, a->y, o->y);
}

void orc_g1_add(const u64 *p, const u64 *q, u64 *o) { jac_add((const jac *)p, (const jac *)q, (jac *)o); }
void orc_g1_add_affine(const u64 *p, const u64 *q, u64 *o) {
    jac_add_aff((const jac *)p, (const aff *)q, (jac *)o);
}
void orc_g1_double(const u64 *p, u64 *o) { jac_double((const jac *)p, (jac *)o); }
void orc_g1_mul_le_bytes(const u64 *p, const uint8_t *k, u64 *o) {
    jac_mul_le((const jac *)p, k, (jac *)o);
}
void orc_g1_mul(const u64 *p, const u64 *s, u64 *o) {
    uint8_t k[32];
    f_to_bytes(FR, s, k);
    jac_mul_le((const jac *)p, k, (jac *)o);
}
void orc_g1_to_affine(const u64 *p, u64 *o) { jac_to_aff((const jac *)p, (aff *)o); }
void orc_g1_from_affine(const u64 *a, u64 *o) { jac_from_aff((const aff *)a, (jac *)o); }
void orc_g1_generator(u64 *o) {
    /* (1, 2): g1.rs:23-28 and :60-64 */
    memcpy(o, FQ_P.r, 32);
    f_from_u64(FQ, 2, o + 4);
}
void orc_g1_affine_to_wire(const u64 *a, uint8_t *o) {
    if (aff_is_inf((const aff *)a)) {
        memset(o, 0, 64);
        return;
    }
    f_to_bytes(FQ, a, o);
    f_to_bytes(FQ, a + 4, o + 32);
}
/* g1.rs:110-127 from_bytes: canonical coordinates, on the curve (cofactor 1) */
int orc_g1_affine_from_wire(const uint8_t *in, u64 *o) {
    int allz = 1;
    for (int i = 0; i < 64; i++)
        if (in[i]) allz = 0;
    if (allz) {
        memset(o, 0, 64);
        return 1;
    }
    u64 x[4], y[4];
    if (!f_from_bytes(FQ, in, x) || !f_from_bytes(FQ, in + 32, y)) return 0;
    u64 y2[4], x3[4], b[4];
    f_sqr(FQ, y, y2);
    f_sqr(FQ, x, x3);
    f_mul(FQ, x3, x, x3);
    f_from_u64(FQ, 3, b);
    f_add(FQ, x3, b, x3);
    if (!eq4(y2, x3)) return 0;
    memcpy(o, x, 32);
    memcpy(o + 4, y, 32);
    return 1;
}

/* ------------------------------------------------------------------------
 * MSM: crates/curve/src/msm.rs
 * ---------------------------------------------------------------------- */

/* msm.rs:112-128 window_bits / log2 (arkworks' rule) */
size_t orc_msm_window_bits(size_t n) {
    if (n < 32) return 3;
    size_t lg;
    if ((n & (n - 1)) == 0) {
        lg = (size_t)__builtin_ctzll(n);
    } else {
        lg = 64 - (size_t)__builtin_clzll(n);
    }
    return (lg * 69) / 100 + 2;
}
/* msm.rs:140-142 window_count */
static size_t window_count(size_t bits, size_t w) { return (bits + 1 + w - 1) / w; }

/* msm.rs:162-173 window_of */
static u64 window_of(const u64 *l, size_t off, size_t w) {
    size_t limb = off / 64, shift = off % 64;
    u64 v = l[limb] >> shift;
    if (shift > 0 && limb + 1 < 4) v |= l[limb + 1] << (64 - shift);
    return v & ((1ULL << w) - 1);
}
/* msm.rs:180-198 recode */
static void recode(const u64 *l, size_t w, int32_t *out, size_t nd) {
    u64 half = 1ULL << (w - 1);
    int64_t radix = (int64_t)1 << w;
    u64 carry = 0;
    for (size_t i = 0; i < nd; i++) {
        u64 coef = window_of(l, i * w, w) + carry;
        if (i == nd - 1) {
            out[i] = (int32_t)coef;
        } else if (coef >= half) {
            out[i] = (int32_t)((int64_t)coef - radix);
            carry = 1;
        } else {
            out[i] = (int32_t)coef;
            carry = 0;
        }
    }
}
/* msm.rs:203-224 recode_u32 */
static void recode_u32(uint32_t s, size_t w, int32_t *out, size_t nd) {
    u64 half = 1ULL << (w - 1), mask = (1ULL << w) - 1, value = s;
    int64_t radix = (int64_t)1 << w;
    u64 carry = 0;
    for (size_t i = 0; i < nd; i++) {
        u64 coef = ((i * w < 64 ? (value >> (i * w)) : 0) & mask) + carry;
        if (i == nd - 1) {
            out[i] = (int32_t)coef;
        } else if (coef >= half) {
            out[i] = (int32_t)((int64_t)coef - radix);
            carry = 1;
        } else {
            out[i] = (int32_t)coef;
            carry = 0;
        }
    }
}

/* msm.rs:277-311 window_sum: one window over one input slice */
static void window_sum(const aff *bases, size_t len, const int32_t *digits, size_t windows,
                       size_t window, size_t buckets, jac *out) {
    if (len == 0) {
        jac_identity(out);
        return;
    }
    jac *bucket = (jac *)malloc(buckets * sizeof(jac));
    for (size_t b = 0; b < buckets; b++) jac_identity(&bucket[b]);
    for (size_t j = 0; j < len; j++) {
        int32_t d = digits[j * windows + window];
        if (d == 0) continue;
        jac *slot = &bucket[(size_t)(d > 0 ? d : -d) - 1];
        if (d > 0) {
            jac_add_aff(slot, &bases[j], slot);
        } else {
            aff nb;
            aff_neg(&bases[j], &nb);
            jac_add_aff(slot, &nb, slot);
        }
    }
    jac running, total;
    jac_identity(&running);
    jac_identity(&total);
    for (size_t b = buckets; b-- > 0;) {
        jac_add(&running, &bucket[b], &running);
        jac_add(&total, &running, &total);
    }
    free(bucket);
    *out = total;
}

typedef struct {
    const aff *bases;
    const int32_t *digits;
    size_t n, windows, chunks, chunk_len, buckets, ntasks, stride, first;
    jac *partials;
} pip_job;

static void *pip_worker(void *arg) {
    pip_job *j = (pip_job *)arg;
    for (size_t task = j->first; task < j->ntasks; task += j->stride) {
        size_t start = (task % j->chunks) * j->chunk_len;
        if (start > j->n) start = j->n;
        size_t end = start + j->chunk_len;
        if (end > j->n) end = j->n;
        window_sum(j->bases + start, end - start, j->digits + start * j->windows, j->windows,
                   task / j->chunks, j->buckets, &j->partials[task]);
    }
    return NULL;
}

/* msm.rs:232-273 pippenger: (window x chunk) tasks, then Horner */
static void pippenger(const aff *bases, const int32_t *digits, size_t n, size_t windows,
                      size_t w, size_t threads, size_t workers, jac *out) {
    size_t buckets = (size_t)1 << (w - 1);
    size_t chunks = (threads + windows - 1) / windows;
    size_t hi = n / buckets;
    if (hi < 1) hi = 1;
    if (chunks < 1) chunks = 1;
    if (chunks > hi) chunks = hi;
    size_t chunk_len = (n + chunks - 1) / chunks;
    size_t ntasks = windows * chunks;
    jac *partials = (jac *)malloc(ntasks * sizeof(jac));
    if (workers < 1) workers = 1;
    if (workers > ntasks) workers = ntasks;
    pip_job *jobs = (pip_job *)malloc(workers * sizeof(pip_job));
    pthread_t *tids = (pthread_t *)malloc(workers * sizeof(pthread_t));
    for (size_t k = 0; k < workers; k++) {
        pip_job jb = {bases, digits, n, windows, chunks, chunk_len, buckets, ntasks, workers, k,
                      partials};
        jobs[k] = jb;
        if (workers == 1)
            pip_worker(&jobs[k]);
        else
            pthread_create(&tids[k], NULL, pip_worker, &jobs[k]);
    }
    if (workers > 1)
        for (size_t k = 0; k < workers; k++) pthread_join(tids[k], NULL);
    jac acc;
    jac_identity(&acc);
    for (size_t win = windows; win-- > 0;) {
        for (size_t d = 0; d < w; d++) jac_double(&acc, &acc);
        for (size_t c = 0; c < chunks; c++) jac_add(&acc, &partials[win * chunks + c], &acc);
    }
    *out = acc;
    free(partials);
    free(jobs);
    free(tids);
}

/* One Montgomery Fr scalar's signed digits under the reference's recoding
 * (msm.rs:149-198) at window width w; returns the window count. */
size_t orc_msm_recode(const u64 *scalar_mont, size_t w, int32_t *digits) {
    const size_t windows = window_count(254, w);
    u64 one[4] = {1, 0, 0, 0}, c[4];
    f_mul(FR, scalar_mont, one, c);
    recode(c, w, digits, windows);
    return windows;
}

/* Nonzero signed digits of n Montgomery Fr scalars at window width w (0: the
 * reference's rule for n, msm.rs:112-118): the digit count in the survey's
 * algorithmic MSM work, one mixed addition per nonzero digit. */
uint64_t orc_msm_nonzero_digits(const u64 *scalars, size_t n, size_t w) {
    if (w == 0) w = orc_msm_window_bits(n);
    const size_t windows = window_count(254, w);
    int32_t d[256];
    uint64_t nz = 0;
    for (size_t i = 0; i < n; i++) {
        orc_msm_recode(scalars + 4 * i, w, d);
        for (size_t j = 0; j < windows; j++) nz += d[j] != 0;
    }
    return nz;
}

/* msm.rs:48-67 msm: recode every scalar's canonical limbs, then pippenger */
int orc_msm(const u64 *bases, const u64 *scalars, size_t n, size_t threads, size_t workers,
            u64 *out) {
    if (n == 0) {
        jac_identity((jac *)out);
        return 0;
    }
    size_t w = orc_msm_window_bits(n);
    size_t windows = window_count(254, w);
    int32_t *digits = (int32_t *)malloc(n * windows * sizeof(int32_t));
    u64 one[4] = {1, 0, 0, 0};
    for (size_t i = 0; i < n; i++) {
        u64 c[4];
        f_mul(FR, scalars + 4 * i, one, c); /* canonical_limbs, msm.rs:149-159 */
        recode(c, w, digits + i * windows, windows);
    }
    pippenger((const aff *)bases, digits, n, windows, w, threads, workers, (jac *)out);
    free(digits);
    return 0;
}
/* msm.rs:79-100 msm_small_u32 */
int orc_msm_small_u32(const u64 *bases, const uint32_t *scalars, size_t n, size_t threads,
                      size_t workers, u64 *out) {
    if (n == 0) {
        jac_identity((jac *)out);
        return 0;
    }
    size_t w = orc_msm_window_bits(n);
    size_t windows = window_count(32, w);
    int32_t *digits = (int32_t *)malloc(n * windows * sizeof(int32_t));
    for (size_t i = 0; i < n; i++) recode_u32(scalars[i], w, digits + i * windows, windows);
    pippenger((const aff *)bases, digits, n, windows, w, threads, workers, (jac *)out);
    free(digits);
    return 0;
}

typedef struct {
    const u64 *powers;
    u64 *out;
    size_t n, stride, first;
} srs_job;
static void *srs_worker(void *arg) {
    srs_job *j = (srs_job *)arg;
    jac g;
    u64 ga[8];
    orc_g1_generator(ga);
    jac_from_aff((const aff *)ga, &g);
    for (size_t i = j->first; i < j->n; i += j->stride) {
        jac p;
        uint8_t k[32];
        f_to_bytes(FR, j->powers + 4 * i, k);
        jac_mul_le(&g, k, &p);
        jac_to_aff(&p, (aff *)(j->out + 8 * i));
    }
    return NULL;
}
/* [s_i]G for every scalar, affine: crates/curve/tests/msm.rs:41-55 builds
 * KAT bases this way (G1Projective::GENERATOR.mul(k).to_affine()). */
void orc_g1_gen_mul_batch(const u64 *scalars, size_t n, size_t workers, u64 *out) {
    if (workers < 1) workers = 1;
    srs_job *jobs = (srs_job *)malloc(workers * sizeof(srs_job));
    pthread_t *tids = (pthread_t *)malloc(workers * sizeof(pthread_t));
    for (size_t k = 0; k < workers; k++) {
        srs_job jb = {scalars, out, n, workers, k};
        jobs[k] = jb;
        if (workers == 1)
            srs_worker(&jobs[k]);
        else
            pthread_create(&tids[k], NULL, srs_worker, &jobs[k]);
    }
    if (workers > 1)
        for (size_t k = 0; k < workers; k++) pthread_join(tids[k], NULL);
    free(jobs);
    free(tids);
}

/* crates/pcs/tests/common/mod.rs:43-60 toy_srs: [tau^i]G */
void orc_srs_powers(const u64 *tau, size_t n, size_t workers, u64 *out) {
    u64 *pw = (u64 *)malloc(n * 32 + 32);
    u64 acc[4];
    memcpy(acc, FR_P.r, 32);
    for (size_t i = 0; i < n; i++) {
        memcpy(pw + 4 * i, acc, 32);
        f_mul(FR, acc, tau, acc);
    }
    if (workers < 1) workers = 1;
    srs_job *jobs = (srs_job *)malloc(workers * sizeof(srs_job));
    pthread_t *tids = (pthread_t *)malloc(workers * sizeof(pthread_t));
    for (size_t k = 0; k < workers; k++) {
        srs_job jb = {pw, out, n, workers, k};
        jobs[k] = jb;
        if (workers == 1)
            srs_worker(&jobs[k]);
        else
            pthread_create(&tids[k], NULL, srs_worker, &jobs[k]);
    }
    if (workers > 1)
        for (size_t k = 0; k < workers; k++) pthread_join(tids[k], NULL);
    free(jobs);
    free(tids);
    free(pw);
}

/* ------------------------------------------------------------------------
 * poly: crates/poly/src/lib.rs
 * ---------------------------------------------------------------------- */
/* lib.rs:156-185 bind (the Fr-backed branch), in place */
void orc_bind(u64 *v, size_t len, const u64 *r) {
    size_t half = len / 2;
    for (size_t i = 0; i < half; i++) {
        u64 lo[4], hi[4], d[4];
        memcpy(lo, v + 8 * i, 32);
        memcpy(hi, v + 8 * i + 4, 32);
        f_sub(FR, hi, lo, d);
        f_mul(FR, r, d, d);
        f_add(FR, lo, d, v + 4 * i);
    }
}
/* lib.rs:233-248 eq_table by doubling */
void orc_eq_table(const u64 *r, size_t n, u64 *t) {
    memcpy(t, FR_P.r, 32);
    for (size_t j = 0; j < n; j++) {
        size_t half = (size_t)1 << j;
        for (size_t i = 0; i < half; i++) {
            u64 v[4], w1[4];
            memcpy(v, t + 4 * i, 32);
            f_mul(FR, v, r + 4 * j, w1);
            memcpy(t + 4 * (i + half), w1, 32);
            f_sub(FR, v, w1, t + 4 * i);
        }
    }
}
/* lib.rs:252-265 eq_eval */
void orc_eq_eval(const u64 *r, const u64 *y, size_t n, u64 *out) {
    u64 acc[4];
    memcpy(acc, FR_P.r, 32);
    for (size_t j = 0; j < n; j++) {
        u64 a[4], b[4], c[4], t[4];
        f_mul(FR, r + 4 * j, y + 4 * j, a);
        f_sub(FR, FR_P.r, r + 4 * j, b);
        f_sub(FR, FR_P.r, y + 4 * j, c);
        f_mul(FR, b, c, t);
        f_add(FR, a, t, t);
        f_mul(FR, acc, t, acc);
    }
    memcpy(out, acc, 32);
}
/* lib.rs:189-219 evaluate (non-destructive fold) */
void orc_evaluate(const u64 *table, size_t num_vars, const u64 *point, u64 *out) {
    size_t len = (size_t)1 << num_vars;
    u64 *s = (u64 *)malloc(len * 32);
    memcpy(s, table, len * 32);
    for (size_t j = 0; j < num_vars; j++) {
        orc_bind(s, len, point + 4 * j);
        len /= 2;
    }
    memcpy(out, s, 32);
    free(s);
}

/* ------------------------------------------------------------------------
 * transcript: crates/transcript/src/lib.rs
 * ---------------------------------------------------------------------- */
static u64 RC_INI[4][3][4], RC_MID[56][4], RC_TER[4][3][4];
static pthread_once_t rc_once = PTHREAD_ONCE_INIT;
static void rc_init(void) {
    for (int r = 0; r < 4; r++)
        for (int l = 0; l < 3; l++) {
            f_from_hex(FR, ORC_RC3_INITIAL[r][l], RC_INI[r][l]);
            f_from_hex(FR, ORC_RC3_TERMINAL[r][l], RC_TER[r][l]);
        }
    for (int i = 0; i < 56; i++) f_from_hex(FR, ORC_RC3_INTERNAL[i], RC_MID[i]);
}
/* lib.rs:45-50 sbox x^5 */
static void sbox(u64 *x) {
    u64 x2[4], x4[4];
    f_sqr(FR, x, x2);
    f_sqr(FR, x2, x4);
    f_mul(FR, x4, x, x);
}
/* lib.rs:55-61 external matrix circ(2,1,1) */
static void ext_matrix(u64 s[3][4]) {
    u64 sum[4];
    f_add(FR, s[0], s[1], sum);
    f_add(FR, sum, s[2], sum);
    for (int i = 0; i < 3; i++) f_add(FR, s[i], sum, s[i]);
}
/* lib.rs:66-72 internal matrix 1 + diag(1,1,2) */
static void int_matrix(u64 s[3][4]) {
    u64 sum[4];
    f_add(FR, s[0], s[1], sum);
    f_add(FR, sum, s[2], sum);
    f_add(FR, s[0], sum, s[0]);
    f_add(FR, s[1], sum, s[1]);
    f_add(FR, s[2], s[2], s[2]);
    f_add(FR, s[2], sum, s[2]);
}
/* lib.rs:121-146 poseidon2_permute */
void orc_poseidon2_permute(u64 *st) {
    pthread_once(&rc_once, rc_init);
    u64(*s)[4] = (u64(*)[4])st;
    ext_matrix(s);
    for (int r = 0; r < 4; r++) {
        for (int l = 0; l < 3; l++) {
            f_add(FR, s[l], RC_INI[r][l], s[l]);
            sbox(s[l]);
        }
        ext_matrix(s);
    }
    for (int i = 0; i < 56; i++) {
        f_add(FR, s[0], RC_MID[i], s[0]);
        sbox(s[0]);
        int_matrix(s);
    }
    for (int r = 0; r < 4; r++) {
        for (int l = 0; l < 3; l++) {
            f_add(FR, s[l], RC_TER[r][l], s[l]);
            sbox(s[l]);
        }
        ext_matrix(s);
    }
}
void orc_tr_init(orc_transcript *t) { memset(t, 0, sizeof(*t)); }
/* lib.rs:273-292 duplex */
static void duplex(orc_transcript *t) {
    size_t n = (size_t)t->input_len;
    for (size_t i = 0; i < n; i++) {
        memcpy(t->state[i], t->input[i], 32);
        memset(t->input[i], 0, 32);
    }
    t->input_len = 0;
    if (n > 0) {
        for (size_t l = n; l < 2; l++) memset(t->state[l], 0, 32);
        u64 c[4];
        f_from_u64(FR, n, c);
        f_add(FR, t->state[2], c, t->state[2]);
    }
    orc_poseidon2_permute(&t->state[0][0]);
    memcpy(t->output[0], t->state[0], 32);
    memcpy(t->output[1], t->state[1], 32);
    t->output_len = 2;
}
/* lib.rs:241-250 observe */
void orc_tr_observe(orc_transcript *t, const u64 *x) {
    memset(t->output, 0, sizeof(t->output));
    t->output_len = 0;
    memcpy(t->input[t->input_len], x, 32);
    t->input_len++;
    if (t->input_len == 2) duplex(t);
}
/* lib.rs:253-262 sample */
void orc_tr_sample(orc_transcript *t, u64 *out) {
    if (t->input_len > 0 || t->output_len == 0) duplex(t);
    t->output_len--;
    memcpy(out, t->output[t->output_len], 32);
    memset(t->output[t->output_len], 0, 32);
}
static void observe_u64(orc_transcript *t, u64 v) {
    u64 x[4];
    f_from_u64(FR, v, x);
    orc_tr_observe(t, x);
}
/* lib.rs:306-316 append_scalars */
void orc_tr_append_scalars(orc_transcript *t, u64 tag, const u64 *xs, size_t n) {
    observe_u64(t, tag);
    observe_u64(t, n);
    for (size_t i = 0; i < n; i++) orc_tr_observe(t, xs + 4 * i);
}
/* lib.rs:323-338 append_bytes: 31-byte little-endian chunks */
void orc_tr_append_bytes(orc_transcript *t, u64 tag, const uint8_t *b, size_t len) {
    observe_u64(t, tag);
    observe_u64(t, len);
    for (size_t off = 0; off < len; off += 31) {
        uint8_t limb[32];
        memset(limb, 0, 32);
        size_t c = len - off < 31 ? len - off : 31;
        memcpy(limb, b + off, c);
        u64 x[4];
        f_from_bytes(FR, limb, x);
        orc_tr_observe(t, x);
    }
}
/* lib.rs:341-346 challenge_scalar */
void orc_tr_challenge(orc_transcript *t, u64 tag, u64 *out) {
    observe_u64(t, tag);
    orc_tr_sample(t, out);
}
/* lib.rs:493-510 g1_limbs */
void orc_g1_limbs(const uint8_t *p, u64 *out) {
    int allz = 1;
    for (int i = 0; i < 64; i++)
        if (p[i]) allz = 0;
    uint8_t b[32];
    if (allz) {
        memset(b, 0, 32);
        b[16] = 1; /* 2^128, constants::G1_INFINITY_SENTINEL */
        for (int k = 0; k < 4; k++) f_from_bytes(FR, b, out + 4 * k);
        return;
    }
    for (int k = 0; k < 4; k++) {
        memset(b, 0, 32);
        memcpy(b, p + 16 * k, 16);
        f_from_bytes(FR, b, out + 4 * k);
    }
}
/* lib.rs:515-521 append_g1_points: one message of 4k limbs */
void orc_tr_append_g1(orc_transcript *t, u64 tag, const uint8_t *pts, size_t k) {
    u64 *l = (u64 *)malloc(k * 128 + 32);
    for (size_t i = 0; i < k; i++) orc_g1_limbs(pts + 64 * i, l + 16 * i);
    orc_tr_append_scalars(t, tag, l, 4 * k);
    free(l);
}

/* ------------------------------------------------------------------------
 * sumcheck: crates/sumcheck/src/lib.rs, crates/gkr/src/lib.rs:703-899
 * ---------------------------------------------------------------------- */
enum {
    TAG_SUMCHECK_ROUND = 4,
    TAG_SUMCHECK_CHALLENGE = 5,
    TAG_WITNESS_DIGEST = 8,
    TAG_SUMCHECK_FINAL_EVALS = 9
};

static void fr_inv_u64(u64 k, u64 *out) {
    u64 a[4];
    f_from_u64(FR, k, a);
    f_inv(FR, a, out);
}

/* sumcheck/src/lib.rs:294-305 interpolate_cubic (d == 3); for d != 3 the
 * Newton forward-difference form on nodes 0..d expanded into monomials, the
 * same construction carried one degree further (the degree-4 extension). */
void orc_interpolate(const u64 *v, size_t d, u64 *c) {
    if (d == 3) {
        u64 inv2[4], inv3[4], inv6[4], d1[4], d2[4], d3[4], t[4], u[4];
        fr_inv_u64(2, inv2);
        fr_inv_u64(3, inv3);
        fr_inv_u64(6, inv6);
        f_sub(FR, v + 4, v, d1);
        f_sub(FR, v + 8, v + 4, t);
        f_sub(FR, t, d1, d2);
        f_sub(FR, v + 12, v + 8, u);
        f_sub(FR, u, t, u);
        f_sub(FR, u, d2, d3);
        memcpy(c, v, 32);
        u64 a[4], b[4];
        f_mul(FR, d2, inv2, a);
        f_mul(FR, d3, inv3, b);
        f_sub(FR, d1, a, t);
        f_add(FR, t, b, c + 4);
        f_sub(FR, d2, d3, t);
        f_mul(FR, t, inv2, c + 8);
        f_mul(FR, d3, inv6, c + 12);
        return;
    }
    /* forward differences delta[k] = Delta^k v(0) */
    u64 diff[8][4], delta[8][4];
    for (size_t i = 0; i <= d; i++) memcpy(diff[i], v + 4 * i, 32);
    for (size_t k = 0; k <= d; k++) {
        memcpy(delta[k], diff[0], 32);
        for (size_t i = 0; i + k < d; i++) f_sub(FR, diff[i + 1], diff[i], diff[i]);
    }
    /* basis_k(X) = X(X-1)...(X-k+1)/k!, accumulated in monomial form */
    u64 basis[8][4], acc[8][4];
    memset(basis, 0, sizeof(basis));
    memset(acc, 0, sizeof(acc));
    memcpy(basis[0], FR_P.r, 32);
    memcpy(acc[0], delta[0], 32);
    for (size_t k = 1; k <= d; k++) {
        /* basis <- basis * (X - (k-1)) / k */
        u64 km1[4], invk[4], nb[8][4];
        f_from_u64(FR, k - 1, km1);
        fr_inv_u64(k, invk);
        memset(nb, 0, sizeof(nb));
        for (size_t i = 0; i < k; i++) {
            u64 t[4];
            f_add(FR, nb[i + 1], basis[i], nb[i + 1]);
            f_mul(FR, basis[i], km1, t);
            f_sub(FR, nb[i], t, nb[i]);
        }
        for (size_t i = 0; i <= k; i++) f_mul(FR, nb[i], invk, basis[i]);
        for (size_t i = 0; i <= k; i++) {
            u64 t[4];
            f_mul(FR, basis[i], delta[k], t);
            f_add(FR, acc[i], t, acc[i]);
        }
    }
    for (size_t i = 0; i <= d; i++) memcpy(c + 4 * i, acc[i], 32);
}

/* Gate::evaluate, sumcheck/src/lib.rs:151-169 (a third factor extends it) */
static void gate_eval(const orc_term *terms, size_t nt, const u64 *x, u64 *out) {
    u64 acc[4] = {0, 0, 0, 0};
    for (size_t t = 0; t < nt; t++) {
        u64 v[4];
        f_mul(FR, terms[t].coef, x + 4 * terms[t].f[0], v);
        for (int q = 1; q < 3; q++)
            if (terms[t].f[q] >= 0) f_mul(FR, v, x + 4 * terms[t].f[q], v);
        f_add(FR, acc, v, acc);
    }
    memcpy(out, acc, 32);
}
static size_t gate_degree(const orc_term *terms, size_t nt) {
    size_t mx = 1;
    for (size_t t = 0; t < nt; t++) {
        size_t f = 1 + (terms[t].f[1] >= 0) + (terms[t].f[2] >= 0);
        if (f > mx) mx = f;
    }
    /* times eq; rounds are never below a cubic (sumcheck/src/lib.rs:179-186) */
    return mx + 1 < 3 ? 3 : mx + 1;
}

typedef struct {
    const orc_term *terms;
    size_t nt, k, d, lo, hi;
    u64 **tables;
    const u64 *eq;
    u64 acc[8][4];
} pair_job;

/* gkr/src/lib.rs:833-887 (and sumcheck round_evaluations, lib.rs:324-349):
 * per pair, walk every line X = 0..d by adding its step */
static void *pair_worker(void *arg) {
    pair_job *j = (pair_job *)arg;
    memset(j->acc, 0, sizeof(j->acc));
    u64 *cur = (u64 *)malloc(j->k * 32 + 32), *step = (u64 *)malloc(j->k * 32 + 32);
    for (size_t i = j->lo; i < j->hi; i++) {
        for (size_t c = 0; c < j->k; c++) {
            memcpy(cur + 4 * c, j->tables[c] + 8 * i, 32);
            f_sub(FR, j->tables[c] + 8 * i + 4, cur + 4 * c, step + 4 * c);
        }
        u64 eq_at[4], eq_step[4];
        memcpy(eq_at, j->eq + 8 * i, 32);
        f_sub(FR, j->eq + 8 * i + 4, eq_at, eq_step);
        for (size_t node = 0; node <= j->d; node++) {
            u64 g[4];
            gate_eval(j->terms, j->nt, cur, g);
            f_mul(FR, eq_at, g, g);
            f_add(FR, j->acc[node], g, j->acc[node]);
            f_add(FR, eq_at, eq_step, eq_at);
            for (size_t c = 0; c < j->k; c++) f_add(FR, cur + 4 * c, step + 4 * c, cur + 4 * c);
        }
    }
    free(cur);
    free(step);
    return NULL;
}

static void round_values(const orc_term *terms, size_t nt, u64 **tables, size_t k,
                         const u64 *eq, size_t half, size_t d, size_t workers, u64 *vals) {
    if (workers < 1) workers = 1;
    if (workers > half) workers = half ? half : 1;
    pair_job *jobs = (pair_job *)calloc(workers, sizeof(pair_job));
    pthread_t *tids = (pthread_t *)malloc(workers * sizeof(pthread_t));
    size_t per = (half + workers - 1) / workers;
    for (size_t w = 0; w < workers; w++) {
        jobs[w].terms = terms;
        jobs[w].nt = nt;
        jobs[w].k = k;
        jobs[w].d = d;
        jobs[w].tables = tables;
        jobs[w].eq = eq;
        jobs[w].lo = w * per < half ? w * per : half;
        jobs[w].hi = (w + 1) * per < half ? (w + 1) * per : half;
        if (workers == 1)
            pair_worker(&jobs[w]);
        else
            pthread_create(&tids[w], NULL, pair_worker, &jobs[w]);
    }
    if (workers > 1)
        for (size_t w = 0; w < workers; w++) pthread_join(tids[w], NULL);
    memset(vals, 0, 32 * (d + 1));
    for (size_t w = 0; w < workers; w++)
        for (size_t node = 0; node <= d; node++)
            f_add(FR, vals + 4 * node, jobs[w].acc[node], vals + 4 * node);
    free(jobs);
    free(tids);
}

/* gkr/src/lib.rs:792-899 prove_sumcheck, with a gate as the summand */
void orc_prove_sumcheck(const u64 *eq_point, size_t n, const orc_term *terms, size_t nt,
                        u64 **tables, size_t k, orc_transcript *t, size_t workers,
                        u64 *rounds_out, u64 *point_out) {
    size_t d = gate_degree(terms, nt);
    size_t len = (size_t)1 << n;
    u64 *eq = (u64 *)malloc(len * 32);
    orc_eq_table(eq_point, n, eq);
    u64 vals[8 * 4];
    for (size_t round = 0; round < n; round++) {
        size_t half = len / 2;
        round_values(terms, nt, tables, k, eq, half, d, workers, vals);
        u64 *g = rounds_out + 4 * (d + 1) * round;
        orc_interpolate(vals, d, g);
        orc_tr_append_scalars(t, TAG_SUMCHECK_ROUND, g, d + 1);
        u64 r[4];
        orc_tr_challenge(t, TAG_SUMCHECK_CHALLENGE, r);
        for (size_t c = 0; c < k; c++) orc_bind(tables[c], len, r);
        orc_bind(eq, len, r);
        len = half;
        memcpy(point_out + 4 * round, r, 32);
    }
    free(eq);
}

/* sumcheck/src/lib.rs:359-408 prove_zerocheck */
void orc_prove_zerocheck(const orc_term *terms, size_t nt, u64 **tables, size_t k, size_t n,
                         orc_transcript *t, u64 *rounds_out, u64 *final_evals_out) {
    u64 *r = (u64 *)malloc(n * 32 + 32);
    for (size_t i = 0; i < n; i++) orc_tr_challenge(t, TAG_SUMCHECK_CHALLENGE, r + 4 * i);
    u64 *point = (u64 *)malloc(n * 32 + 32);
    orc_prove_sumcheck(r, n, terms, nt, tables, k, t, 1, rounds_out, point);
    for (size_t c = 0; c < k; c++) memcpy(final_evals_out + 4 * c, tables[c], 32);
    orc_tr_append_scalars(t, TAG_SUMCHECK_FINAL_EVALS, final_evals_out, k);
    free(r);
    free(point);
}

/* The random-linear-combination fold of k same-size columns:
 * out_o[i] = sum_c w[o k + c] * col_c[i] (pcs::batch_open_stacked builds its
 * combined column this way, crates/pcs/src/lib.rs:768-780, with weights
 * rho^i from batch_preamble, crates/pcs-verify/src/lib.rs:333-345). */
void orc_combine(const u64 *cols, size_t n, size_t k, const u64 *w, size_t outs, u64 *out) {
    for (size_t o = 0; o < outs; o++)
        for (size_t i = 0; i < n; i++) {
            u64 acc[4] = {0, 0, 0, 0}, t[4];
            for (size_t c = 0; c < k; c++) {
                f_mul(FR, w + 4 * (o * k + c), cols + 4 * (c * n + i), t);
                f_add(FR, acc, t, acc);
            }
            memcpy(out + 4 * (o * n + i), acc, 32);
        }
}

/* y[i] += a * x[i] over n elements of Fr: one column's share of a
 * random-linear-combination fold (orc_combine), one column at a time. */
void orc_axpy(size_t n, const u64 *a, const u64 *x, u64 *y) {
    for (size_t i = 0; i < n; i++) {
        u64 t[4];
        f_mul(FR, a, x + 4 * i, t);
        f_add(FR, y + 4 * i, t, y + 4 * i);
    }
}

/* The claim of a product sumcheck, sum_x eq(r, x) prod_{o<f} t_o(x) over the
 * 2^n hypercube: what g_0(0) + g_0(1) must equal (the verifier's first check,
 * crates/sumcheck/src/lib.rs:441), with eq by eq_table (lib.rs:233-248). */
void orc_product_claim(const u64 *r, size_t n, const u64 *const *tables, size_t f, u64 *out) {
    const size_t len = (size_t)1 << n;
    u64 *eq = (u64 *)malloc(len * 32);
    orc_eq_table(r, n, eq);
    u64 acc[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < len; i++) {
        u64 p[4];
        memcpy(p, eq + 4 * i, 32);
        for (size_t o = 0; o < f; o++) f_mul(FR, p, tables[o] + 4 * i, p);
        f_add(FR, acc, p, acc);
    }
    memcpy(out, acc, 32);
    free(eq);
}

/* sumcheck/src/lib.rs:228-255 witness_digest */
void orc_witness_digest(u64 **tables, size_t k, size_t n, u64 *out) {
    orc_transcript s;
    orc_tr_init(&s);
    u64 hdr[8];
    f_from_u64(FR, k, hdr);
    f_from_u64(FR, n, hdr + 4);
    orc_tr_append_scalars(&s, TAG_WITNESS_DIGEST, hdr, 2);
    for (size_t c = 0; c < k; c++)
        orc_tr_append_scalars(&s, TAG_WITNESS_DIGEST, tables[c], (size_t)1 << n);
    orc_tr_sample(&s, out);
}

/* ------------------------------------------------------------------------
 * Seeded inputs.
 * splitmix64: tools/test-support/src/lib.rs:26-66; next_fr draw-mask-reject:
 * crates/pcs/tests/common/mod.rs:114-123.
 * ChaCha20 block function: RFC 8439 section 2.3.
 * ---------------------------------------------------------------------- */
u64 orc_splitmix_next(u64 *s) {
    *s += 0x9e3779b97f4a7c15ULL;
    u64 z = *s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
void orc_splitmix_next_fr(u64 *s, u64 *out) {
    for (;;) {
        uint8_t b[32];
        for (int i = 0; i < 4; i++) {
            u64 v = orc_splitmix_next(s);
            for (int k = 0; k < 8; k++) b[8 * i + k] = (uint8_t)(v >> (8 * k));
        }
        b[31] &= 0x3f;
        if (f_from_bytes(FR, b, out)) return;
    }
}

#define ROTL(a, b) (((a) << (b)) | ((a) >> (32 - (b))))
#define QR(a, b, c, d)                                                                           \
    (a += b, d ^= a, d = ROTL(d, 16), c += d, b ^= c, b = ROTL(b, 12), a += b, d ^= a,          \
     d = ROTL(d, 8), c += d, b ^= c, b = ROTL(b, 7))
void orc_chacha_block(const uint32_t key[8], uint32_t counter, const uint32_t nonce[3],
                      uint8_t out[64]) {
    uint32_t s[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};
    for (int i = 0; i < 8; i++) s[4 + i] = key[i];
    s[12] = counter;
    s[13] = nonce[0];
    s[14] = nonce[1];
    s[15] = nonce[2];
    uint32_t x[16];
    memcpy(x, s, 64);
    for (int i = 0; i < 10; i++) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) {
        uint32_t v = x[i] + s[i];
        out[4 * i] = (uint8_t)v;
        out[4 * i + 1] = (uint8_t)(v >> 8);
        out[4 * i + 2] = (uint8_t)(v >> 16);
        out[4 * i + 3] = (uint8_t)(v >> 24);
    }
}
/* key = seed as little-endian u64 in key words 0..1, zero elsewhere; nonce 0;
 * counter from 0; bytes consumed in stream order. */
void orc_chacha_init(orc_chacha *c, u64 seed) {
    memset(c, 0, sizeof(*c));
    c->key[0] = (uint32_t)seed;
    c->key[1] = (uint32_t)(seed >> 32);
    c->pos = 64;
}
static uint8_t chacha_byte(orc_chacha *c) {
    if (c->pos == 64) {
        uint32_t nonce[3] = {0, 0, 0};
        orc_chacha_block(c->key, c->counter++, nonce, c->buf);
        c->pos = 0;
    }
    return c->buf[c->pos++];
}
uint32_t orc_chacha_u32(orc_chacha *c) {
    uint32_t v = 0;
    for (int k = 0; k < 4; k++) v |= (uint32_t)chacha_byte(c) << (8 * k);
    return v;
}
void orc_chacha_fr(orc_chacha *c, u64 *out) {
    for (;;) {
        uint8_t b[32];
        for (int i = 0; i < 32; i++) b[i] = chacha_byte(c);
        b[31] &= 0x3f;
        if (f_from_bytes(FR, b, out)) return;
    }
}
void orc_chacha_witness(orc_chacha *c, u64 *out) {
    uint32_t cat = orc_chacha_u32(c) % 10;
    if (cat < 2) {
        memset(out, 0, 32);
    } else if (cat == 2) {
        memcpy(out, FR_P.r, 32);
    } else if (cat < 8) {
        f_from_u64(FR, orc_chacha_u32(c), out);
    } else {
        orc_chacha_fr(c, out);
    }
}
/* Counter-mode draws: element i from ChaCha20 blocks with counter i and nonce
 * (attempt, domain, 0); domain "fr" = 0x7266 for uniform Fr, "wi" = 0x6977
 * for the witness mix. A block offers 32-byte Fr candidates at bytes 0-31
 * and 32-63 (top two bits cleared, reject >= p); the witness category is the
 * u32 at bytes 0-3 of attempt 0 mod 10, its u32 value bytes 4-7, its Fr
 * candidates bytes 32-63 of attempt 0 then both halves of attempts 1, 2... */
static int ctr_candidate(const uint8_t *b32, u64 *out) {
    uint8_t b[32];
    memcpy(b, b32, 32);
    b[31] &= 0x3f;
    return f_from_bytes(FR, b, out);
}
static void ctr_block(u64 seed, uint32_t i, uint32_t attempt, uint32_t dom, uint8_t out[64]) {
    uint32_t key[8] = {(uint32_t)seed, (uint32_t)(seed >> 32), 0, 0, 0, 0, 0, 0};
    uint32_t nonce[3] = {attempt, dom, 0};
    orc_chacha_block(key, i, nonce, out);
}
static void ctr_fr(u64 seed, uint32_t i, uint32_t dom, uint32_t first, u64 *out) {
    for (uint32_t a = first;; a++) {
        uint8_t blk[64];
        ctr_block(seed, i, a, dom, blk);
        if (ctr_candidate(blk, out) || ctr_candidate(blk + 32, out)) return;
    }
}
void orc_fill_ctr(u64 seed, int kind, u64 *out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        u64 *o = out + 4 * i;
        if (kind == 0) {
            ctr_fr(seed, (uint32_t)i, 0x7266, 0, o);
            continue;
        }
        uint8_t blk[64];
        ctr_block(seed, (uint32_t)i, 0, 0x6977, blk);
        uint32_t cat = (uint32_t)blk[0] | ((uint32_t)blk[1] << 8) | ((uint32_t)blk[2] << 16) |
                       ((uint32_t)blk[3] << 24);
        uint32_t v = (uint32_t)blk[4] | ((uint32_t)blk[5] << 8) | ((uint32_t)blk[6] << 16) |
                     ((uint32_t)blk[7] << 24);
        cat %= 10;
        if (cat < 2) {
            memset(o, 0, 32);
        } else if (cat == 2) {
            memcpy(o, FR_P.r, 32);
        } else if (cat < 8) {
            f_from_u64(FR, v, o);
        } else if (!ctr_candidate(blk + 32, o)) {
            ctr_fr(seed, (uint32_t)i, 0x6977, 1, o);
        }
    }
}

void orc_fill_uniform(u64 seed, u64 *out, size_t n) {
    orc_chacha c;
    orc_chacha_init(&c, seed);
    for (size_t i = 0; i < n; i++) orc_chacha_fr(&c, out + 4 * i);
}
void orc_fill_witness(u64 seed, u64 *out, size_t n) {
    orc_chacha c;
    orc_chacha_init(&c, seed);
    for (size_t i = 0; i < n; i++) orc_chacha_witness(&c, out + 4 * i);
}
