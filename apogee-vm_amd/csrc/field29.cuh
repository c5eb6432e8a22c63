// Fq in nine 29-bit words with lazy reduction: the form the MSM's bucket
// phase computes in.
//
// A value is nine words v[j] < 2^29 (bits [29j, 29j + 29)) of an integer x
// that stands for x mod p in the Montgomery domain of R' = 2^261 (x = a R'
// mod p, up to a multiple of p): nine reduction rounds of one 29-bit word
// each. The multiply accumulates in 64-bit words with no carry handling in
// its rounds: a column takes at most 9 products a_j b_i and 9 products
// m_i p_j, each below 2^58, so it stays below 2^63. Its result is
// (a b + M p) / 2^261 < a b / R' + p, and R' / p is about 169, so inputs up to
// ~12p give a result below 1.1 p: subtraction is a - b + K p with K p above
// b's bound and no conditional correction, addition is a plain add, and only
// the final conversion reduces fully. Each formula in curve29.cuh states the
// bounds of its values. On gfx950 a multiply is ~230 VALU instructions against
// ~430 for the 32-bit CIOS of field.cuh, whose add-with-carry chains cost a
// carry select and a hazard wait per word.
//
// Entering the domain multiplies a 32-bit Montgomery value (R = 2^256) by 2^5
// (five modular doublings, done where the MSM's base tables are built);
// leaving it multiplies by 2^256 mod p in this domain, which divides by 2^5.
#pragma once
#include "field.cuh"

namespace apg {

struct F29 {
    uint32_t v[9];
};

constexpr uint32_t W29 = (1u << 29) - 1;

// word j of an 8 x 32-bit little-endian integer, re-sliced to 29 bits
template <int J>
APG_HD uint32_t slice29(const uint32_t l[8]) {
    constexpr int bit = 29 * J, w = bit >> 5, s = bit & 31;
    uint64_t x = (uint64_t)l[w] >> s;
    if (w + 1 < 8) x |= (uint64_t)l[w + 1] << (32 - s);
    return (uint32_t)x & W29;
}

APG_HD uint32_t slice29_rt(const uint32_t l[8], int j) {
    const int bit = 29 * j, w = bit >> 5, s = bit & 31;
    uint64_t x = (uint64_t)l[w] >> s;
    if (w + 1 < 8) x |= (uint64_t)l[w + 1] << (32 - s);
    return (uint32_t)x & W29;
}

// words of K p (K p < 2^261)
template <class P, uint32_t K = 1>
APG_HD uint32_t modk29(int j) {
    uint32_t l[9];
    uint64_t c = 0;
    for (int i = 0; i < 8; i++) {
        const uint64_t s = (uint64_t)P::mod(i) * K + c;
        l[i] = (uint32_t)s;
        c = s >> 32;
    }
    l[8] = (uint32_t)c;
    const int bit = 29 * j, w = bit >> 5, s = bit & 31;
    uint64_t x = (uint64_t)l[w] >> s;
    if (w + 1 < 9) x |= (uint64_t)l[w + 1] << (32 - s);
    return (uint32_t)x & W29;
}

template <class P>
APG_HD uint32_t mod29(int j) {
    return modk29<P, 1>(j);
}

// 2^256 mod p (P::one) in words: the multiplier that leaves the domain
template <class P>
APG_HD uint32_t r256_29(int j) {
    uint32_t l[8];
    for (int i = 0; i < 8; i++) l[i] = P::one(i);
    return slice29_rt(l, j);
}

// 2^261 mod p in words: one in the domain (32 x 2^256 mod p, five doublings)
template <class P>
APG_HD uint32_t one29w(int j) {
    uint32_t l[8];
    for (int i = 0; i < 8; i++) l[i] = P::one(i);
    for (int d = 0; d < 5; d++) {
        uint32_t t[8], c = 0;
        for (int i = 0; i < 8; i++) {
            t[i] = (l[i] << 1) | c;
            c = l[i] >> 31;
        }
        // t >= p (t < 2p < 2^256): subtract
        bool ge = true;
        for (int i = 7; i >= 0; i--) {
            if (t[i] != P::mod(i)) {
                ge = t[i] > P::mod(i);
                break;
            }
        }
        uint32_t br = 0;
        for (int i = 0; i < 8; i++) {
            const uint64_t s = (uint64_t)t[i] - P::mod(i) - br;
            l[i] = ge ? (uint32_t)s : t[i];
            br = (uint32_t)(s >> 63);
        }
    }
    return slice29_rt(l, j);
}

template <class P>
APG_HD F29 f29_const_one() {
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = one29w<P>(j);
    return r;
}

APG_HD F29 f29_from32(const Fe &a) {
    F29 r;
    r.v[0] = slice29<0>(a.l);
    r.v[1] = slice29<1>(a.l);
    r.v[2] = slice29<2>(a.l);
    r.v[3] = slice29<3>(a.l);
    r.v[4] = slice29<4>(a.l);
    r.v[5] = slice29<5>(a.l);
    r.v[6] = slice29<6>(a.l);
    r.v[7] = slice29<7>(a.l);
    r.v[8] = slice29<8>(a.l);
    return r;
}

// a 32-bit Montgomery value (R = 2^256, below 2p) times 2^5 mod p, still
// in 32-bit words: the domain's representative, as the base tables hold it
template <class P>
APG_HD Fe f_to_r261(const Fe &a) {
    Fe r = f_add<P>(a, f_zero<P>());  // a < 2p -> a mod p
#pragma unroll
    for (int d = 0; d < 5; d++) r = f_add<P>(r, r);
    return r;
}

// d = a - b - borrow_in over the nine words; returns the borrow out.
APG_HD uint32_t sub29_words(const uint32_t a[9], const uint32_t b[9], uint32_t d[9]) {
    uint32_t br = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const int32_t x = (int32_t)a[j] - (int32_t)b[j] - (int32_t)br;
        br = x < 0 ? 1u : 0u;
        d[j] = (uint32_t)x & W29;
    }
    return br;
}

template <class P>
APG_HD F29 f29_mul(const F29 &a, const F29 &b) {
    uint64_t t[10];
#pragma unroll
    for (int j = 0; j < 10; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] += (uint64_t)a.v[j] * b.v[i];
        const uint32_t m = ((uint32_t)t[0] * P::inv) & W29;
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] += (uint64_t)m * mod29<P>(j);
        const uint64_t c = t[0] >> 29;
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] = t[j + 1];
        t[0] += c;
    }
    // normalise to 29-bit words (the value is below 2^255)
    F29 r;
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint64_t s = t[j] + c;
        r.v[j] = (uint32_t)s & W29;
        c = s >> 29;
    }
    return r;
}

// a^2 with the products of distinct words taken once and doubled: row i adds
// a_i^2 at word i and (2 a_j) a_i at words j > i, so the pair (i, j) lands in
// column i + j exactly once (45 multiply-adds for the product instead of 81).
// A column then holds at most 4 doubled products below 2^59, one square
// below 2^58 and 9 products m_i p_j below 2^58: below 2^63. Bounds as f29_mul.
template <class P>
APG_HD F29 f29_sqr(const F29 &a) {
    uint64_t t[10];
    uint32_t a2[9];
#pragma unroll
    for (int j = 0; j < 10; j++) t[j] = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) a2[j] = a.v[j] << 1;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        t[i] += (uint64_t)a.v[i] * a.v[i];
#pragma unroll
        for (int j = i + 1; j < 9; j++) t[j] += (uint64_t)a2[j] * a.v[i];
        const uint32_t m = ((uint32_t)t[0] * P::inv) & W29;
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] += (uint64_t)m * mod29<P>(j);
        const uint64_t c = t[0] >> 29;
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] = t[j + 1];
        t[0] += c;
    }
    F29 r;
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint64_t s = t[j] + c;
        r.v[j] = (uint32_t)s & W29;
        c = s >> 29;
    }
    return r;
}

// a b + c d with one reduction: (a b + c d + M p) / 2^261, below
// (a b + c d) / R' + p. A column holds at most 27 products below 2^58:
// below 2^63. A difference a b - c e is a b + c (K p - e).
template <class P>
APG_HD F29 f29_mul2(const F29 &a, const F29 &b, const F29 &c, const F29 &d) {
    uint64_t t[10];
#pragma unroll
    for (int j = 0; j < 10; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] += (uint64_t)a.v[j] * b.v[i] + (uint64_t)c.v[j] * d.v[i];
        const uint32_t m = ((uint32_t)t[0] * P::inv) & W29;
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] += (uint64_t)m * mod29<P>(j);
        const uint64_t cy = t[0] >> 29;
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] = t[j + 1];
        t[0] += cy;
    }
    F29 r;
    uint64_t cy = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint64_t s = t[j] + cy;
        r.v[j] = (uint32_t)s & W29;
        cy = s >> 29;
    }
    return r;
}

// sum_g a_g b_g over G <= 6 pairs with one reduction: (sum a_g b_g + M p) /
// 2^261, below sum a_g b_g / R' + p. A column holds at most 9 (G + 1)
// products below 2^58: 63 of them at G = 6, below 2^64.
template <class P, int G>
APG_HD F29 f29_dot(const F29 *a, const F29 *b) {
    static_assert(G >= 1 && G <= 6, "at most six products per reduction");
    uint64_t t[10];
#pragma unroll
    for (int j = 0; j < 10; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int j = 0; j < 9; j++) t[j] += (uint64_t)a[g].v[j] * b[g].v[i];
        const uint32_t m = ((uint32_t)t[0] * P::inv) & W29;
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] += (uint64_t)m * mod29<P>(j);
        const uint64_t cy = t[0] >> 29;
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] = t[j + 1];
        t[0] += cy;
    }
    F29 r;
    uint64_t cy = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint64_t s = t[j] + cy;
        r.v[j] = (uint32_t)s & W29;
        cy = s >> 29;
    }
    return r;
}

// a + b, not reduced (the bound is the sum of the bounds)
APG_HD F29 f29_add(const F29 &a, const F29 &b) {
    F29 r;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint32_t x = a.v[j] + b.v[j] + c;
        r.v[j] = x & W29;
        c = x >> 29;
    }
    return r;
}

// a + b (a + b + c) word by word with no carries: words below 2^30 (3 x
// 2^29), bound the sum of the bounds. Only for the subtrahend of f29_sub,
// whose signed word arithmetic takes such words (a multiply's operands need
// 29-bit words).
APG_HD F29 f29_add_words(const F29 &a, const F29 &b) {
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = a.v[j] + b.v[j];
    return r;
}

APG_HD F29 f29_sum3_words(const F29 &a, const F29 &b, const F29 &c) {
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = a.v[j] + b.v[j] + c.v[j];
    return r;
}

// a - b + K p for b < K p (the bound is a's plus K p); b's words may be
// below 3 x 2^29 (f29_add_words, f29_sum3_words), a's below 2^29
template <class P, uint32_t K>
APG_HD F29 f29_sub(const F29 &a, const F29 &b) {
    F29 r;
    int32_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const int32_t x = (int32_t)(a.v[j] + modk29<P, K>(j)) - (int32_t)b.v[j] + c;
        r.v[j] = (uint32_t)x & W29;
        c = x >> 29;  // arithmetic: the borrow is -1
    }
    return r;
}

// p - a for a <= p (the negation of a fully reduced coordinate)
template <class P>
APG_HD F29 f29_neg_reduced(const F29 &a) {
    uint32_t pw[9], d[9];
#pragma unroll
    for (int j = 0; j < 9; j++) pw[j] = mod29<P>(j);
    (void)sub29_words(pw, a.v, d);
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = d[j];
    return r;
}

// x = k p for some 0 <= k <= MAXK, i.e. x = 0 mod p for x below (MAXK + 1) p.
// The low words of k p are distinct (p is odd), so one word picks the only
// candidate and the full comparison runs only when it matches.
template <class P, int MAXK>
APG_HD bool f29_is_zero(const F29 &x) {
    const uint32_t p0 = mod29<P>(0);
    int cand = -1;
#pragma unroll
    for (int k = 0; k <= MAXK; k++)
        if (x.v[0] == (((uint32_t)k * p0) & W29)) cand = k;
    if (cand < 0) return false;
    uint64_t c = 0;
    bool eq = true;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint64_t s = (uint64_t)(uint32_t)cand * mod29<P>(j) + c;
        eq = eq && x.v[j] == ((uint32_t)s & W29);
        c = s >> 29;
    }
    return eq;
}

// leaves the domain: x / 2^5 as a fully reduced 32-bit Montgomery value
// (R = 2^256); x may be any value below ~12p
template <class P>
APG_HD Fe f29_to32(const F29 &x) {
    F29 c;
#pragma unroll
    for (int j = 0; j < 9; j++) c.v[j] = r256_29<P>(j);
    const F29 m = f29_mul<P>(x, c);  // below 1.1 p
    uint32_t p[9], d[9];
#pragma unroll
    for (int j = 0; j < 9; j++) p[j] = mod29<P>(j);
    const uint32_t br = sub29_words(m.v, p, d);
    uint32_t w[9];
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = br ? m.v[j] : d[j];
    Fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int bit = 32 * i, j = bit / 29, sh = bit % 29;
        uint64_t v = (uint64_t)w[j] >> sh;
        v |= (uint64_t)w[j + 1] << (29 - sh);
        if (j + 2 < 9) v |= (uint64_t)w[j + 2] << (58 - sh);
        r.l[i] = (uint32_t)v;
    }
    return r;
}

// x mod p for x < 32p, packed into eight 32-bit words: subtracts 16p, 8p,
// 4p, 2p and p where they fit (the words of x are below 2^29)
template <class P, uint32_t K>
APG_HD void f29_sub_if_ge(uint32_t w[9]) {
    uint32_t k[9], d[9];
#pragma unroll
    for (int j = 0; j < 9; j++) k[j] = modk29<P, K>(j);
    const uint32_t br = sub29_words(w, k, d);
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = br ? w[j] : d[j];
}

template <class P>
APG_HD Fe f29_pack_reduced(const F29 &x) {
    uint32_t w[9];
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = x.v[j];
    f29_sub_if_ge<P, 16>(w);
    f29_sub_if_ge<P, 8>(w);
    f29_sub_if_ge<P, 4>(w);
    f29_sub_if_ge<P, 2>(w);
    f29_sub_if_ge<P, 1>(w);
    Fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int bit = 32 * i, j = bit / 29, sh = bit % 29;
        uint64_t v = (uint64_t)w[j] >> sh;
        v |= (uint64_t)w[j + 1] << (29 - sh);
        if (j + 2 < 9) v |= (uint64_t)w[j + 2] << (58 - sh);
        r.l[i] = (uint32_t)v;
    }
    return r;
}

// x mod p for x < 2p, packed into eight 32-bit words
template <class P>
APG_HD Fe f29_pack_small(const F29 &x) {
    uint32_t w[9];
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = x.v[j];
    f29_sub_if_ge<P, 1>(w);
    Fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int bit = 32 * i, j = bit / 29, sh = bit % 29;
        uint64_t v = (uint64_t)w[j] >> sh;
        v |= (uint64_t)w[j + 1] << (29 - sh);
        if (j + 2 < 9) v |= (uint64_t)w[j + 2] << (58 - sh);
        r.l[i] = (uint32_t)v;
    }
    return r;
}

// a b / 2^261 mod p of two fully reduced 32-bit words values, fully reduced:
// for 32-bit Montgomery images a = x R, b = y R it is the image of x y / 2^5
template <class P>
APG_HD Fe f_mul_r261(const Fe &a, const Fe &b) {
    return f29_pack_small<P>(f29_mul<P>(f29_from32(a), f29_from32(b)));
}

// enters the domain from a fully reduced 32-bit Montgomery value
template <class P>
APG_HD F29 f29_from_mont32(const Fe &a) {
    return f29_from32(f_to_r261<P>(a));
}

// A test composition over the lazy bounds: ((a + b)(a - b) + a^2) - b, in
// and out as fully reduced 32-bit Montgomery values; the host build and a
// kernel are checked against the oracle.
template <class P>
APG_HD Fe f29_test_chain(const Fe &a32, const Fe &b32, int op) {
    const F29 a = f29_from_mont32<P>(a32), b = f29_from_mont32<P>(b32);
    F29 r;
    switch (op) {
        case 0: r = f29_mul<P>(a, b); break;
        case 1: r = f29_add(a, b); break;
        case 2: r = f29_sub<P, 2>(a, b); break;
        case 4: r = f29_sqr<P>(a); break;
        case 5: r = f29_mul2<P>(a, b, a, a); break;
        default: r = f29_sub<P, 2>(f29_add(f29_mul<P>(f29_add(a, b), f29_sub<P, 2>(a, b)), f29_mul<P>(a, a)), b);
    }
    return f29_to32<P>(r);
}

}  // namespace apg
