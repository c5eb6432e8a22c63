// Fq in nine 29-bit words with lazy reduction: the form the MSM's bucket
// accumulation computes in.
//
// A value is nine words v[j] < 2^29 (bits [29j, 29j + 29)) of an integer
// x < 2p. It stands for x mod p, Montgomery with the same R = 2^256 as the
// 32-bit form (field.cuh), so converting is re-slicing bits, not a multiply.
//
// The multiply accumulates in 64-bit words with no carry handling in the
// inner loops: a column takes at most 9 products a_j b_i and 9 products
// m_i p_j, each below 2^58, so it stays below 2^63. Each of the nine rounds
// reduces by one word (2^29), the last by 2^24 (8 * 29 + 24 = 256). For inputs
// below 2p the result (a b + M p) / 2^256 is below 2p again because 4p <
// 2^256, so no conditional subtraction is needed between multiplies. On gfx950
// this is ~250 VALU instructions per multiply against ~430 for the 32-bit
// CIOS, whose add-with-carry chains cost a carry select and a hazard wait per
// word.
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

template <class P>
APG_HD uint32_t mod29(int j) {
    uint32_t l[8];
    for (int i = 0; i < 8; i++) l[i] = P::mod(i);
    const int bit = 29 * j, w = bit >> 5, s = bit & 31;
    uint64_t x = (uint64_t)l[w] >> s;
    if (w + 1 < 8) x |= (uint64_t)l[w + 1] << (32 - s);
    return (uint32_t)x & W29;
}

// words of 2p (for the lazy range [0, 2p))
template <class P>
APG_HD uint32_t mod2x29(int j) {
    uint32_t l[8];
    uint32_t c = 0;
    for (int i = 0; i < 8; i++) {
        const uint64_t s = ((uint64_t)P::mod(i) << 1) | c;
        l[i] = (uint32_t)s;
        c = P::mod(i) >> 31;
    }
    const int bit = 29 * j, w = bit >> 5, s = bit & 31;
    uint64_t x = (uint64_t)l[w] >> s;
    if (w + 1 < 8) x |= (uint64_t)l[w + 1] << (32 - s);
    return (uint32_t)x & W29;
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

// x mod p as a fully reduced 32-bit-word value (x < 2p)
template <class P>
APG_HD Fe f29_to32(const F29 &a) {
    uint32_t p[9], d[9];
#pragma unroll
    for (int j = 0; j < 9; j++) p[j] = mod29<P>(j);
    const uint32_t br = sub29_words(a.v, p, d);
    const uint32_t *s = br ? a.v : d;  // a < p: keep a
    uint32_t w[9];
#pragma unroll
    for (int j = 0; j < 9; j++) w[j] = s[j];
    Fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        // 32-bit word i: bits [32i, 32i + 32)
        const int bit = 32 * i, j = bit / 29, sh = bit % 29;
        uint64_t x = (uint64_t)w[j] >> sh;
        x |= (uint64_t)w[j + 1] << (29 - sh);
        if (j + 2 < 9) x |= (uint64_t)w[j + 2] << (58 - sh);
        r.l[i] = (uint32_t)x;
    }
    return r;
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
        const uint32_t mask = i < 8 ? W29 : (1u << 24) - 1;
        uint32_t m = ((uint32_t)t[0] * P::inv) & mask;
#if defined(__HIP_DEVICE_COMPILE__)
        // Hide m's known width from the gfx950 backend. Told that the last
        // round's m fits 24 bits, it selects 24-bit multiplies for the words
        // of p below 2^24 (words 6 and 8 of Fq), drops the mask they make
        // redundant, then fuses them into v_mad_u64_u32, which reads all 32
        // bits of the unmasked value: wrong products. The empty asm keeps the
        // masked value in a register of unknown width.
        asm volatile("" : "+v"(m));
#endif
#pragma unroll
        for (int j = 0; j < 9; j++) t[j] += (uint64_t)m * mod29<P>(j);
        if (i < 8) {
            const uint64_t c = t[0] >> 29;
#pragma unroll
            for (int j = 0; j < 9; j++) t[j] = t[j + 1];
            t[0] += c;
        }
    }
    // t[0] = 0 mod 2^24: normalise to 29-bit words, then shift right by 24
    uint32_t n[10];
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint64_t s = t[j] + c;
        n[j] = (uint32_t)s & W29;
        c = s >> 29;
    }
    n[9] = (uint32_t)c;
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = ((n[j] >> 24) | (n[j + 1] << 5)) & W29;
    return r;
}

// (a + b) with a, b < 2p, reduced below 2p
template <class P>
APG_HD F29 f29_add(const F29 &a, const F29 &b) {
    uint32_t s[9], tp[9], d[9];
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint32_t x = a.v[j] + b.v[j] + c;
        s[j] = x & W29;
        c = x >> 29;
    }
#pragma unroll
    for (int j = 0; j < 9; j++) tp[j] = mod2x29<P>(j);
    const uint32_t br = sub29_words(s, tp, d);  // s < 4p < 2^256: no tenth word
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = br ? s[j] : d[j];
    return r;
}

// (a - b) with a, b < 2p, reduced into [0, 2p)
template <class P>
APG_HD F29 f29_sub(const F29 &a, const F29 &b) {
    uint32_t d[9], e[9];
    const uint32_t br = sub29_words(a.v, b.v, d);
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        const uint32_t x = d[j] + (br ? mod2x29<P>(j) : 0u) + c;
        e[j] = x & W29;
        c = x >> 29;
    }
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = e[j];
    return r;
}

// x = 0 mod p for x < 2p: x is 0 or p
template <class P>
APG_HD bool f29_is_zero(const F29 &a) {
    uint32_t z = 0, q = 0;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        z |= a.v[j];
        q |= a.v[j] ^ mod29<P>(j);
    }
    return z == 0 || q == 0;
}

// A test composition over the whole lazy range: ((a + b)(a - b) + a^2) - b,
// every intermediate in [0, 2p); the host build checks it against the oracle.
template <class P>
APG_HD Fe f29_test_chain(const Fe &a32, const Fe &b32, int op) {
    const F29 a = f29_from32(a32), b = f29_from32(b32);
    F29 r;
    switch (op) {
        case 0: r = f29_mul<P>(a, b); break;
        case 1: r = f29_add<P>(a, b); break;
        case 2: r = f29_sub<P>(a, b); break;
        default: r = f29_sub<P>(f29_add<P>(f29_mul<P>(f29_add<P>(a, b), f29_sub<P>(a, b)), f29_mul<P>(a, a)), b);
    }
    return f29_to32<P>(r);
}

}  // namespace apg
