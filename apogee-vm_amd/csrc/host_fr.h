// Host arithmetic over Fr on four 64-bit words: the Montgomery form with
// R = 2^256, so a value's memory image is the same as Fe's eight 32-bit
// words. Used where the host does field work in a loop: the Poseidon2
// permutation of the transcript and the last sumcheck rounds.
#pragma once
#include <stdint.h>
#if defined(__x86_64__) && !defined(__HIP_DEVICE_COMPILE__)
#include <cpuid.h>
#define APG_HOST_ADX 1
#endif

#include "field.cuh"

namespace apg {
namespace hfr {

// The permutation runs on four 64-bit Montgomery words (the same R = 2^256,
// so the memory image is unchanged): a CIOS multiply unrolled for the Fr
// modulus, whose top word is below 2^62, so the running sum never needs a
// fifth carry word.
typedef unsigned __int128 u128;
struct W4 {
    uint64_t v[4];
};

constexpr uint64_t RP0 = 0x43e1f593f0000001ull, RP1 = 0x2833e84879b97091ull,
                   RP2 = 0xb85045b68181585dull, RP3 = 0x30644e72e131a029ull;
static_assert(RP0 == ((uint64_t)FrP::mod(0) | ((uint64_t)FrP::mod(1) << 32)), "Fr modulus word 0");
static_assert(RP3 == ((uint64_t)FrP::mod(6) | ((uint64_t)FrP::mod(7) << 32)), "Fr modulus word 3");

inline W4 w4_of(const Fe &a) {
    W4 r;
    for (int i = 0; i < 4; i++) r.v[i] = (uint64_t)a.l[2 * i] | ((uint64_t)a.l[2 * i + 1] << 32);
    return r;
}

inline Fe fe_of(const W4 &a) {
    Fe r;
    for (int i = 0; i < 4; i++) {
        r.l[2 * i] = (uint32_t)a.v[i];
        r.l[2 * i + 1] = (uint32_t)(a.v[i] >> 32);
    }
    return r;
}

// t - p if t >= p
inline W4 w4_reduce(uint64_t t0, uint64_t t1, uint64_t t2, uint64_t t3) {
    u128 x = (u128)t0 - RP0;
    const uint64_t d0 = (uint64_t)x;
    x = (u128)t1 - RP1 - (uint64_t)((x >> 64) & 1);
    const uint64_t d1 = (uint64_t)x;
    x = (u128)t2 - RP2 - (uint64_t)((x >> 64) & 1);
    const uint64_t d2 = (uint64_t)x;
    x = (u128)t3 - RP3 - (uint64_t)((x >> 64) & 1);
    const uint64_t d3 = (uint64_t)x;
    const bool lt = (x >> 64) & 1;
    return lt ? W4{{t0, t1, t2, t3}} : W4{{d0, d1, d2, d3}};
}

// Lazy forms for the permutation: values stay below 2p. With p < 2^254 the
// CIOS product of two values below 2p is below (4p^2 + 2^256 p) / 2^256 < 2p
// without the final subtraction, and the running sum stays below 3p < 2^256;
// a sum of two values below 2p is below 4p < 2^256 and one conditional
// subtraction of 2p brings it back below 2p. The state is made canonical
// (below p) when the permutation ends.
constexpr uint64_t R2P0 = RP0 << 1, R2P1 = (RP1 << 1) | (RP0 >> 63), R2P2 = (RP2 << 1) | (RP1 >> 63),
                   R2P3 = (RP3 << 1) | (RP2 >> 63);
static_assert((RP3 >> 62) == 0, "p < 2^254");

inline W4 w4_mul_lazy(const W4 &a, const W4 &b) {
    uint64_t t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    for (int i = 0; i < 4; i++) {
        const uint64_t bi = b.v[i];
        u128 x = (u128)a.v[0] * bi + t0;
        t0 = (uint64_t)x;
        x = (u128)a.v[1] * bi + t1 + (uint64_t)(x >> 64);
        t1 = (uint64_t)x;
        x = (u128)a.v[2] * bi + t2 + (uint64_t)(x >> 64);
        t2 = (uint64_t)x;
        x = (u128)a.v[3] * bi + t3 + (uint64_t)(x >> 64);
        t3 = (uint64_t)x;
        const uint64_t t4 = (uint64_t)(x >> 64);
        const uint64_t m = t0 * FrP::inv64;
        x = (u128)m * RP0 + t0;
        x = (u128)m * RP1 + t1 + (uint64_t)(x >> 64);
        t0 = (uint64_t)x;
        x = (u128)m * RP2 + t2 + (uint64_t)(x >> 64);
        t1 = (uint64_t)x;
        x = (u128)m * RP3 + t3 + (uint64_t)(x >> 64);
        t2 = (uint64_t)x;
        t3 = t4 + (uint64_t)(x >> 64);
    }
    return W4{{t0, t1, t2, t3}};
}

#ifdef APG_HOST_ADX
// The same CIOS product on x86-64 with BMI2 and ADX: mulx leaves the flags
// alone, so each row runs two carry chains at once (adox adds the low halves
// of the partial products, adcx the high halves one word up) and the
// running sum rotates through five registers instead of being shifted. The
// bounds are those of w4_mul_lazy: inputs below 2p, output below 2p, every
// intermediate sum below 2^320. Callers check w4_have_adx() first.
alignas(64) inline const uint64_t W4_MOD_INV[5] = {RP0, RP1, RP2, RP3, FrP::inv64};
static_assert((uint64_t)(RP0 * FrP::inv64) == ~uint64_t(0), "-1/p mod 2^64");

inline W4 w4_mul_lazy_adx(const W4 &a, const W4 &b) {
    uint64_t t0, t1, t2, t3, t4, lo, hi, zero;
    asm(
        "movq %[b0], %%rdx\n\t"
        "mulxq %[a0], %[t0], %[t1]\n\t"
        "mulxq %[a1], %[lo], %[t2]\n\t"
        "addq %[lo], %[t1]\n\t"
        "mulxq %[a2], %[lo], %[t3]\n\t"
        "adcq %[lo], %[t2]\n\t"
        "mulxq %[a3], %[lo], %[t4]\n\t"
        "adcq %[lo], %[t3]\n\t"
        "adcq $0, %[t4]\n\t"
        "movq %[t0], %%rdx\n\t"
        "imulq %[inv], %%rdx\n\t"
        "xorl %k[zero], %k[zero]\n\t"
        "mulxq %[p0], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t0]\n\t"
        "adcxq %[hi], %[t1]\n\t"
        "mulxq %[p1], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t1]\n\t"
        "adcxq %[hi], %[t2]\n\t"
        "mulxq %[p2], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t2]\n\t"
        "adcxq %[hi], %[t3]\n\t"
        "mulxq %[p3], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t3]\n\t"
        "adcxq %[hi], %[t4]\n\t"
        "adoxq %[zero], %[t4]\n\t"
        "movq %[b1], %%rdx\n\t"
        "xorl %k[zero], %k[zero]\n\t"
        "movq %[zero], %[t0]\n\t"
        "mulxq %[a0], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t1]\n\t"
        "adcxq %[hi], %[t2]\n\t"
        "mulxq %[a1], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t2]\n\t"
        "adcxq %[hi], %[t3]\n\t"
        "mulxq %[a2], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t3]\n\t"
        "adcxq %[hi], %[t4]\n\t"
        "mulxq %[a3], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t4]\n\t"
        "adcxq %[hi], %[t0]\n\t"
        "adoxq %[zero], %[t0]\n\t"
        "movq %[t1], %%rdx\n\t"
        "imulq %[inv], %%rdx\n\t"
        "xorl %k[zero], %k[zero]\n\t"
        "mulxq %[p0], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t1]\n\t"
        "adcxq %[hi], %[t2]\n\t"
        "mulxq %[p1], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t2]\n\t"
        "adcxq %[hi], %[t3]\n\t"
        "mulxq %[p2], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t3]\n\t"
        "adcxq %[hi], %[t4]\n\t"
        "mulxq %[p3], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t4]\n\t"
        "adcxq %[hi], %[t0]\n\t"
        "adoxq %[zero], %[t0]\n\t"
        "movq %[b2], %%rdx\n\t"
        "xorl %k[zero], %k[zero]\n\t"
        "movq %[zero], %[t1]\n\t"
        "mulxq %[a0], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t2]\n\t"
        "adcxq %[hi], %[t3]\n\t"
        "mulxq %[a1], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t3]\n\t"
        "adcxq %[hi], %[t4]\n\t"
        "mulxq %[a2], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t4]\n\t"
        "adcxq %[hi], %[t0]\n\t"
        "mulxq %[a3], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t0]\n\t"
        "adcxq %[hi], %[t1]\n\t"
        "adoxq %[zero], %[t1]\n\t"
        "movq %[t2], %%rdx\n\t"
        "imulq %[inv], %%rdx\n\t"
        "xorl %k[zero], %k[zero]\n\t"
        "mulxq %[p0], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t2]\n\t"
        "adcxq %[hi], %[t3]\n\t"
        "mulxq %[p1], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t3]\n\t"
        "adcxq %[hi], %[t4]\n\t"
        "mulxq %[p2], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t4]\n\t"
        "adcxq %[hi], %[t0]\n\t"
        "mulxq %[p3], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t0]\n\t"
        "adcxq %[hi], %[t1]\n\t"
        "adoxq %[zero], %[t1]\n\t"
        "movq %[b3], %%rdx\n\t"
        "xorl %k[zero], %k[zero]\n\t"
        "movq %[zero], %[t2]\n\t"
        "mulxq %[a0], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t3]\n\t"
        "adcxq %[hi], %[t4]\n\t"
        "mulxq %[a1], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t4]\n\t"
        "adcxq %[hi], %[t0]\n\t"
        "mulxq %[a2], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t0]\n\t"
        "adcxq %[hi], %[t1]\n\t"
        "mulxq %[a3], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t1]\n\t"
        "adcxq %[hi], %[t2]\n\t"
        "adoxq %[zero], %[t2]\n\t"
        "movq %[t3], %%rdx\n\t"
        "imulq %[inv], %%rdx\n\t"
        "xorl %k[zero], %k[zero]\n\t"
        "mulxq %[p0], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t3]\n\t"
        "adcxq %[hi], %[t4]\n\t"
        "mulxq %[p1], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t4]\n\t"
        "adcxq %[hi], %[t0]\n\t"
        "mulxq %[p2], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t0]\n\t"
        "adcxq %[hi], %[t1]\n\t"
        "mulxq %[p3], %[lo], %[hi]\n\t"
        "adoxq %[lo], %[t1]\n\t"
        "adcxq %[hi], %[t2]\n\t"
        "adoxq %[zero], %[t2]\n\t"
        : [t0] "=&r"(t0), [t1] "=&r"(t1), [t2] "=&r"(t2), [t3] "=&r"(t3), [t4] "=&r"(t4), [lo] "=&r"(lo),
          [hi] "=&r"(hi), [zero] "=&r"(zero)
        : [a0] "m"(a.v[0]), [a1] "m"(a.v[1]), [a2] "m"(a.v[2]), [a3] "m"(a.v[3]), [b0] "m"(b.v[0]),
          [b1] "m"(b.v[1]), [b2] "m"(b.v[2]), [b3] "m"(b.v[3]), [p0] "m"(W4_MOD_INV[0]),
          [p1] "m"(W4_MOD_INV[1]), [p2] "m"(W4_MOD_INV[2]), [p3] "m"(W4_MOD_INV[3]), [inv] "m"(W4_MOD_INV[4])
        : "rdx", "cc");
    return W4{{t4, t0, t1, t2}};
}

// BMI2 (mulx) and ADX (adcx, adox): CPUID leaf 7, EBX bits 8 and 19.
inline bool w4_have_adx() {
    unsigned a = 0, b = 0, c = 0, d = 0;
    if (!__get_cpuid_count(7, 0, &a, &b, &c, &d)) return false;
    return ((b >> 8) & 1) && ((b >> 19) & 1);
}
#else
inline bool w4_have_adx() { return false; }
#endif

inline W4 w4_add_lazy(const W4 &a, const W4 &b) {
    u128 x = (u128)a.v[0] + b.v[0];
    const uint64_t s0 = (uint64_t)x;
    x = (u128)a.v[1] + b.v[1] + (uint64_t)(x >> 64);
    const uint64_t s1 = (uint64_t)x;
    x = (u128)a.v[2] + b.v[2] + (uint64_t)(x >> 64);
    const uint64_t s2 = (uint64_t)x;
    x = (u128)a.v[3] + b.v[3] + (uint64_t)(x >> 64);
    const uint64_t s3 = (uint64_t)x;
    // s - 2p if s >= 2p
    u128 y = (u128)s0 - R2P0;
    const uint64_t d0 = (uint64_t)y;
    y = (u128)s1 - R2P1 - (uint64_t)((y >> 64) & 1);
    const uint64_t d1 = (uint64_t)y;
    y = (u128)s2 - R2P2 - (uint64_t)((y >> 64) & 1);
    const uint64_t d2 = (uint64_t)y;
    y = (u128)s3 - R2P3 - (uint64_t)((y >> 64) & 1);
    const uint64_t d3 = (uint64_t)y;
    const bool lt = (y >> 64) & 1;
    return lt ? W4{{s0, s1, s2, s3}} : W4{{d0, d1, d2, d3}};
}

// a b mod p, fully reduced (a, b below p)
inline W4 w4_mul(const W4 &a, const W4 &b) {
    const W4 t = w4_mul_lazy(a, b);
    return w4_reduce(t.v[0], t.v[1], t.v[2], t.v[3]);
}

// a + b mod p (a, b below p: the sum is below 2p < 2^256)
inline W4 w4_add(const W4 &a, const W4 &b) {
    u128 x = (u128)a.v[0] + b.v[0];
    const uint64_t s0 = (uint64_t)x;
    x = (u128)a.v[1] + b.v[1] + (uint64_t)(x >> 64);
    const uint64_t s1 = (uint64_t)x;
    x = (u128)a.v[2] + b.v[2] + (uint64_t)(x >> 64);
    const uint64_t s2 = (uint64_t)x;
    x = (u128)a.v[3] + b.v[3] + (uint64_t)(x >> 64);
    const uint64_t s3 = (uint64_t)x;
    return w4_reduce(s0, s1, s2, s3);
}

// a - b mod p (a, b below p)
inline W4 w4_sub(const W4 &a, const W4 &b) {
    u128 x = (u128)a.v[0] - b.v[0];
    const uint64_t d0 = (uint64_t)x;
    x = (u128)a.v[1] - b.v[1] - (uint64_t)((x >> 64) & 1);
    const uint64_t d1 = (uint64_t)x;
    x = (u128)a.v[2] - b.v[2] - (uint64_t)((x >> 64) & 1);
    const uint64_t d2 = (uint64_t)x;
    x = (u128)a.v[3] - b.v[3] - (uint64_t)((x >> 64) & 1);
    const uint64_t d3 = (uint64_t)x;
    if (!((x >> 64) & 1)) return W4{{d0, d1, d2, d3}};
    // borrow: add p back
    u128 y = (u128)d0 + RP0;
    const uint64_t e0 = (uint64_t)y;
    y = (u128)d1 + RP1 + (uint64_t)(y >> 64);
    const uint64_t e1 = (uint64_t)y;
    y = (u128)d2 + RP2 + (uint64_t)(y >> 64);
    const uint64_t e2 = (uint64_t)y;
    y = (u128)d3 + RP3 + (uint64_t)(y >> 64);
    return W4{{e0, e1, e2, (uint64_t)y}};
}

}  // namespace hfr
}  // namespace apg
