// BN254 Fr / Fq Montgomery arithmetic on eight 32-bit limbs, for gfx950 VALU
// and the host.
//
// Memory form: eight little-endian u32 limbs of x * 2^256 mod m, fully reduced.
// That is byte-identical to the reference's four u64 Montgomery limbs
// (crates/field/src/lib.rs:30-36, R = 2^256), so device buffers take the
// reference's in-memory Fr/Fq images with no conversion.
//
// Multiplication is CIOS (Koc-Acar-Kaliski) over 32-bit words: every
// partial product is one 32x32->64 multiply-add (v_mad_u64_u32 on CDNA), the
// carries run through 64-bit adds (v_add_co / v_addc). Both moduli are below
// 2^254, so the accumulator stays below 2m and one conditional subtraction
// reduces it, as in the reference's u64 CIOS (crates/field/src/lib.rs:148-184).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define APG_HD __host__ __device__ __forceinline__

namespace apg {

struct Fe {
    uint32_t l[8];
};

// Fr: crates/constants/src/lib.rs:22-57. -p^-1 mod 2^32 = 0xefffffff.
struct FrP {
    APG_HD static constexpr uint32_t mod(int i) {
        constexpr uint32_t m[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                                   0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
        return m[i];
    }
    APG_HD static constexpr uint32_t one(int i) {
        constexpr uint32_t r[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                   0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
        return r[i];
    }
    APG_HD static constexpr uint32_t r2(int i) {
        constexpr uint32_t r[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                   0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
        return r[i];
    }
    static constexpr uint32_t inv = 0xefffffffu;
    static constexpr uint64_t inv64 = 0xc2e1f593efffffffull;
};

// Fq: crates/constants/src/lib.rs:97-143. -q^-1 mod 2^32 = 0xe4866389.
struct FqP {
    APG_HD static constexpr uint32_t mod(int i) {
        constexpr uint32_t m[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                                   0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
        return m[i];
    }
    APG_HD static constexpr uint32_t one(int i) {
        constexpr uint32_t r[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                   0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
        return r[i];
    }
    APG_HD static constexpr uint32_t r2(int i) {
        constexpr uint32_t r[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                   0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
        return r[i];
    }
    static constexpr uint32_t inv = 0xe4866389u;
    static constexpr uint64_t inv64 = 0x87d20782e4866389ull;
};

template <class P>
APG_HD Fe f_zero() {
    Fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.l[i] = 0;
    return r;
}

template <class P>
APG_HD Fe f_one() {
    Fe r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.l[i] = P::one(i);
    return r;
}

APG_HD bool f_is_zero(const Fe &a) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) o |= a.l[i];
    return o == 0;
}

APG_HD bool f_eq(const Fe &a, const Fe &b) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) o |= a.l[i] ^ b.l[i];
    return o == 0;
}

// r = a - m if a >= m else a (a < 2m)
template <class P>
APG_HD void f_reduce_once(Fe &a) {
    uint32_t s[8];
    uint64_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t d = (uint64_t)a.l[i] - P::mod(i) - br;
        s[i] = (uint32_t)d;
        br = (d >> 32) & 1;
    }
    // br == 0 means a >= m
#pragma unroll
    for (int i = 0; i < 8; i++) a.l[i] = br ? a.l[i] : s[i];
}

#if !defined(__HIP_DEVICE_COMPILE__)
template <class P>
inline void host_sub_mod_if_ge(uint64_t t[4]) {
    uint64_t d[4], br = 0;
    for (int i = 0; i < 4; i++) {
        const uint64_t m = (uint64_t)P::mod(2 * i) | ((uint64_t)P::mod(2 * i + 1) << 32);
        unsigned __int128 x = (unsigned __int128)t[i] - m - br;
        d[i] = (uint64_t)x;
        br = (uint64_t)(x >> 64) & 1;
    }
    if (!br)
        for (int i = 0; i < 4; i++) t[i] = d[i];
}
template <class P>
inline Fe f_add_host64(const Fe &a, const Fe &b) {
    uint64_t t[4], c = 0;
    for (int i = 0; i < 4; i++) {
        const uint64_t x = (uint64_t)a.l[2 * i] | ((uint64_t)a.l[2 * i + 1] << 32);
        const uint64_t y = (uint64_t)b.l[2 * i] | ((uint64_t)b.l[2 * i + 1] << 32);
        unsigned __int128 s = (unsigned __int128)x + y + c;
        t[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    host_sub_mod_if_ge<P>(t);
    Fe r;
    for (int i = 0; i < 4; i++) {
        r.l[2 * i] = (uint32_t)t[i];
        r.l[2 * i + 1] = (uint32_t)(t[i] >> 32);
    }
    return r;
}
template <class P>
inline Fe f_sub_host64(const Fe &a, const Fe &b) {
    uint64_t t[4], br = 0;
    for (int i = 0; i < 4; i++) {
        const uint64_t x = (uint64_t)a.l[2 * i] | ((uint64_t)a.l[2 * i + 1] << 32);
        const uint64_t y = (uint64_t)b.l[2 * i] | ((uint64_t)b.l[2 * i + 1] << 32);
        unsigned __int128 s = (unsigned __int128)x - y - br;
        t[i] = (uint64_t)s;
        br = (uint64_t)(s >> 64) & 1;
    }
    if (br) {
        uint64_t c = 0;
        for (int i = 0; i < 4; i++) {
            const uint64_t m = (uint64_t)P::mod(2 * i) | ((uint64_t)P::mod(2 * i + 1) << 32);
            unsigned __int128 s = (unsigned __int128)t[i] + m + c;
            t[i] = (uint64_t)s;
            c = (uint64_t)(s >> 64);
        }
    }
    Fe r;
    for (int i = 0; i < 4; i++) {
        r.l[2 * i] = (uint32_t)t[i];
        r.l[2 * i + 1] = (uint32_t)(t[i] >> 32);
    }
    return r;
}
#endif

template <class P>
APG_HD Fe f_add(const Fe &a, const Fe &b) {
#if !defined(__HIP_DEVICE_COMPILE__)
    return f_add_host64<P>(a, b);
#endif
    Fe r;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t s = (uint64_t)a.l[i] + b.l[i] + c;
        r.l[i] = (uint32_t)s;
        c = s >> 32;
    }
    // m < 2^254, so a + b < 2^255 never carries out
    f_reduce_once<P>(r);
    return r;
}

template <class P>
APG_HD Fe f_sub(const Fe &a, const Fe &b) {
#if !defined(__HIP_DEVICE_COMPILE__)
    return f_sub_host64<P>(a, b);
#endif
    Fe r;
    uint64_t br = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t d = (uint64_t)a.l[i] - b.l[i] - br;
        r.l[i] = (uint32_t)d;
        br = (d >> 32) & 1;
    }
    // add m back under the borrow mask
    uint32_t mask = 0u - (uint32_t)br;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t s = (uint64_t)r.l[i] + (P::mod(i) & mask) + c;
        r.l[i] = (uint32_t)s;
        c = s >> 32;
    }
    return r;
}

template <class P>
APG_HD Fe f_neg(const Fe &a) {
    if (f_is_zero(a)) return a;
    Fe m;
#pragma unroll
    for (int i = 0; i < 8; i++) m.l[i] = P::mod(i);
    return f_sub<P>(m, a);
}

template <class P>
APG_HD Fe f_dbl(const Fe &a) {
    return f_add<P>(a, a);
}

#if !defined(__HIP_DEVICE_COMPILE__)
// Host build: the same CIOS over four 64-bit words with 128-bit intermediates
// (the reference's own limb width), so host tails and the host transcript run
// at CPU speed. Same R = 2^256, so the memory images agree.
template <class P>
inline Fe f_mul_host64(const Fe &a, const Fe &b) {
    typedef unsigned __int128 u128;
    uint64_t A[4], Bv[4], M[4], t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        A[i] = (uint64_t)a.l[2 * i] | ((uint64_t)a.l[2 * i + 1] << 32);
        Bv[i] = (uint64_t)b.l[2 * i] | ((uint64_t)b.l[2 * i + 1] << 32);
        M[i] = (uint64_t)P::mod(2 * i) | ((uint64_t)P::mod(2 * i + 1) << 32);
    }
    for (int i = 0; i < 4; i++) {
        uint64_t c = 0;
        for (int j = 0; j < 4; j++) {
            u128 x = (u128)A[j] * Bv[i] + t[j] + c;
            t[j] = (uint64_t)x;
            c = (uint64_t)(x >> 64);
        }
        u128 s = (u128)t[4] + c;
        t[4] = (uint64_t)s;
        t[5] = (uint64_t)(s >> 64);
        const uint64_t m = t[0] * P::inv64;
        u128 x = (u128)m * M[0] + t[0];
        c = (uint64_t)(x >> 64);
        for (int j = 1; j < 4; j++) {
            x = (u128)m * M[j] + t[j] + c;
            t[j - 1] = (uint64_t)x;
            c = (uint64_t)(x >> 64);
        }
        s = (u128)t[4] + c;
        t[3] = (uint64_t)s;
        t[4] = t[5] + (uint64_t)(s >> 64);
    }
    // conditional subtraction of m
    uint64_t d[4], br = 0;
    for (int i = 0; i < 4; i++) {
        u128 x = (u128)t[i] - M[i] - br;
        d[i] = (uint64_t)x;
        br = (uint64_t)(x >> 64) & 1;
    }
    Fe r;
    for (int i = 0; i < 4; i++) {
        const uint64_t v = br ? t[i] : d[i];
        r.l[2 * i] = (uint32_t)v;
        r.l[2 * i + 1] = (uint32_t)(v >> 32);
    }
    return r;
}
#endif

// lo + 2^32 hi = a * b + x + y for 32-bit a, b, x, y (never overflows 64
// bits): one add-with-carry builds the 64-bit addend, one v_mad_u64_u32 does
// the rest, with no register shuffling of zero-extended operands.
APG_HD uint32_t mac2(uint32_t a, uint32_t b, uint32_t x, uint32_t y, uint32_t &hi) {
    unsigned co;
    const uint32_t lo = __builtin_addc(x, y, 0u, &co);
    const uint64_t p = (uint64_t)a * b + (((uint64_t)co << 32) | lo);
    hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

// a * b * 2^-256 mod m over eight 32-bit words: CIOS with the product and
// reduction rows interleaved and no words above the eighth. That is exact
// when the modulus' top word is below 2^31 - 1 (0x30644e72 for both BN254
// moduli): every row's accumulator stays below 2m < 2^255, so the two carries
// of a row fit in the top word (the "no-carry" CIOS bound). Inputs reduced,
// output reduced by one conditional subtraction.
template <class P>
APG_HD Fe f_mul32(const Fe &a, const Fe &b) {
    static_assert(P::mod(7) < 0x7fffffffu, "no-carry CIOS needs a spare top bit");
    uint32_t t[8];
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint32_t bi = b.l[i];
        const uint64_t p0 = (uint64_t)a.l[0] * bi + t[0];
        t[0] = (uint32_t)p0;
        uint32_t A = (uint32_t)(p0 >> 32);
        const uint32_t m = t[0] * P::inv;
        uint32_t C = (uint32_t)(((uint64_t)m * P::mod(0) + t[0]) >> 32);
#pragma unroll
        for (int j = 1; j < 8; j++) {
            t[j] = mac2(a.l[j], bi, t[j], A, A);
            t[j - 1] = mac2(m, P::mod(j), t[j], C, C);
        }
        t[7] = C + A;
    }
    Fe r;
#pragma unroll
    for (int j = 0; j < 8; j++) r.l[j] = t[j];
    f_reduce_once<P>(r);
    return r;
}

template <class P>
APG_HD Fe f_mul(const Fe &a, const Fe &b) {
#if !defined(__HIP_DEVICE_COMPILE__)
    return f_mul_host64<P>(a, b);
#else
    return f_mul32<P>(a, b);
#endif
}

template <class P>
APG_HD Fe f_sqr(const Fe &a) {
    return f_mul<P>(a, a);
}

// Montgomery form -> canonical integer limbs (one reduction: a * 1 * R^-1).
template <class P>
APG_HD Fe f_from_mont(const Fe &a) {
    Fe one;
#pragma unroll
    for (int i = 0; i < 8; i++) one.l[i] = 0;
    one.l[0] = 1;
    return f_mul<P>(a, one);
}

// canonical integer limbs (< m) -> Montgomery form (times R^2, reduce once).
template <class P>
APG_HD Fe f_to_mont(const Fe &a) {
    Fe r2;
#pragma unroll
    for (int i = 0; i < 8; i++) r2.l[i] = P::r2(i);
    return f_mul<P>(a, r2);
}

template <class P>
APG_HD Fe f_from_u64(uint64_t x) {
    Fe a = f_zero<P>();
    a.l[0] = (uint32_t)x;
    a.l[1] = (uint32_t)(x >> 32);
    return f_to_mont<P>(a);
}

// a^e for a 256-bit little-endian exponent, MSB first.
template <class P>
APG_HD Fe f_pow(const Fe &a, const uint32_t e[8]) {
    Fe acc = f_one<P>();
    for (int w = 7; w >= 0; w--)
        for (int bit = 31; bit >= 0; bit--) {
            acc = f_sqr<P>(acc);
            if ((e[w] >> bit) & 1) acc = f_mul<P>(acc, a);
        }
    return acc;
}

// Fermat inverse a^(m-2); zero maps to zero.
template <class P>
APG_HD Fe f_inv(const Fe &a) {
    uint32_t e[8];
#pragma unroll
    for (int i = 0; i < 8; i++) e[i] = P::mod(i);
    e[0] -= 2;  // both moduli end in an odd word > 2
    return f_pow<P>(a, e);
}

// canonical little-endian bytes (value < m) check
template <class P>
APG_HD bool f_is_canonical(const Fe &a) {
    for (int i = 7; i >= 0; i--) {
        if (a.l[i] != P::mod(i)) return a.l[i] < P::mod(i);
    }
    return false;
}

using Fr = FrP;
using Fq = FqP;

}  // namespace apg
