// The MSM's bucket phase in XYZZ coordinates over the lazy 29-bit Fq of
// field29.cuh (Montgomery domain R' = 2^261): the same formulas as
// curve.cuh's mixed addition, addition and doublings. The group element, and
// so its fully reduced coordinates after leaving the domain, are the same as
// the 32-bit path's.
//
// Values are kept below fixed multiples of p instead of below p. With table
// coordinates below p, every accumulator or bucket point of the phase has
// X < 5.2p, Y < 3.3p, ZZ < 1.1p and ZZZ < 1.1p (the bounds each formula
// below restates); a product of two values below 12p is below 1.1p, so the
// subtractions add 2p, 4p or 8p above the subtrahend's bound and nothing is
// reduced until the phase's output. The identity is ZZ = ZZZ = 0.
#pragma once
#include "curve.cuh"
#include "field29.cuh"

namespace apg {

struct Xyzz29 {
    F29 x, y, zz, zzz;
};

APG_HD F29 f29_zero() {
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = 0;
    return r;
}

APG_HD Xyzz29 xyzz29_from_affine(const F29 &x, const F29 &y) {
    Xyzz29 r;
    r.x = x;
    r.y = y;
    r.zz = f29_const_one<Fq>();
    r.zzz = r.zz;
    return r;
}

APG_HD Xyzz29 xyzz29_identity() {
    Xyzz29 r;
    r.x = f29_const_one<Fq>();
    r.y = r.x;
    r.zz = f29_zero();
    r.zzz = r.zz;
    return r;
}

// ZZ = 0 mod p; ZZ < 1.1p, so ZZ is 0 or p
APG_HD bool xyzz29_is_id(const Xyzz29 &p) { return f29_is_zero<Fq, 2>(p.zz); }

// dbl-2008-s-1 with ZZ = ZZZ = 1 for x, y < p (bounds in the comments):
// U = 2y, V = U^2, W = U V, S = x V, M = 3 x^2, X3 = M^2 - 2S,
// Y3 = M (S - X3) - W y, ZZ3 = V, ZZZ3 = W. X3 is M^2 - (S + S) + 4p, the
// sum taken word by word and subtracted once (2S < 4p).
APG_HD Xyzz29 xyzz29_mdbl(const F29 &x, const F29 &y) {
    const F29 u = f29_add(y, y);                 // < 2p
    const F29 v = f29_sqr<Fq>(u);                // < 1.03p
    const F29 w = f29_mul<Fq>(u, v);             // < 1.02p
    const F29 s = f29_mul<Fq>(x, v);             // < 1.01p
    const F29 x2 = f29_sqr<Fq>(x);               // < 1.01p
    const F29 m = f29_add(f29_add(x2, x2), x2);  // < 3.03p
    Xyzz29 r;
    r.x = f29_sub<Fq, 4>(f29_sqr<Fq>(m), f29_add_words(s, s));  // < 5.06p
    // M (S - X3) - W y as one reduction: M (S - X3 + 8p) + W (2p - y)
    r.y = f29_mul2<Fq>(m, f29_sub<Fq, 8>(s, r.x), w, f29_sub<Fq, 2>(F29{}, y));  // < 1.18p
    r.zz = v;
    r.zzz = w;
    return r;
}

// acc + (x, y) for a non-identity accumulator (add-2008-s with Z2 = 1), the
// point below p; P = -Q leaves the identity, signalled through id.
APG_HD void xyzz29_add_affine(Xyzz29 &acc, bool &id, const F29 &x, const F29 &y) {
    const F29 u2 = f29_mul<Fq>(x, acc.zz);     // < 1.01p
    const F29 s2 = f29_mul<Fq>(y, acc.zzz);    // < 1.01p
    const F29 pd = f29_sub<Fq, 8>(u2, acc.x);  // < 9.01p (acc.x < 5.2p)
    const F29 r = f29_sub<Fq, 4>(s2, acc.y);   // < 5.01p (acc.y < 3.3p)
    if (f29_is_zero<Fq, 9>(pd)) {
        if (f29_is_zero<Fq, 5>(r))
            acc = xyzz29_mdbl(x, y);
        else
            id = true;
        return;
    }
    // ordered so that each coordinate of acc is replaced as soon as its last
    // use is done (fewer live values)
    const F29 pp = f29_sqr<Fq>(pd);          // < 1.49p
    acc.zz = f29_mul<Fq>(acc.zz, pp);        // < 1.01p
    const F29 ppp = f29_mul<Fq>(pd, pp);     // < 1.08p
    acc.zzz = f29_mul<Fq>(acc.zzz, ppp);     // < 1.01p
    const F29 q = f29_mul<Fq>(acc.x, pp);    // < 1.07p
    // R^2 - (PPP + 2Q) + 4p, the sum word by word (PPP + 2Q < 3.22p)
    const F29 x3 = f29_sub<Fq, 4>(f29_sqr<Fq>(r), f29_sum3_words(ppp, q, q));  // < 5.15p
    // R (Q - X3) - Y1 PPP as one reduction: R (Q - X3 + 8p) + Y1 (2p - PPP);
    // (5.01 x 9.07 + 3.3 x 2) p^2 / R' + p with R' / p ~ 169: < 1.31p
    acc.y = f29_mul2<Fq>(r, f29_sub<Fq, 8>(q, x3), acc.y, f29_sub<Fq, 2>(F29{}, ppp));
    acc.x = x3;
}

// Accumulates one base-table point (coordinates in the domain, below p;
// all-zero = infinity), sign-flipped if neg, into (acc, id).
APG_HD void xyzz29_accumulate(Xyzz29 &acc, bool &id, const Affine &a, bool neg) {
    if (aff_is_inf(a)) return;
    const F29 x = f29_from32(a.x);
    const F29 y0 = f29_from32(a.y);
    const F29 y = neg ? f29_neg_reduced<Fq>(y0) : y0;
    if (id) {
        acc = xyzz29_from_affine(x, y);
        id = false;
        return;
    }
    xyzz29_add_affine(acc, id, x, y);
}

// dbl-2008-s-1 (a = 0), as curve.cuh's xyzz_dbl, on a point within the bounds
APG_HD Xyzz29 xyzz29_dbl(const Xyzz29 &p) {
    if (xyzz29_is_id(p)) return p;
    const F29 u = f29_add(p.y, p.y);             // < 6.6p
    const F29 v = f29_sqr<Fq>(u);                // < 1.26p
    const F29 w = f29_mul<Fq>(u, v);             // < 1.05p
    const F29 s = f29_mul<Fq>(p.x, v);           // < 1.06p
    const F29 x2 = f29_sqr<Fq>(p.x);             // < 1.31p
    const F29 m = f29_add(f29_add(x2, x2), x2);  // < 3.93p
    Xyzz29 r;
    r.x = f29_sub<Fq, 4>(f29_sqr<Fq>(m), f29_add_words(s, s));  // < 5.1p (2S < 2.12p)
    // M (S - X3) - W Y as one reduction: M (S - X3 + 8p) + W (4p - Y), Y < 3.3p
    r.y = f29_mul2<Fq>(m, f29_sub<Fq, 8>(s, r.x), w, f29_sub<Fq, 4>(F29{}, p.y));  // < 1.24p
    r.zz = f29_mul<Fq>(v, p.zz);                 // < 1.01p
    r.zzz = f29_mul<Fq>(w, p.zzz);               // < 1.01p
    return r;
}

// p + q (add-2008-s), as curve.cuh's xyzz_add, on points within the bounds
APG_HD Xyzz29 xyzz29_add(const Xyzz29 &p, const Xyzz29 &q) {
    if (xyzz29_is_id(p)) return q;
    if (xyzz29_is_id(q)) return p;
    const F29 u1 = f29_mul<Fq>(p.x, q.zz);   // < 1.05p
    const F29 u2 = f29_mul<Fq>(q.x, p.zz);
    const F29 s1 = f29_mul<Fq>(p.y, q.zzz);  // < 1.03p
    const F29 s2 = f29_mul<Fq>(q.y, p.zzz);
    const F29 pd = f29_sub<Fq, 2>(u2, u1);   // < 3.05p
    const F29 r = f29_sub<Fq, 2>(s2, s1);    // < 3.03p
    if (f29_is_zero<Fq, 3>(pd)) {
        if (f29_is_zero<Fq, 3>(r)) return xyzz29_dbl(p);
        return xyzz29_identity();
    }
    const F29 pp = f29_sqr<Fq>(pd);          // < 1.06p
    const F29 ppp = f29_mul<Fq>(pd, pp);     // < 1.02p
    const F29 qq = f29_mul<Fq>(u1, pp);      // < 1.01p
    Xyzz29 o;
    // R^2 - (PPP + 2QQ) + 4p, the sum word by word (PPP + 2QQ < 3.04p)
    o.x = f29_sub<Fq, 4>(f29_sqr<Fq>(r), f29_sum3_words(ppp, qq, qq));  // < 5.06p
    // R (QQ - X3) - S1 PPP as one reduction: R (QQ - X3 + 8p) + S1 (2p - PPP)
    o.y = f29_mul2<Fq>(r, f29_sub<Fq, 8>(qq, o.x), s1, f29_sub<Fq, 2>(F29{}, ppp));  // < 1.18p
    o.zz = f29_mul<Fq>(f29_mul<Fq>(p.zz, q.zz), pp);      // < 1.01p
    o.zzz = f29_mul<Fq>(f29_mul<Fq>(p.zzz, q.zzz), ppp);  // < 1.01p
    return o;
}

// a 32-bit Montgomery point (R = 2^256, coordinates below 2p) in the domain
APG_HD Xyzz29 xyzz29_from32(const Xyzz &p) {
    Xyzz29 r;
    r.x = f29_from_mont32<Fq>(p.x);
    r.y = f29_from_mont32<Fq>(p.y);
    r.zz = f29_from_mont32<Fq>(p.zz);
    r.zzz = f29_from_mont32<Fq>(p.zzz);
    return r;
}

// leaves the domain: fully reduced 32-bit Montgomery coordinates
APG_HD Xyzz xyzz29_to32(const Xyzz29 &a, bool id) {
    if (id) return xyzz_identity();
    Xyzz r;
    r.x = f29_to32<Fq>(a.x);
    r.y = f29_to32<Fq>(a.y);
    r.zz = f29_to32<Fq>(a.zz);
    r.zzz = f29_to32<Fq>(a.zzz);
    return r;
}

APG_HD Xyzz xyzz29_store(const Xyzz29 &a) { return xyzz29_to32(a, xyzz29_is_id(a)); }

// An affine point (32-bit Montgomery, R = 2^256) as the base tables hold it:
// coordinates times 2^5 mod p, fully reduced; infinity stays all-zero.
APG_HD Affine affine_to_table(const Affine &a) {
    if (aff_is_inf(a)) return a;
    Affine r;
    r.x = f_to_r261<Fq>(a.x);
    r.y = f_to_r261<Fq>(a.y);
    return r;
}

// Memory image of an Xyzz29 (36 words, 144 B): the bucket phase keeps points
// in the lazy form between kernels, so a write is stores only. Arrays of it
// are 16-byte aligned and move as nine 16-byte vectors.
__device__ __forceinline__ Xyzz29 ld29(const Xyzz29 *p) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    uint32_t w[36];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const uint4 v = q[i];
        w[4 * i] = v.x;
        w[4 * i + 1] = v.y;
        w[4 * i + 2] = v.z;
        w[4 * i + 3] = v.w;
    }
    Xyzz29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        r.x.v[j] = w[j];
        r.y.v[j] = w[9 + j];
        r.zz.v[j] = w[18 + j];
        r.zzz.v[j] = w[27 + j];
    }
    return r;
}

__device__ __forceinline__ void st29(Xyzz29 *p, const Xyzz29 &a) {
    uint32_t w[36];
#pragma unroll
    for (int j = 0; j < 9; j++) {
        w[j] = a.x.v[j];
        w[9 + j] = a.y.v[j];
        w[18 + j] = a.zz.v[j];
        w[27 + j] = a.zzz.v[j];
    }
    uint4 *q = reinterpret_cast<uint4 *>(p);
#pragma unroll
    for (int i = 0; i < 9; i++) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// the accumulator of a bucket loop as stored: the identity when id is set
__device__ __forceinline__ void st29_acc(Xyzz29 *p, const Xyzz29 &a, bool id) {
    Xyzz29 v = a;
#pragma unroll
    for (int j = 0; j < 9; j++) {
        v.zz.v[j] = id ? 0u : a.zz.v[j];
        v.zzz.v[j] = id ? 0u : a.zzz.v[j];
    }
    st29(p, v);
}

}  // namespace apg
