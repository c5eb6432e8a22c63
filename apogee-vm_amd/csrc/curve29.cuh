// The MSM's bucket accumulation in XYZZ coordinates over the lazy 29-bit Fq
// (field29.cuh): the same formulas as curve.cuh's mixed addition and affine
// doubling, with every intermediate kept in [0, 2p) and one full reduction
// when the accumulator is written back in the 32-bit form. The group element,
// and so its fully reduced coordinates, are the same as the 32-bit path's.
#pragma once
#include "curve.cuh"
#include "field29.cuh"

namespace apg {

struct Xyzz29 {
    F29 x, y, zz, zzz;
};

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

APG_HD F29 f29_one_fq() {
    const Fe one = f_one<Fq>();
    return f29_from32(one);
}

APG_HD Xyzz29 xyzz29_from_affine(const F29 &x, const F29 &y) {
    Xyzz29 r;
    r.x = x;
    r.y = y;
    r.zz = f29_one_fq();
    r.zzz = r.zz;
    return r;
}

// dbl-2008-s-1 with ZZ = ZZZ = 1: U = 2y, V = U^2, W = U V, S = x V,
// M = 3 x^2, X3 = M^2 - 2S, Y3 = M (S - X3) - W y, ZZ3 = V, ZZZ3 = W.
APG_HD Xyzz29 xyzz29_mdbl(const F29 &x, const F29 &y) {
    const F29 u = f29_add<Fq>(y, y);
    const F29 v = f29_mul<Fq>(u, u);
    const F29 w = f29_mul<Fq>(u, v);
    const F29 s = f29_mul<Fq>(x, v);
    const F29 x2 = f29_mul<Fq>(x, x);
    const F29 m = f29_add<Fq>(f29_add<Fq>(x2, x2), x2);
    Xyzz29 r;
    r.x = f29_sub<Fq>(f29_sub<Fq>(f29_mul<Fq>(m, m), s), s);
    r.y = f29_sub<Fq>(f29_mul<Fq>(m, f29_sub<Fq>(s, r.x)), f29_mul<Fq>(w, y));
    r.zz = v;
    r.zzz = w;
    return r;
}

// acc + (x, y) for a non-identity accumulator (add-2008-s with Z2 = 1);
// P = -Q leaves the identity, signalled through id.
APG_HD void xyzz29_add_affine(Xyzz29 &acc, bool &id, const F29 &x, const F29 &y) {
    const F29 u2 = f29_mul<Fq>(x, acc.zz);
    const F29 s2 = f29_mul<Fq>(y, acc.zzz);
    const F29 pd = f29_sub<Fq>(u2, acc.x);
    const F29 r = f29_sub<Fq>(s2, acc.y);
    if (f29_is_zero<Fq>(pd)) {
        if (f29_is_zero<Fq>(r))
            acc = xyzz29_mdbl(x, y);
        else
            id = true;
        return;
    }
    const F29 pp = f29_mul<Fq>(pd, pd);
    const F29 ppp = f29_mul<Fq>(pd, pp);
    const F29 q = f29_mul<Fq>(acc.x, pp);
    const F29 x3 = f29_sub<Fq>(f29_sub<Fq>(f29_sub<Fq>(f29_mul<Fq>(r, r), ppp), q), q);
    acc.y = f29_sub<Fq>(f29_mul<Fq>(r, f29_sub<Fq>(q, x3)), f29_mul<Fq>(acc.y, ppp));
    acc.x = x3;
    acc.zz = f29_mul<Fq>(acc.zz, pp);
    acc.zzz = f29_mul<Fq>(acc.zzz, ppp);
}

// Accumulates one affine point (sign-flipped if neg) into (acc, id).
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

// General XYZZ points in the lazy form for the bucket reduction, where the
// identity is ZZ = 0 mod p as in the 32-bit form.
APG_HD F29 f29_zero() {
    F29 r;
#pragma unroll
    for (int j = 0; j < 9; j++) r.v[j] = 0;
    return r;
}

APG_HD Xyzz29 xyzz29_identity() {
    Xyzz29 r;
    r.x = f29_one_fq();
    r.y = r.x;
    r.zz = f29_zero();
    r.zzz = r.zz;
    return r;
}

APG_HD bool xyzz29_is_id(const Xyzz29 &p) { return f29_is_zero<Fq>(p.zz); }

APG_HD Xyzz29 xyzz29_from32(const Xyzz &p) {
    Xyzz29 r;
    r.x = f29_from32(p.x);
    r.y = f29_from32(p.y);
    r.zz = f29_from32(p.zz);
    r.zzz = f29_from32(p.zzz);
    return r;
}

// dbl-2008-s-1 (a = 0), as curve.cuh's xyzz_dbl
APG_HD Xyzz29 xyzz29_dbl(const Xyzz29 &p) {
    if (xyzz29_is_id(p)) return p;
    const F29 u = f29_add<Fq>(p.y, p.y);
    const F29 v = f29_mul<Fq>(u, u);
    const F29 w = f29_mul<Fq>(u, v);
    const F29 s = f29_mul<Fq>(p.x, v);
    const F29 x2 = f29_mul<Fq>(p.x, p.x);
    const F29 m = f29_add<Fq>(f29_add<Fq>(x2, x2), x2);
    Xyzz29 r;
    r.x = f29_sub<Fq>(f29_sub<Fq>(f29_mul<Fq>(m, m), s), s);
    r.y = f29_sub<Fq>(f29_mul<Fq>(m, f29_sub<Fq>(s, r.x)), f29_mul<Fq>(w, p.y));
    r.zz = f29_mul<Fq>(v, p.zz);
    r.zzz = f29_mul<Fq>(w, p.zzz);
    return r;
}

// p + q (add-2008-s), as curve.cuh's xyzz_add
APG_HD Xyzz29 xyzz29_add(const Xyzz29 &p, const Xyzz29 &q) {
    if (xyzz29_is_id(p)) return q;
    if (xyzz29_is_id(q)) return p;
    const F29 u1 = f29_mul<Fq>(p.x, q.zz);
    const F29 u2 = f29_mul<Fq>(q.x, p.zz);
    const F29 s1 = f29_mul<Fq>(p.y, q.zzz);
    const F29 s2 = f29_mul<Fq>(q.y, p.zzz);
    const F29 pd = f29_sub<Fq>(u2, u1);
    const F29 r = f29_sub<Fq>(s2, s1);
    if (f29_is_zero<Fq>(pd)) {
        if (f29_is_zero<Fq>(r)) return xyzz29_dbl(p);
        return xyzz29_identity();
    }
    const F29 pp = f29_mul<Fq>(pd, pd);
    const F29 ppp = f29_mul<Fq>(pd, pp);
    const F29 qq = f29_mul<Fq>(u1, pp);
    Xyzz29 o;
    o.x = f29_sub<Fq>(f29_sub<Fq>(f29_sub<Fq>(f29_mul<Fq>(r, r), ppp), qq), qq);
    o.y = f29_sub<Fq>(f29_mul<Fq>(r, f29_sub<Fq>(qq, o.x)), f29_mul<Fq>(s1, ppp));
    o.zz = f29_mul<Fq>(f29_mul<Fq>(p.zz, q.zz), pp);
    o.zzz = f29_mul<Fq>(f29_mul<Fq>(p.zzz, q.zzz), ppp);
    return o;
}

APG_HD Xyzz xyzz29_to32(const Xyzz29 &a, bool id) {
    if (id) return xyzz_identity();
    Xyzz r;
    r.x = f29_to32<Fq>(a.x);
    r.y = f29_to32<Fq>(a.y);
    r.zz = f29_to32<Fq>(a.zz);
    r.zzz = f29_to32<Fq>(a.zzz);
    return r;
}

// fully reduced 32-bit form; the identity is written as xyzz_identity()
APG_HD Xyzz xyzz29_store(const Xyzz29 &a) { return xyzz29_to32(a, xyzz29_is_id(a)); }

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
