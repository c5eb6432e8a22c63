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

APG_HD Xyzz xyzz29_to32(const Xyzz29 &a, bool id) {
    if (id) return xyzz_identity();
    Xyzz r;
    r.x = f29_to32<Fq>(a.x);
    r.y = f29_to32<Fq>(a.y);
    r.zz = f29_to32<Fq>(a.zz);
    r.zzz = f29_to32<Fq>(a.zzz);
    return r;
}

}  // namespace apg
