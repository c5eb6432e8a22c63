// BN254 G1 (y^2 = x^3 + 3 over Fq) in extended Jacobian "XYZZ" coordinates.
//
// (X, Y, ZZ, ZZZ) stands for the affine point (X/ZZ, Y/ZZZ) with ZZ^3 = ZZZ^2;
// ZZ = 0 is the identity. The MSM's bucket loop is a mixed addition
// (add-2008-s with Z2 = 1: 8M + 2S) instead of the reference's Jacobian
// madd-2007-bl (7M + 4S, crates/curve/src/g1.rs:203-243). The group element is
// the same whichever formulas produce it, so the affine result is bit-identical
// to the reference's (crates/curve/src/msm.rs:236-241 relies on the same fact).
// Both exceptional cases the reference dispatches explicitly (P == Q -> double,
// P == -Q -> identity, g1.rs:174-179) are dispatched here too.
#pragma once
#include "field.cuh"

namespace apg {

struct Affine {
    Fe x, y;  // Montgomery; all-zero = infinity (0,0 is off the curve)
};

struct Xyzz {
    Fe x, y, zz, zzz;
};

APG_HD bool aff_is_inf(const Affine &a) { return f_is_zero(a.x) && f_is_zero(a.y); }

APG_HD Xyzz xyzz_identity() {
    Xyzz r;
    r.x = f_one<Fq>();
    r.y = f_one<Fq>();
    r.zz = f_zero<Fq>();
    r.zzz = f_zero<Fq>();
    return r;
}

APG_HD bool xyzz_is_id(const Xyzz &p) { return f_is_zero(p.zz); }

APG_HD Xyzz xyzz_from_affine(const Affine &a) {
    if (aff_is_inf(a)) return xyzz_identity();
    Xyzz r;
    r.x = a.x;
    r.y = a.y;
    r.zz = f_one<Fq>();
    r.zzz = f_one<Fq>();
    return r;
}

// dbl-2008-s-1 (a = 0): U = 2Y, V = U^2, W = U V, S = X V, M = 3 X^2,
// X3 = M^2 - 2S, Y3 = M (S - X3) - W Y, ZZ3 = V ZZ, ZZZ3 = W ZZZ.
APG_HD Xyzz xyzz_dbl(const Xyzz &p) {
    if (xyzz_is_id(p)) return p;
    Fe u = f_dbl<Fq>(p.y);
    Fe v = f_sqr<Fq>(u);
    Fe w = f_mul<Fq>(u, v);
    Fe s = f_mul<Fq>(p.x, v);
    Fe x2 = f_sqr<Fq>(p.x);
    Fe m = f_add<Fq>(f_dbl<Fq>(x2), x2);
    Xyzz r;
    r.x = f_sub<Fq>(f_sqr<Fq>(m), f_dbl<Fq>(s));
    r.y = f_sub<Fq>(f_mul<Fq>(m, f_sub<Fq>(s, r.x)), f_mul<Fq>(w, p.y));
    r.zz = f_mul<Fq>(v, p.zz);
    r.zzz = f_mul<Fq>(w, p.zzz);
    return r;
}

// doubling of an affine point into XYZZ (ZZ = ZZZ = 1 specialised)
APG_HD Xyzz xyzz_mdbl(const Affine &a) {
    Fe u = f_dbl<Fq>(a.y);
    Fe v = f_sqr<Fq>(u);
    Fe w = f_mul<Fq>(u, v);
    Fe s = f_mul<Fq>(a.x, v);
    Fe x2 = f_sqr<Fq>(a.x);
    Fe m = f_add<Fq>(f_dbl<Fq>(x2), x2);
    Xyzz r;
    r.x = f_sub<Fq>(f_sqr<Fq>(m), f_dbl<Fq>(s));
    r.y = f_sub<Fq>(f_mul<Fq>(m, f_sub<Fq>(s, r.x)), f_mul<Fq>(w, a.y));
    r.zz = v;
    r.zzz = w;
    return r;
}

// p + a (mixed, add-2008-s with Z2 = 1)
APG_HD Xyzz xyzz_add_affine(const Xyzz &p, const Affine &a) {
    if (aff_is_inf(a)) return p;
    if (xyzz_is_id(p)) return xyzz_from_affine(a);
    Fe u2 = f_mul<Fq>(a.x, p.zz);
    Fe s2 = f_mul<Fq>(a.y, p.zzz);
    Fe pp_ = f_sub<Fq>(u2, p.x);
    Fe r = f_sub<Fq>(s2, p.y);
    if (f_is_zero(pp_)) {
        if (f_is_zero(r)) return xyzz_mdbl(a);
        return xyzz_identity();
    }
    Fe pp = f_sqr<Fq>(pp_);
    Fe ppp = f_mul<Fq>(pp_, pp);
    Fe q = f_mul<Fq>(p.x, pp);
    Xyzz o;
    o.x = f_sub<Fq>(f_sub<Fq>(f_sqr<Fq>(r), ppp), f_dbl<Fq>(q));
    o.y = f_sub<Fq>(f_mul<Fq>(r, f_sub<Fq>(q, o.x)), f_mul<Fq>(p.y, ppp));
    o.zz = f_mul<Fq>(p.zz, pp);
    o.zzz = f_mul<Fq>(p.zzz, ppp);
    return o;
}

// p + q (add-2008-s)
APG_HD Xyzz xyzz_add(const Xyzz &p, const Xyzz &q) {
    if (xyzz_is_id(p)) return q;
    if (xyzz_is_id(q)) return p;
    Fe u1 = f_mul<Fq>(p.x, q.zz);
    Fe u2 = f_mul<Fq>(q.x, p.zz);
    Fe s1 = f_mul<Fq>(p.y, q.zzz);
    Fe s2 = f_mul<Fq>(q.y, p.zzz);
    Fe pp_ = f_sub<Fq>(u2, u1);
    Fe r = f_sub<Fq>(s2, s1);
    if (f_is_zero(pp_)) {
        if (f_is_zero(r)) return xyzz_dbl(p);
        return xyzz_identity();
    }
    Fe pp = f_sqr<Fq>(pp_);
    Fe ppp = f_mul<Fq>(pp_, pp);
    Fe qq = f_mul<Fq>(u1, pp);
    Xyzz o;
    o.x = f_sub<Fq>(f_sub<Fq>(f_sqr<Fq>(r), ppp), f_dbl<Fq>(qq));
    o.y = f_sub<Fq>(f_mul<Fq>(r, f_sub<Fq>(qq, o.x)), f_mul<Fq>(s1, ppp));
    o.zz = f_mul<Fq>(f_mul<Fq>(p.zz, q.zz), pp);
    o.zzz = f_mul<Fq>(f_mul<Fq>(p.zzz, q.zzz), ppp);
    return o;
}

APG_HD Affine aff_neg(const Affine &a) {
    Affine r = a;
    r.y = f_neg<Fq>(a.y);  // zero stays zero, so infinity stays infinity
    return r;
}

APG_HD Xyzz xyzz_neg(const Xyzz &p) {
    Xyzz r = p;
    r.y = f_neg<Fq>(p.y);
    return r;
}

// Affine normalisation x = X / ZZ, y = Y / ZZZ with one Fermat inversion:
// invert the product ZZ * ZZZ once and split it with one multiply each.
APG_HD Affine xyzz_to_affine(const Xyzz &p) {
    Affine a;
    if (xyzz_is_id(p)) {
        a.x = f_zero<Fq>();
        a.y = f_zero<Fq>();
        return a;
    }
    Fe t = f_mul<Fq>(p.zz, p.zzz);
    Fe ti = f_inv<Fq>(t);                    // 1 / (ZZ ZZZ)
    Fe zz_inv = f_mul<Fq>(ti, p.zzz);        // 1 / ZZ
    Fe zzz_inv = f_mul<Fq>(ti, p.zz);        // 1 / ZZZ
    a.x = f_mul<Fq>(p.x, zz_inv);
    a.y = f_mul<Fq>(p.y, zzz_inv);
    return a;
}

}  // namespace apg
