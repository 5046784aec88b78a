// BLS12-381 G1 (y^2 = x^3 + 4) point arithmetic for gfx950.
//
// Three representations:
//   G1A  affine (x, y), with (0, 0) standing for the identity (it is not on
//        the curve), as in the reference (src/shaders/field.glsl.inc:304-320);
//   G1J  Jacobian (X, Y, Z), x = X/Z^2, y = Y/Z^3, Z = 0 is the identity --
//        used for doubling chains (dbl-2009-l, 2M+5S);
//   G1X  extended Jacobian "XYZZ" (X, Y, ZZ, ZZZ), x = X/ZZ, y = Y/ZZZ,
//        ZZ = 0 is the identity -- used for accumulation (madd-2008-s 8M+2S,
//        add-2008-s 12M+2S), cheaper than the reference's Jacobian madd
//        (7M+4S) and add (11M+5S).
// Every addition handles P == Q (falls back to doubling) and P == -Q.
#pragma once

#include "field.h"

// Point additions and doublings are inlined into the kernels (passing 100-200
// byte points to an out-of-line function goes through the stack in scratch
// memory, which measured at several times the kernels' useful HBM traffic).
// The P == Q fallbacks inside additions are rare and stay out of line.
#define VKZG_DEV_CALL __device__ __forceinline__
#define VKZG_DEV_COLD __device__ __forceinline__

namespace vkzg_amd {

struct G1A {
    Fp x, y;
};
struct G1J {
    Fp x, y, z;
};
// A phase-A table entry: an affine point padded to exactly one 128-byte
// line, so each random gather touches one line instead of two.
struct __attribute__((aligned(128))) G1AT {
    G1A p;
};
struct G1X {
    Fp x, y, zz, zzz;
};

VKZG_DEV G1A g1a_identity() {
    G1A r;
    r.x = fp_zero();
    r.y = fp_zero();
    return r;
}
VKZG_DEV bool g1a_is_identity(const G1A &p) { return fp_is_zero(p.x) && fp_is_zero(p.y); }
VKZG_DEV G1A g1a_neg_if(const G1A &p, bool neg) {
    G1A r;
    r.x = p.x;
    // The identity (0, 0) stays (0, 0): fp_neg(0) == 0.
    r.y = fp_cond_neg(p.y, neg);
    return r;
}

// ---------------------------------------------------------------- Jacobian

VKZG_DEV G1J g1j_identity() {
    G1J r;
    r.x = fp_one();
    r.y = fp_one();
    r.z = fp_zero();
    return r;
}
VKZG_DEV bool g1j_is_identity(const G1J &p) { return fp_is_zero(p.z); }
VKZG_DEV G1J g1j_from_affine(const G1A &a) {
    if (g1a_is_identity(a)) return g1j_identity();
    G1J r;
    r.x = a.x;
    r.y = a.y;
    r.z = fp_one();
    return r;
}

VKZG_DEV_CALL G1J g1j_dbl(const G1J &p) {
    if (fp_is_zero(p.z)) return p;
    const Fp A = fp_sqr(p.x);
    const Fp B = fp_sqr(p.y);
    const Fp C = fp_sqr(B);
    Fp t = fp_add(p.x, B);
    t = fp_sqr(t);
    t = fp_sub(t, A);
    t = fp_sub(t, C);
    const Fp D = fp_dbl(t);
    const Fp E = fp_add(fp_dbl(A), A);
    const Fp F = fp_sqr(E);
    G1J r;
    r.x = fp_sub(F, fp_dbl(D));
    const Fp C8 = fp_dbl(fp_dbl(fp_dbl(C)));
    r.y = fp_sub(fp_mul(E, fp_sub(D, r.x)), C8);
    r.z = fp_dbl(fp_mul(p.y, p.z));
    return r;
}

VKZG_DEV_COLD G1J g1j_dbl_cold(const G1J &p);

// add-2007-bl
VKZG_DEV_CALL G1J g1j_add(const G1J &a, const G1J &b) {
    if (fp_is_zero(a.z)) return b;
    if (fp_is_zero(b.z)) return a;
    const Fp Z1Z1 = fp_sqr(a.z);
    const Fp Z2Z2 = fp_sqr(b.z);
    const Fp U1 = fp_mul(a.x, Z2Z2);
    const Fp U2 = fp_mul(b.x, Z1Z1);
    const Fp S1 = fp_mul(a.y, fp_mul(b.z, Z2Z2));
    const Fp S2 = fp_mul(b.y, fp_mul(a.z, Z1Z1));
    if (fp_eq(U1, U2)) {
        if (fp_eq(S1, S2)) return g1j_dbl_cold(a);
        return g1j_identity();
    }
    const Fp H = fp_sub(U2, U1);
    const Fp I = fp_sqr(fp_dbl(H));
    const Fp J = fp_mul(H, I);
    const Fp K = fp_dbl(fp_sub(S2, S1));
    const Fp V = fp_mul(U1, I);
    G1J r;
    r.x = fp_sub(fp_sub(fp_sqr(K), J), fp_dbl(V));
    r.y = fp_sub(fp_mul(K, fp_sub(V, r.x)), fp_dbl(fp_mul(S1, J)));
    Fp z = fp_sqr(fp_add(a.z, b.z));
    z = fp_sub(fp_sub(z, Z1Z1), Z2Z2);
    r.z = fp_mul(z, H);
    return r;
}

VKZG_DEV_COLD G1J g1j_dbl_cold(const G1J &p) { return g1j_dbl(p); }

// madd-2007-bl (b affine, may be the identity)
VKZG_DEV_CALL G1J g1j_madd(const G1J &a, const G1A &b) {
    if (g1a_is_identity(b)) return a;
    if (fp_is_zero(a.z)) return g1j_from_affine(b);
    const Fp Z1Z1 = fp_sqr(a.z);
    const Fp U2 = fp_mul(b.x, Z1Z1);
    const Fp S2 = fp_mul(b.y, fp_mul(a.z, Z1Z1));
    if (fp_eq(a.x, U2)) {
        if (fp_eq(a.y, S2)) return g1j_dbl_cold(a);
        return g1j_identity();
    }
    const Fp H = fp_sub(U2, a.x);
    const Fp HH = fp_sqr(H);
    const Fp I = fp_dbl(fp_dbl(HH));
    const Fp J = fp_mul(H, I);
    const Fp K = fp_dbl(fp_sub(S2, a.y));
    const Fp V = fp_mul(a.x, I);
    G1J r;
    r.x = fp_sub(fp_sub(fp_sqr(K), J), fp_dbl(V));
    r.y = fp_sub(fp_mul(K, fp_sub(V, r.x)), fp_dbl(fp_mul(a.y, J)));
    Fp z = fp_sqr(fp_add(a.z, H));
    r.z = fp_sub(fp_sub(z, Z1Z1), HH);
    return r;
}

// Jacobian + affine for hot loops (see g1x_madd_fast for the contract).
VKZG_DEV G1J g1j_madd_fast(const G1J &a, const G1A &b, bool &exc) {
    if (g1a_is_identity(b)) return a;
    if (fp_is_zero(a.z)) return g1j_from_affine(b);
    const Fp Z1Z1 = fp_sqr(a.z);
    const Fp U2 = fp_mul(b.x, Z1Z1);
    const Fp S2 = fp_mul(b.y, fp_mul(a.z, Z1Z1));
    const Fp H = fp_sub(U2, a.x);
    exc |= fp_is_zero(H);
    const Fp HH = fp_sqr(H);
    const Fp I = fp_dbl(fp_dbl(HH));
    const Fp J = fp_mul(H, I);
    const Fp K = fp_dbl(fp_sub(S2, a.y));
    const Fp V = fp_mul(a.x, I);
    G1J r;
    r.x = fp_sub(fp_sub(fp_sqr(K), J), fp_dbl(V));
    r.y = fp_sub(fp_mul(K, fp_sub(V, r.x)), fp_dbl(fp_mul(a.y, J)));
    r.z = fp_sub(fp_sub(fp_sqr(fp_add(a.z, H)), Z1Z1), HH);
    return r;
}

VKZG_DEV G1J g1j_neg(const G1J &p) {
    G1J r = p;
    r.y = fp_neg(p.y);
    return r;
}

// ---------------------------------------------------------------- XYZZ

VKZG_DEV G1X g1x_identity() {
    G1X r;
    r.x = fp_one();
    r.y = fp_one();
    r.zz = fp_zero();
    r.zzz = fp_zero();
    return r;
}
VKZG_DEV bool g1x_is_identity(const G1X &p) { return fp_is_zero(p.zz); }
VKZG_DEV G1X g1x_from_affine(const G1A &a) {
    if (g1a_is_identity(a)) return g1x_identity();
    G1X r;
    r.x = a.x;
    r.y = a.y;
    r.zz = fp_one();
    r.zzz = fp_one();
    return r;
}
VKZG_DEV G1X g1x_from_jacobian(const G1J &p) {
    if (fp_is_zero(p.z)) return g1x_identity();
    G1X r;
    r.x = p.x;
    r.y = p.y;
    r.zz = fp_sqr(p.z);
    r.zzz = fp_mul(r.zz, p.z);
    return r;
}
// (X, Y, ZZ, ZZZ) with ZZ^3 == ZZZ^2 -> Jacobian with Z = ZZ*ZZZ:
// X_J = X*ZZ^4, Y_J = Y*ZZZ^4 (no inversion).
VKZG_DEV_CALL G1J g1j_from_xyzz(const G1X &p) {
    if (fp_is_zero(p.zz)) return g1j_identity();
    G1J r;
    const Fp zz2 = fp_sqr(p.zz);
    const Fp zzz2 = fp_sqr(p.zzz);
    r.x = fp_mul(p.x, fp_sqr(zz2));
    r.y = fp_mul(p.y, fp_sqr(zzz2));
    r.z = fp_mul(p.zz, p.zzz);
    return r;
}

// dbl-2008-s-1
VKZG_DEV_CALL G1X g1x_dbl(const G1X &p) {
    if (fp_is_zero(p.zz)) return p;
    const Fp U = fp_dbl(p.y);
    const Fp V = fp_sqr(U);
    const Fp W = fp_mul(U, V);
    const Fp S = fp_mul(p.x, V);
    const Fp X2 = fp_sqr(p.x);
    const Fp M = fp_add(fp_dbl(X2), X2);
    G1X r;
    r.x = fp_sub(fp_sqr(M), fp_dbl(S));
    r.y = fp_sub(fp_mul(M, fp_sub(S, r.x)), fp_mul(W, p.y));
    r.zz = fp_mul(V, p.zz);
    r.zzz = fp_mul(W, p.zzz);
    return r;
}

VKZG_DEV_COLD G1X g1x_dbl_cold(const G1X &p) { return g1x_dbl(p); }

// g1x_dbl without the identity test: the identity (ZZ = ZZZ = 0) doubles to
// ZZ3 = ZZZ3 = 0, the identity again, so no input has to stay live for a
// select.  Never flags: doubling has no exceptional case.
VKZG_DEV G1X g1x_dbl_core(const G1X &p) {
    const Fp U = fp_dbl(p.y);
    const Fp V = fp_sqr(U);
    const Fp W = fp_mul(U, V);
    G1X r;
    r.zz = fp_mul(V, p.zz);
    r.zzz = fp_mul(W, p.zzz);
    const Fp S = fp_mul(p.x, V);
    const Fp X2 = fp_sqr(p.x);
    const Fp M = fp_add(fp_dbl(X2), X2);
    r.x = fp_sub(fp_sqr(M), fp_dbl(S));
    r.y = fp_mul_sum(M, fp_sub_lazy(S, r.x), p.y, fp_neg_lazy(W));
    return r;
}

// madd-2008-s: XYZZ + affine (b may be the identity)
VKZG_DEV_CALL G1X g1x_madd(const G1X &a, const G1A &b) {
    if (g1a_is_identity(b)) return a;
    if (fp_is_zero(a.zz)) return g1x_from_affine(b);
    const Fp U2 = fp_mul(b.x, a.zz);
    const Fp S2 = fp_mul(b.y, a.zzz);
    const Fp P = fp_sub(U2, a.x);
    const Fp R = fp_sub(S2, a.y);
    if (fp_is_zero(P)) {
        if (fp_is_zero(R)) return g1x_dbl_cold(g1x_from_affine(b));
        return g1x_identity();
    }
    const Fp PP = fp_sqr(P);
    const Fp PPP = fp_mul(P, PP);
    const Fp Q = fp_mul(a.x, PP);
    G1X r;
    r.x = fp_sub(fp_sub(fp_sqr(R), PPP), fp_dbl(Q));
    r.y = fp_sub(fp_mul(R, fp_sub(Q, r.x)), fp_mul(a.y, PPP));
    r.zz = fp_mul(a.zz, PP);
    r.zzz = fp_mul(a.zzz, PPP);
    return r;
}

// Hot-loop variants.  When the inputs are equal or opposite (the doubling /
// cancellation cases, which random data never reaches) they set `exc` and
// return an unspecified point; callers then recompute that result with the
// checked forms above outside their loops.  Keeping the rare path out of the
// loop body is what keeps the accumulator in VGPRs (a call inside the loop
// that takes the point by address pins it to scratch memory).  P, R and
// Q - X3 only feed products, so they are lazy (fp_sub_lazy), and P == 0 is
// detected on PP = P^2.

// Accumulator step a + (+-b) for an accumulator that is never the identity
// (callers start it from their first point).  b is an affine point in
// normalised limbs, negated when `neg`; `exc` is also set when b is the
// identity (all-zero limbs: no point of the prime-order group has y = 0).
VKZG_DEV G1X g1x_madd_acc(const G1X &a, const G1A &b, bool neg, bool &exc) {
    exc |= limbs_zero<kFpLimbs>(b.y.v);
    const Fp by = fp_neg_lazy_if(b.y, neg);
    const Fp U2 = fp_mul(b.x, a.zz);
    const Fp S2 = fp_mul(by, a.zzz);
    const Fp P = fp_sub_lazy(U2, a.x);
    const Fp R = fp_sub_lazy(S2, a.y);
    const Fp PP = fp_sqr(P);
    exc |= fp_is_zero(PP);
    const Fp PPP = fp_mul(P, PP);
    const Fp Q = fp_mul(a.x, PP);
    G1X r;
    r.x = fp_sub(fp_sub(fp_sqr(R), PPP), fp_dbl(Q));
    r.y = fp_mul_sum(R, fp_sub_lazy(Q, r.x), a.y, fp_neg_lazy(PPP));
    r.zz = fp_mul(a.zz, PP);
    r.zzz = fp_mul(a.zzz, PPP);
    return r;
}

// The first point of an accumulator: (+-b) in XYZZ; flags the identity.
VKZG_DEV G1X g1x_start(const G1A &b, bool neg, bool &exc) {
    exc |= limbs_zero<kFpLimbs>(b.y.v);
    G1X r;
    r.x = b.x;
    r.y = fp_cond_neg(b.y, neg);
    r.zz = fp_one();
    r.zzz = fp_one();
    return r;
}

// As g1x_madd_acc, but a may be the identity (used where accumulators can
// start empty).
VKZG_DEV G1X g1x_madd_fast(const G1X &a, const G1A &b, bool &exc) {
    if (g1a_is_identity(b)) return a;
    if (fp_is_zero(a.zz)) return g1x_from_affine(b);
    return g1x_madd_acc(a, b, false, exc);
}

// add-2008-s without the identity checks: a and b are not the identity.
// The products are ordered so that every input coordinate dies as early as
// possible (ZZ1 ZZ2 and ZZZ1 ZZZ2 first), which keeps the sum within a
// register budget where two other points are live.
VKZG_DEV G1X g1x_add_core(const G1X &a, const G1X &b, bool &exc) {
    const Fp U1 = fp_mul(a.x, b.zz);
    const Fp U2 = fp_mul(b.x, a.zz);
    const Fp ZZ12 = fp_mul(a.zz, b.zz);
    const Fp S1 = fp_mul(a.y, b.zzz);
    const Fp S2 = fp_mul(b.y, a.zzz);
    const Fp ZZZ12 = fp_mul(a.zzz, b.zzz);
    const Fp P = fp_sub_lazy(U2, U1);
    const Fp R = fp_sub_lazy(S2, S1);
    const Fp PP = fp_sqr(P);
    exc |= fp_is_zero(PP);
    const Fp PPP = fp_mul(P, PP);
    const Fp Q = fp_mul(U1, PP);
    G1X r;
    r.zz = fp_mul(ZZ12, PP);
    r.zzz = fp_mul(ZZZ12, PPP);
    r.x = fp_sub(fp_sub(fp_sqr(R), PPP), fp_dbl(Q));
    r.y = fp_mul_sum(R, fp_sub_lazy(Q, r.x), S1, fp_neg_lazy(PPP));
    return r;
}

// a + b for hot loops (see g1x_madd_acc for the `exc` contract); either may
// be the identity.  Branches rather than selects around the sum, so neither
// input has to stay live to the end.
VKZG_DEV G1X g1x_add_fast(const G1X &a, const G1X &b, bool &exc) {
    G1X r = a;
    if (!fp_is_zero(b.zz)) {
        if (fp_is_zero(a.zz)) r = b;
        else r = g1x_add_core(a, b, exc);
    }
    return r;
}

// add-2008-s: XYZZ + XYZZ
VKZG_DEV_CALL G1X g1x_add(const G1X &a, const G1X &b) {
    if (fp_is_zero(a.zz)) return b;
    if (fp_is_zero(b.zz)) return a;
    const Fp U1 = fp_mul(a.x, b.zz);
    const Fp U2 = fp_mul(b.x, a.zz);
    const Fp S1 = fp_mul(a.y, b.zzz);
    const Fp S2 = fp_mul(b.y, a.zzz);
    const Fp P = fp_sub(U2, U1);
    const Fp R = fp_sub(S2, S1);
    if (fp_is_zero(P)) {
        if (fp_is_zero(R)) return g1x_dbl_cold(a);
        return g1x_identity();
    }
    const Fp PP = fp_sqr(P);
    const Fp PPP = fp_mul(P, PP);
    const Fp Q = fp_mul(U1, PP);
    G1X r;
    r.x = fp_sub(fp_sub(fp_sqr(R), PPP), fp_dbl(Q));
    r.y = fp_sub(fp_mul(R, fp_sub(Q, r.x)), fp_mul(S1, PPP));
    r.zz = fp_mul(fp_mul(a.zz, b.zz), PP);
    r.zzz = fp_mul(fp_mul(a.zzz, b.zzz), PPP);
    return r;
}

// The four lanes of a quad hold the same point(s) and share one operation:
// its products are split into dependent rounds of at most four, and DPP quad
// broadcasts give every lane each round's products.  Same formulas and
// operand order as g1x_dbl (3 rounds instead of 9 products) and g1x_add_fast
// (4 rounds instead of 14); sums that only feed products stay lazy.
VKZG_DEV G1X g1x_dbl_quad(const G1X &p, uint32_t q) {
    if (fp_is_zero(p.zz)) return p;  // uniform in the quad
    const Fp U = fp_add_lazy(p.y, p.y);  // U and M feed products only
    const Fp r1 = fp_sqr((q & 1u) ? p.x : U);
    const Fp V = fp_quad_bcast<0>(r1), X2 = fp_quad_bcast<1>(r1);
    const Fp M = fp_triple_lazy(X2);
    const Fp r2 = fp_mul(q == 0 ? U : (q == 1 ? p.x : (q == 2 ? M : V)), q == 2 ? M : (q == 3 ? p.zz : V));
    const Fp W = fp_quad_bcast<0>(r2), S = fp_quad_bcast<1>(r2);
    G1X r;
    r.x = fp_sub(fp_quad_bcast<2>(r2), fp_dbl(S));  // M^2 - 2S
    r.zz = fp_quad_bcast<3>(r2);                    // V ZZ
    const Fp r3 = fp_mul(q == 0 ? M : W, q == 0 ? fp_sub_lazy(S, r.x) : (q == 1 ? p.y : p.zzz));
    r.y = fp_sub(fp_quad_bcast<0>(r3), fp_quad_bcast<1>(r3));  // M (S - X3) - W Y
    r.zzz = fp_quad_bcast<2>(r3);                               // W ZZZ
    return r;
}

VKZG_DEV G1X g1x_add_quad(const G1X &a, const G1X &b, uint32_t q, bool &exc) {
    if (fp_is_zero(a.zz)) return b;  // uniform in the quad
    if (fp_is_zero(b.zz)) return a;
    const Fp r1 = fp_mul(q == 0 ? a.x : (q == 1 ? b.x : (q == 2 ? a.y : b.y)),
                         q == 0 ? b.zz : (q == 1 ? a.zz : (q == 2 ? b.zzz : a.zzz)));
    const Fp U1 = fp_quad_bcast<0>(r1), S1 = fp_quad_bcast<2>(r1);
    const Fp P = fp_sub_lazy(fp_quad_bcast<1>(r1), U1);  // U2 - U1
    const Fp R = fp_sub_lazy(fp_quad_bcast<3>(r1), S1);  // S2 - S1
    const Fp r2 = fp_mul(q == 0 ? P : (q == 1 ? R : (q == 2 ? a.zz : a.zzz)),
                         q == 0 ? P : (q == 1 ? R : (q == 2 ? b.zz : b.zzz)));
    const Fp PP = fp_quad_bcast<0>(r2), RR = fp_quad_bcast<1>(r2);
    const Fp ZZab = fp_quad_bcast<2>(r2), ZZZab = fp_quad_bcast<3>(r2);
    exc |= fp_is_zero(PP);
    const Fp r3 = fp_mul(q == 0 ? P : (q == 1 ? U1 : ZZab), PP);
    const Fp PPP = fp_quad_bcast<0>(r3), Q = fp_quad_bcast<1>(r3);
    G1X r;
    r.zz = fp_quad_bcast<2>(r3);
    r.x = fp_sub(fp_sub(RR, PPP), fp_dbl(Q));
    const Fp r4 = fp_mul(q == 0 ? R : (q == 1 ? S1 : ZZZab), q == 0 ? fp_sub_lazy(Q, r.x) : PPP);
    r.y = fp_sub(fp_quad_bcast<0>(r4), fp_quad_bcast<1>(r4));
    r.zzz = fp_quad_bcast<2>(r4);
    return r;
}

// g1x_madd_acc shared by the four lanes of a quad that all hold a, b and
// neg: its products in four dependent rounds -- {U2, S2}, {P^2, R^2},
// {P PP, X1 PP, ZZ1 PP}, {R (Q - X3), Y1 PPP, ZZZ1 PPP} -- instead of ten
// products in a row; same operands and lazy sums, so the same field elements.
VKZG_DEV G1X g1x_madd_acc_quad(const G1X &a, const G1A &b, bool neg, uint32_t q, bool &exc) {
    exc |= limbs_zero<kFpLimbs>(b.y.v);
    const bool odd = (q & 1u) != 0;
    const Fp by = fp_neg_lazy_if(b.y, neg);
    const Fp r1 = fp_mul(fp_select(odd, by, b.x), fp_select(odd, a.zzz, a.zz));
    const Fp P = fp_sub_lazy(fp_quad_bcast<0>(r1), a.x);  // U2 - X1
    const Fp R = fp_sub_lazy(fp_quad_bcast<1>(r1), a.y);  // S2 - Y1
    const Fp r2 = fp_sqr(fp_select(odd, R, P));
    const Fp PP = fp_quad_bcast<0>(r2), RR = fp_quad_bcast<1>(r2);
    exc |= fp_is_zero(PP);
    const Fp r3 = fp_mul(q == 0 ? P : (q == 1 ? a.x : a.zz), PP);
    const Fp PPP = fp_quad_bcast<0>(r3), Q = fp_quad_bcast<1>(r3);
    G1X r;
    r.zz = fp_quad_bcast<2>(r3);
    r.x = fp_sub(fp_sub(RR, PPP), fp_dbl(Q));
    const Fp r4 = fp_mul(q == 0 ? R : (q == 1 ? a.y : a.zzz), q == 0 ? fp_sub_lazy(Q, r.x) : PPP);
    r.y = fp_sub(fp_quad_bcast<0>(r4), fp_quad_bcast<1>(r4));  // R (Q - X3) - Y1 PPP
    r.zzz = fp_quad_bcast<2>(r4);
    return r;
}

VKZG_DEV G1X g1x_neg(const G1X &p) {
    G1X r = p;
    r.y = fp_neg(p.y);
    return r;
}

// --------------------------------------------------------------- encoding

// Affine (Montgomery) -> 48-byte ZCash compressed encoding as 12 big-endian
// words: x big-endian, 0x80 compressed flag, 0x20 when y > (p-1)/2, identity
// = 0xc0 followed by zeros (k_compress_proofs.comp:21-46).
VKZG_DEV_CALL void g1a_compress_words(uint32_t *w, const G1A &p) {
    if (g1a_is_identity(p)) {
        w[0] = __builtin_bswap32(0xc0000000u);
#pragma unroll
        for (int i = 1; i < 12; i++) w[i] = 0;
        return;
    }
    const Fp x = fp_to_canonical(p.x);
    const Fp y = fp_to_canonical(p.y);
    uint32_t xw[12];
    fp_to_words(xw, x);
#pragma unroll
    for (int i = 0; i < 12; i++) w[11 - i] = xw[i];
    uint32_t flags = 0x80000000u;
    if (fp_is_lex_largest(y)) flags |= 0x20000000u;
    w[0] |= flags;
#pragma unroll
    for (int i = 0; i < 12; i++) w[i] = __builtin_bswap32(w[i]);
}

// 48-byte ZCash compressed encoding (src: its 12 words as stored, 4-byte
// aligned) -> affine.  The identity decodes to (0, 0).  False for a
// malformed encoding: no compression flag, stray bits after the infinity
// flag, x >= p or an x with no point on the curve.  No subgroup check.
VKZG_DEV bool g1a_decompress(G1A &out, const uint32_t *src) {
    uint32_t w[12];
#pragma unroll
    for (int k = 0; k < 12; k++) w[k] = __builtin_bswap32(src[k]);
    const uint32_t b0 = w[0] >> 24;
    bool ok = (b0 & 0x80u) != 0;
    if (b0 & 0x40u) {
        uint32_t rest = w[0] & 0x3fffffffu;
#pragma unroll
        for (int k = 1; k < 12; k++) rest |= w[k];
        out = g1a_identity();
        return ok && rest == 0;
    }
    uint32_t xw[12];
#pragma unroll
    for (int k = 0; k < 12; k++) xw[k] = w[11 - k];
    xw[11] &= 0x1fffffffu;
    const Fp x = fp_from_words(xw);
    if (!fp_limbs_lt_p(x)) ok = false;
    const Fp xm = fp_from_canonical(x);
    const Fp rhs = fp_add(fp_mul(fp_sqr(xm), xm), fp_const(cfp::FOUR_M));
    Fp y = fp_pow_const_window<12>(rhs, cfp::SQRT_EXP);
    if (!fp_eq(fp_sqr(y), rhs)) ok = false;
    const bool want_large = (b0 & 0x20u) != 0;
    if (fp_is_lex_largest(fp_to_canonical(y)) != want_large) y = fp_neg(y);
    out.x = xm;
    out.y = y;
    return ok;
}

}  // namespace vkzg_amd
