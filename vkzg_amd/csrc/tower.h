// Extension-field and G2 arithmetic for the pairing of cell-proof
// verification (the reference has no verifier: include/vkzg.h:14-18; the
// verification equation is the consensus specs' verify_cell_kzg_proof_batch).
//
//   Fp2  = Fp[u]/(u^2 + 1)
//   Fp12 = Fp2[w]/(w^6 - xi), xi = 1 + u; an element is sum_m a_m w^m with
//          a_m in Fp2, held as 12 Fp values, index 2m + e (e = 0: real part).
//   G2 is on the twist E': y^2 = x^3 + 4 xi over Fp2; psi(x, y) =
//   (x w^-2, y w^-3) maps it into E(Fp12).
//
// Fp2 operations run on one lane.  Fp12 products run on a whole workgroup:
// the 144 Fp products a_i * b_j go to 144 lanes at once (one Montgomery
// multiplication of latency instead of 54-144 in sequence), then 72 lanes
// combine them into Fp2 partial sums and 12 lanes add those up.  That keeps
// the pairing's long dependent chain (about 500 Fp12 products) at roughly a
// millisecond on one CU, where a one-lane tower would take over ten.
#pragma once

#include "curve.h"

namespace vkzg_amd {

struct Fp2 {
    Fp c0, c1;
};
struct G2A {
    Fp2 x, y;
};

VKZG_DEV Fp2 fp2_zero() { return {fp_zero(), fp_zero()}; }
VKZG_DEV Fp2 fp2_one() { return {fp_one(), fp_zero()}; }
VKZG_DEV Fp2 fp2_add(const Fp2 &a, const Fp2 &b) { return {fp_add(a.c0, b.c0), fp_add(a.c1, b.c1)}; }
VKZG_DEV Fp2 fp2_sub(const Fp2 &a, const Fp2 &b) { return {fp_sub(a.c0, b.c0), fp_sub(a.c1, b.c1)}; }
VKZG_DEV Fp2 fp2_neg(const Fp2 &a) { return {fp_neg(a.c0), fp_neg(a.c1)}; }
VKZG_DEV Fp2 fp2_conj(const Fp2 &a) { return {a.c0, fp_neg(a.c1)}; }
VKZG_DEV bool fp2_is_zero(const Fp2 &a) { return fp_is_zero(a.c0) && fp_is_zero(a.c1); }
VKZG_DEV bool fp2_eq(const Fp2 &a, const Fp2 &b) { return fp_eq(a.c0, b.c0) && fp_eq(a.c1, b.c1); }
// Karatsuba: (a0 + a1 u)(b0 + b1 u) = a0 b0 - a1 b1 + ((a0 + a1)(b0 + b1) - a0 b0 - a1 b1) u
VKZG_DEV Fp2 fp2_mul(const Fp2 &a, const Fp2 &b) {
    const Fp t0 = fp_mul(a.c0, b.c0);
    const Fp t1 = fp_mul(a.c1, b.c1);
    const Fp t2 = fp_mul(fp_add(a.c0, a.c1), fp_add(b.c0, b.c1));
    return {fp_sub(t0, t1), fp_sub(fp_sub(t2, t0), t1)};
}
// (a0 + a1 u)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 u
VKZG_DEV Fp2 fp2_sqr(const Fp2 &a) {
    return {fp_mul(fp_add(a.c0, a.c1), fp_sub(a.c0, a.c1)), fp_dbl(fp_mul(a.c0, a.c1))};
}
VKZG_DEV Fp2 fp2_mul_fp(const Fp2 &a, const Fp &s) { return {fp_mul(a.c0, s), fp_mul(a.c1, s)}; }
// a * xi = (a0 - a1) + (a0 + a1) u
VKZG_DEV Fp2 fp2_mul_xi(const Fp2 &a) { return {fp_sub(a.c0, a.c1), fp_add(a.c0, a.c1)}; }
// 1/a = conj(a) / (a0^2 + a1^2); 0 -> 0
VKZG_DEV Fp2 fp2_inv(const Fp2 &a) {
    const Fp n = fp_add(fp_sqr(a.c0), fp_sqr(a.c1));
    const Fp ni = fp_inv(n);
    return {fp_mul(a.c0, ni), fp_neg(fp_mul(a.c1, ni))};
}
template <int NE>
VKZG_DEV Fp2 fp2_pow_const(const Fp2 &a, const uint32_t (&e)[NE]) {
    Fp2 acc = fp2_one();
#pragma unroll 1
    for (int i = NE - 1; i >= 0; i--) {
        const uint32_t limb = e[i];
#pragma unroll 1
        for (int b = 31; b >= 0; b--) {
            acc = fp2_sqr(acc);
            if ((limb >> b) & 1u) acc = fp2_mul(acc, a);
        }
    }
    return acc;
}

// Square root in Fp2 for p = 3 mod 4 (Adj and Rodriguez-Henriquez, "Square
// root computation over even extension fields", 2014, Algorithm 9):
//   a1 = a^((p-3)/4), alpha = a1^2 a, x0 = a1 a;
//   alpha = -1: x = u x0; otherwise x = (1 + alpha)^((p-1)/2) x0.
// The result is checked by squaring; false when a is not a square.
VKZG_DEV bool fp2_sqrt(Fp2 &out, const Fp2 &a) {
    const Fp2 a1 = fp2_pow_const<12>(a, cfp::P_MINUS_3_DIV_4);
    const Fp2 alpha = fp2_mul(fp2_sqr(a1), a);
    const Fp2 x0 = fp2_mul(a1, a);
    const Fp2 minus_one = {fp_neg(fp_one()), fp_zero()};
    Fp2 x;
    if (fp2_eq(alpha, minus_one)) {
        x = {fp_neg(x0.c1), x0.c0};  // u * x0
    } else {
        const Fp2 b = fp2_pow_const<12>(fp2_add(fp2_one(), alpha), cfp::P_MINUS_1_DIV_2);
        x = fp2_mul(b, x0);
    }
    out = x;
    return fp2_eq(fp2_sqr(x), a);
}

// Sign of the compressed G2 encoding: y is "larger" when y.c1 > (p-1)/2, or
// y.c1 == 0 and y.c0 > (p-1)/2 (the ZCash serialisation of BLS12-381).
VKZG_DEV bool fp2_is_lex_largest(const Fp2 &y) {
    const Fp c1 = fp_to_canonical(y.c1);
    if (!limbs_zero<kFpLimbs>(c1.v)) return fp_is_lex_largest(c1);
    return fp_is_lex_largest(fp_to_canonical(y.c0));
}

// ------------------------------------------------------------ Fp12 in LDS
// Work buffers of one workgroup: 144 products and 72 partial sums.
struct Fp12Scratch {
    Fp prod[144];
    Fp part[72];
};

// dst = a * b for Fp12 values held in LDS (12 Fp each).  Every thread of the
// block calls it (it contains barriers); blockDim.x >= 144.  dst may alias a
// or b.  Product a_{m1} b_{m2} lands on w^(m1+m2); a power m + 6 is
// multiplied by xi: (x + y u) xi = (x - y) + (x + y) u.  Inlined, so that
// every operand access is an LDS access (a generic pointer would make them
// flat loads with several times the latency).
__device__ __forceinline__ void fp12_mul_lds(Fp *dst, const Fp *a, const Fp *b, Fp12Scratch &s) {
    const uint32_t t = threadIdx.x;
    if (t < 144) s.prod[t] = fp_mul(a[t / 12], b[t % 12]);
    __syncthreads();
    if (t < 72) {
        // output o = 2m + e, term m1 (m2 = m - m1 mod 6)
        const uint32_t o = t / 6, m1 = t % 6, m = o >> 1, e = o & 1u;
        const uint32_t m2 = (m + 6 - m1) % 6;
        const Fp &p00 = s.prod[(2 * m1) * 12 + 2 * m2];
        const Fp &p01 = s.prod[(2 * m1) * 12 + 2 * m2 + 1];
        const Fp &p10 = s.prod[(2 * m1 + 1) * 12 + 2 * m2];
        const Fp &p11 = s.prod[(2 * m1 + 1) * 12 + 2 * m2 + 1];
        const Fp x = fp_sub(p00, p11);  // real part of a_{m1} b_{m2}
        const Fp y = fp_add(p01, p10);  // imaginary part
        Fp v;
        if (m1 <= m) v = e ? y : x;
        else v = e ? fp_add(x, y) : fp_sub(x, y);
        s.part[t] = v;
    }
    __syncthreads();
    // six partials per output: pairs first (36 lanes), then three (12 lanes)
    if (t < 36) {
        const uint32_t o = t / 3, q = t % 3;
        s.prod[t] = fp_add(s.part[6 * o + 2 * q], s.part[6 * o + 2 * q + 1]);
    }
    __syncthreads();
    if (t < 12) dst[t] = fp_add(fp_add(s.prod[3 * t], s.prod[3 * t + 1]), s.prod[3 * t + 2]);
    __syncthreads();
}

VKZG_DEV void fp12_set_one(Fp *a, uint32_t t) {
    if (t < 12) a[t] = t == 0 ? fp_one() : fp_zero();
}
// w -> -w: odd powers negated (the p^6-power Frobenius)
VKZG_DEV void fp12_conj_to(Fp *dst, const Fp *a, uint32_t t) {
    if (t < 12) dst[t] = ((t >> 1) & 1u) ? fp_neg(a[t]) : a[t];
}

}  // namespace vkzg_amd
