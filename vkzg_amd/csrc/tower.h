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
// multiplication of latency instead of 54-144 in sequence), then cross-lane
// exchanges inside each wave (DPP quad broadcasts and an xor tree) combine
// them into the 12 result coefficients.  That keeps the pairing's long
// dependent chain (about 480 Fp12 products) near a millisecond on one CU,
// where a one-lane tower would take over ten.
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

VKZG_DEV Fp fp_shfl_xor(const Fp &v, int mask) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = (uint32_t)__shfl_xor((int)v.v[i], mask);
    return r;
}

constexpr uint32_t kFp12Threads = 192;  // lanes an Fp12 product uses (blockDim.x >= this)

// Limb-wise sum without carries (limbs of values below 2p stay below 2^31
// for up to 8 terms).
VKZG_DEV Fp fp_add_raw(const Fp &a, const Fp &b) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = a.v[i] + b.v[i];
    return r;
}
// Carries of a raw sum propagated: 28-bit limbs, same value.
VKZG_DEV Fp fp_normalize(const Fp &a) {
    Fp r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs - 1; i++) {
        const uint32_t x = a.v[i] + c;
        r.v[i] = x & kFpMask;
        c = x >> 28;
    }
    r.v[kFpLimbs - 1] = a.v[kFpLimbs - 1] + c;
    return r;
}
// A normalised value below 16p -> [0, 2p) (conditional subtractions of 8p, 4p, 2p).
VKZG_DEV Fp fp_reduce16(const Fp &a) {
    Fp r = a;
#pragma unroll
    for (int k = 3; k >= 1; k--) {
        Fp t;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < kFpLimbs; i++) {
            // subtract (2^k p) limb by limb: 2^k * P[i] fits 31 bits
            const int64_t y = (int64_t)r.v[i] - ((int64_t)cfp::P[i] << k) + c;
            t.v[i] = (uint32_t)y & kFpMask;
            c = (int32_t)(y >> 28);
        }
        if (c >= 0) r = t;
    }
    return r;
}

// dst = a * b for Fp12 values held in LDS (12 Fp each).  Every thread of the
// block calls it (it contains barriers); blockDim.x >= 192.  dst may alias a
// or b.  Lane t works on output power M = t / 32 and on term m1 = (t % 32) / 4
// (terms 6 and 7 of a group are empty): the quad's four lanes multiply
// a_{m1} and b_{m2}, m2 = M - m1 mod 6, component by component, DPP quad
// broadcasts give every lane the four products, lane 0 of the quad forms the
// real and lane 1 the imaginary part of a_{m1} b_{m2} (times xi when
// m1 + m2 >= 6: (x + y u) xi = (x - y) + (x + y) u), and an xor tree over the
// group's eight quads adds the six terms without carries, normalised once.
// The result is below 12p in 28-bit limbs: a valid Montgomery input (any
// input below 45p gives a product below 2p), but callers that add, negate or
// compare it reduce it first (fp12_reduce).  Operands and result stay in
// LDS; the partial sums never leave registers.
__device__ __forceinline__ void fp12_mul_lds(Fp *dst, const Fp *a, const Fp *b) {
    const uint32_t t = threadIdx.x;
    const uint32_t M = (t / 32) % 6, m1 = (t % 32) / 4, q = t & 3u;
    const bool term = t < kFp12Threads && m1 < 6;
    const uint32_t m1c = term ? m1 : 0;
    const uint32_t m2 = (M + 6 - m1c) % 6;
    const Fp x_in = a[2 * m1c + (q >> 1)];
    const Fp y_in = b[2 * m2 + (q & 1u)];
    __syncthreads();  // dst may alias a or b
    const Fp prod = fp_mul(x_in, y_in);
    const Fp p00 = fp_quad_bcast<0>(prod), p01 = fp_quad_bcast<1>(prod);
    const Fp p10 = fp_quad_bcast<2>(prod), p11 = fp_quad_bcast<3>(prod);
    const Fp x = fp_sub(p00, p11), y = fp_add(p01, p10);
    const bool wrap = m1c > M;
    Fp v = (q == 0) ? (wrap ? fp_sub(x, y) : x) : (wrap ? fp_add(x, y) : y);
    if (!term || q > 1) v = fp_zero();
    v = fp_add_raw(v, fp_shfl_xor(v, 4));
    v = fp_add_raw(v, fp_shfl_xor(v, 8));
    v = fp_add_raw(v, fp_shfl_xor(v, 16));
    if (t < kFp12Threads && (t % 32) < 2) dst[2 * M + (t % 32)] = fp_normalize(v);
    __syncthreads();
}

// The 12 coefficients of an Fp12 product (below 12p) brought to [0, 2p).
__device__ __forceinline__ void fp12_reduce(Fp *a) {
    if (threadIdx.x < 12) a[threadIdx.x] = fp_reduce16(a[threadIdx.x]);
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
