// BLS12-381 Fp and Fr Montgomery arithmetic for gfx950, held in VGPRs.
//
// Fr: 8 x 32-bit limbs, radix 2^256, canonical results.  Multiplication is
// Montgomery by product scanning (mont_mul below): each 32x32->64 limb
// product is one v_mad_u64_u32 whose carry-out feeds a v_addc_co_u32 into a
// 96-bit column accumulator; r's top limb has a clear top bit, so one
// conditional subtraction finishes.
//
// Fp: 14 x 28-bit limbs, radix 2^392, values in [0, 2p) (see the Fp section).
//
// Semantics follow the reference's field library (src/shaders/field.glsl.inc
// :69-297): the same field elements in Montgomery form; byte encodings are
// canonical big-endian.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "constants.h"
#include "inverse.h"

#define VKZG_DEV __device__ __forceinline__

namespace vkzg_amd {

struct Fr {
    uint32_t v[8];
};
struct Fp {
    uint32_t v[14];  // 28-bit limbs, see below
};

// ------------------------------------------------------------ limb helpers

// r = t - m when t >= m, else t.
template <int N>
VKZG_DEV void reduce_once(uint32_t *r, const uint32_t *t, const uint32_t (&m)[N]) {
    uint32_t d[N];
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint64_t s = (uint64_t)t[i] - m[i] - br;
        d[i] = (uint32_t)s;
        br = (uint32_t)(s >> 32) & 1u;
    }
#pragma unroll
    for (int i = 0; i < N; i++) r[i] = br ? t[i] : d[i];
}

template <int N>
VKZG_DEV void add_mod(uint32_t *r, const uint32_t *a, const uint32_t *b, const uint32_t (&m)[N]) {
    uint32_t t[N];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint64_t s = (uint64_t)a[i] + b[i] + c;
        t[i] = (uint32_t)s;
        c = (uint32_t)(s >> 32);
    }
    // a + b < 2m < 2^(32N): no carry out of the top limb.
    reduce_once<N>(r, t, m);
}

template <int N>
VKZG_DEV void sub_mod(uint32_t *r, const uint32_t *a, const uint32_t *b, const uint32_t (&m)[N]) {
    uint32_t t[N];
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint64_t s = (uint64_t)a[i] - b[i] - br;
        t[i] = (uint32_t)s;
        br = (uint32_t)(s >> 32) & 1u;
    }
    const uint32_t mask = 0u - br;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint64_t s = (uint64_t)t[i] + (m[i] & mask) + c;
        r[i] = (uint32_t)s;
        c = (uint32_t)(s >> 32);
    }
}

// acc (64-bit) += x * y and hi += carry-out: v_mad_u64_u32 leaves its carry
// in VCC and v_addc folds it into the third accumulator word, two VALU
// instructions per limb product (the compiler's own lowering of the same
// expression needs five to six).  `y` is an SGPR operand when it is a
// wave-uniform modulus limb.
VKZG_DEV void mac_vv(uint64_t &acc, uint32_t &hi, uint32_t x, uint32_t y) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\tv_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(hi)
        : "v"(x), "v"(y)
        : "vcc");
}
VKZG_DEV void mac_vs(uint64_t &acc, uint32_t &hi, uint32_t x, uint32_t y) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\tv_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(hi)
        : "v"(x), "s"(y)
        : "vcc");
}
// Both products of one FIPS step, a[j] * b[i - j] and m[j] * p[i - j], in one
// block: the compiler separates consecutive asm blocks with an s_nop.
VKZG_DEV void mac_pair(uint64_t &acc, uint32_t &hi, uint32_t x, uint32_t y, uint32_t mx,
                       uint32_t py) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc\n\t"
        "v_mad_u64_u32 %0, vcc, %4, %5, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(hi)
        : "v"(x), "v"(y), "v"(mx), "s"(py)
        : "vcc");
}

// Fr (mod r) carry chains written out.  The carry of v_add_co/v_addc and the
// borrow of v_sub_co/v_subb pass through VCC, and a VALU carry read of VCC
// needs no wait states; VCC read as a v_cndmask select does, so one s_nop
// (two wait states) sits before the selects.  A VALU instruction reads at
// most one scalar value (VCC, an SGPR or a literal), so r's limbs 2..7 enter
// the borrow chain from VGPRs; limbs 0 and 1 are the inline constants 1 and
// -1.  The compiler's own lowering of the same chains goes through 64-bit
// adds, about twice the instructions.
//
// r = t - r_mod when t >= r_mod, else t (t < 2^256).
VKZG_DEV void fr_reduce_once(uint32_t *r, const uint32_t *t) {
    uint32_t u[8];
    asm("v_subrev_co_u32_e32 %[u0], vcc, 1, %[t0]\n\t"
        "v_subbrev_co_u32_e32 %[u1], vcc, -1, %[t1], vcc\n\t"
        "v_subb_co_u32_e32 %[u2], vcc, %[t2], %[p2], vcc\n\t"
        "v_subb_co_u32_e32 %[u3], vcc, %[t3], %[p3], vcc\n\t"
        "v_subb_co_u32_e32 %[u4], vcc, %[t4], %[p4], vcc\n\t"
        "v_subb_co_u32_e32 %[u5], vcc, %[t5], %[p5], vcc\n\t"
        "v_subb_co_u32_e32 %[u6], vcc, %[t6], %[p6], vcc\n\t"
        "v_subb_co_u32_e32 %[u7], vcc, %[t7], %[p7], vcc\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e32 %[r0], %[u0], %[t0], vcc\n\t"
        "v_cndmask_b32_e32 %[r1], %[u1], %[t1], vcc\n\t"
        "v_cndmask_b32_e32 %[r2], %[u2], %[t2], vcc\n\t"
        "v_cndmask_b32_e32 %[r3], %[u3], %[t3], vcc\n\t"
        "v_cndmask_b32_e32 %[r4], %[u4], %[t4], vcc\n\t"
        "v_cndmask_b32_e32 %[r5], %[u5], %[t5], vcc\n\t"
        "v_cndmask_b32_e32 %[r6], %[u6], %[t6], vcc\n\t"
        "v_cndmask_b32_e32 %[r7], %[u7], %[t7], vcc"
        : [r0] "=&v"(r[0]), [r1] "=&v"(r[1]), [r2] "=&v"(r[2]), [r3] "=&v"(r[3]),
          [r4] "=&v"(r[4]), [r5] "=&v"(r[5]), [r6] "=&v"(r[6]), [r7] "=&v"(r[7]),
          [u0] "=&v"(u[0]), [u1] "=&v"(u[1]), [u2] "=&v"(u[2]), [u3] "=&v"(u[3]),
          [u4] "=&v"(u[4]), [u5] "=&v"(u[5]), [u6] "=&v"(u[6]), [u7] "=&v"(u[7])
        : [t0] "v"(t[0]), [t1] "v"(t[1]), [t2] "v"(t[2]), [t3] "v"(t[3]), [t4] "v"(t[4]),
          [t5] "v"(t[5]), [t6] "v"(t[6]), [t7] "v"(t[7]), [p2] "v"(cfr::P[2]),
          [p3] "v"(cfr::P[3]), [p4] "v"(cfr::P[4]), [p5] "v"(cfr::P[5]), [p6] "v"(cfr::P[6]),
          [p7] "v"(cfr::P[7])
        : "vcc");
}

// r = a + b mod r_mod for a, b < r_mod (a + b < 2^256: no carry out).
VKZG_DEV void fr_add_limbs(uint32_t *r, const uint32_t *a, const uint32_t *b) {
    uint32_t t[8];
    asm("v_add_co_u32_e32 %[t0], vcc, %[a0], %[b0]\n\t"
        "v_addc_co_u32_e32 %[t1], vcc, %[a1], %[b1], vcc\n\t"
        "v_addc_co_u32_e32 %[t2], vcc, %[a2], %[b2], vcc\n\t"
        "v_addc_co_u32_e32 %[t3], vcc, %[a3], %[b3], vcc\n\t"
        "v_addc_co_u32_e32 %[t4], vcc, %[a4], %[b4], vcc\n\t"
        "v_addc_co_u32_e32 %[t5], vcc, %[a5], %[b5], vcc\n\t"
        "v_addc_co_u32_e32 %[t6], vcc, %[a6], %[b6], vcc\n\t"
        "v_addc_co_u32_e32 %[t7], vcc, %[a7], %[b7], vcc"
        : [t0] "=&v"(t[0]), [t1] "=&v"(t[1]), [t2] "=&v"(t[2]), [t3] "=&v"(t[3]),
          [t4] "=&v"(t[4]), [t5] "=&v"(t[5]), [t6] "=&v"(t[6]), [t7] "=&v"(t[7])
        : [a0] "v"(a[0]), [a1] "v"(a[1]), [a2] "v"(a[2]), [a3] "v"(a[3]), [a4] "v"(a[4]),
          [a5] "v"(a[5]), [a6] "v"(a[6]), [a7] "v"(a[7]), [b0] "v"(b[0]), [b1] "v"(b[1]),
          [b2] "v"(b[2]), [b3] "v"(b[3]), [b4] "v"(b[4]), [b5] "v"(b[5]), [b6] "v"(b[6]),
          [b7] "v"(b[7])
        : "vcc");
    fr_reduce_once(r, t);
}

// r = a - b mod r_mod for a, b < r_mod: the difference, plus r_mod masked by
// the final borrow.
VKZG_DEV void fr_sub_limbs(uint32_t *r, const uint32_t *a, const uint32_t *b) {
    uint32_t t[8], mask, c;
    asm("v_sub_co_u32_e32 %[t0], vcc, %[a0], %[b0]\n\t"
        "v_subb_co_u32_e32 %[t1], vcc, %[a1], %[b1], vcc\n\t"
        "v_subb_co_u32_e32 %[t2], vcc, %[a2], %[b2], vcc\n\t"
        "v_subb_co_u32_e32 %[t3], vcc, %[a3], %[b3], vcc\n\t"
        "v_subb_co_u32_e32 %[t4], vcc, %[a4], %[b4], vcc\n\t"
        "v_subb_co_u32_e32 %[t5], vcc, %[a5], %[b5], vcc\n\t"
        "v_subb_co_u32_e32 %[t6], vcc, %[a6], %[b6], vcc\n\t"
        "v_subb_co_u32_e32 %[t7], vcc, %[a7], %[b7], vcc\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %[m], 0, -1, vcc"
        : [t0] "=&v"(t[0]), [t1] "=&v"(t[1]), [t2] "=&v"(t[2]), [t3] "=&v"(t[3]),
          [t4] "=&v"(t[4]), [t5] "=&v"(t[5]), [t6] "=&v"(t[6]), [t7] "=&v"(t[7]), [m] "=&v"(mask)
        : [a0] "v"(a[0]), [a1] "v"(a[1]), [a2] "v"(a[2]), [a3] "v"(a[3]), [a4] "v"(a[4]),
          [a5] "v"(a[5]), [a6] "v"(a[6]), [a7] "v"(a[7]), [b0] "v"(b[0]), [b1] "v"(b[1]),
          [b2] "v"(b[2]), [b3] "v"(b[3]), [b4] "v"(b[4]), [b5] "v"(b[5]), [b6] "v"(b[6]),
          [b7] "v"(b[7])
        : "vcc");
    asm("v_and_b32_e32 %[c], 1, %[m]\n\t"
        "v_add_co_u32_e32 %[r0], vcc, %[c], %[t0]\n\t"
        "v_addc_co_u32_e32 %[r1], vcc, %[m], %[t1], vcc\n\t"
        "v_and_b32_e32 %[c], 0xfffe5bfe, %[m]\n\t"
        "v_addc_co_u32_e32 %[r2], vcc, %[c], %[t2], vcc\n\t"
        "v_and_b32_e32 %[c], 0x53bda402, %[m]\n\t"
        "v_addc_co_u32_e32 %[r3], vcc, %[c], %[t3], vcc\n\t"
        "v_and_b32_e32 %[c], 0x09a1d805, %[m]\n\t"
        "v_addc_co_u32_e32 %[r4], vcc, %[c], %[t4], vcc\n\t"
        "v_and_b32_e32 %[c], 0x3339d808, %[m]\n\t"
        "v_addc_co_u32_e32 %[r5], vcc, %[c], %[t5], vcc\n\t"
        "v_and_b32_e32 %[c], 0x299d7d48, %[m]\n\t"
        "v_addc_co_u32_e32 %[r6], vcc, %[c], %[t6], vcc\n\t"
        "v_and_b32_e32 %[c], 0x73eda753, %[m]\n\t"
        "v_addc_co_u32_e32 %[r7], vcc, %[c], %[t7], vcc"
        : [r0] "=&v"(r[0]), [r1] "=&v"(r[1]), [r2] "=&v"(r[2]), [r3] "=&v"(r[3]),
          [r4] "=&v"(r[4]), [r5] "=&v"(r[5]), [r6] "=&v"(r[6]), [r7] "=&v"(r[7]), [c] "=&v"(c)
        : [t0] "v"(t[0]), [t1] "v"(t[1]), [t2] "v"(t[2]), [t3] "v"(t[3]), [t4] "v"(t[4]),
          [t5] "v"(t[5]), [t6] "v"(t[6]), [t7] "v"(t[7]), [m] "v"(mask)
        : "vcc");
}

// Montgomery multiplication by finely integrated product scanning (FIPS):
// column k of a*b and of m*p are summed together in a 96-bit accumulator,
// m_k = low word * n0 makes the low word vanish, and the accumulator shifts
// down one word per column.  Inputs below the modulus give a result below
// twice the modulus (the modulus' top bit is clear), reduced once.
template <int N>
VKZG_DEV void mont_mul(uint32_t *r, const uint32_t *a, const uint32_t *b, const uint32_t (&p)[N],
                       uint32_t n0) {
    uint32_t m[N];
    uint64_t acc = 0;
    uint32_t hi = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
        for (int j = 0; j < i; j++) mac_pair(acc, hi, a[j], b[i - j], m[j], p[i - j]);
        mac_vv(acc, hi, a[i], b[0]);
        m[i] = (uint32_t)acc * n0;
        mac_vs(acc, hi, m[i], p[0]);
        acc = (acc >> 32) | ((uint64_t)hi << 32);
        hi = 0;
    }
    uint32_t t[N];
#pragma unroll
    for (int i = N; i < 2 * N; i++) {
#pragma unroll
        for (int j = i - N + 1; j < N; j++) mac_pair(acc, hi, a[j], b[i - j], m[j], p[i - j]);
        t[i - N] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)hi << 32);
        hi = 0;
    }
    if constexpr (N == 8) {
        fr_reduce_once(r, t);
    } else {
        reduce_once<N>(r, t, p);
    }
}

// The same product by operand scanning (CIOS) in plain C++; the test hooks
// compute both forms and flag any disagreement.
template <int N>
VKZG_DEV void mont_mul_cios(uint32_t *r, const uint32_t *a, const uint32_t *b,
                            const uint32_t (&m)[N], uint32_t n0) {
    uint32_t t[N];
#pragma unroll
    for (int j = 0; j < N; j++) t[j] = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint32_t bi = b[i];
        uint64_t s = (uint64_t)a[0] * bi + t[0];
        t[0] = (uint32_t)s;
        uint32_t A = (uint32_t)(s >> 32);
        const uint32_t mm = t[0] * n0;
        uint64_t q = (uint64_t)mm * m[0] + t[0];
        uint32_t C = (uint32_t)(q >> 32);
#pragma unroll
        for (int j = 1; j < N; j++) {
            s = (uint64_t)a[j] * bi + t[j] + A;
            t[j] = (uint32_t)s;
            A = (uint32_t)(s >> 32);
            q = (uint64_t)mm * m[j] + t[j] + C;
            t[j - 1] = (uint32_t)q;
            C = (uint32_t)(q >> 32);
        }
        t[N - 1] = C + A;
    }
    reduce_once<N>(r, t, m);
}

template <int N>
VKZG_DEV bool limbs_zero(const uint32_t *a) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < N; i++) acc |= a[i];
    return acc == 0;
}

template <int N>
VKZG_DEV bool limbs_eq(const uint32_t *a, const uint32_t *b) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < N; i++) acc |= a[i] ^ b[i];
    return acc == 0;
}

// a < m as plain integers
template <int N>
VKZG_DEV bool limbs_lt(const uint32_t *a, const uint32_t (&m)[N]) {
    uint32_t br = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const uint64_t s = (uint64_t)a[i] - m[i] - br;
        br = (uint32_t)(s >> 32) & 1u;
    }
    return br != 0;
}

// ------------------------------------------------------------------ Fr

VKZG_DEV Fr fr_zero() {
    Fr r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    return r;
}
VKZG_DEV Fr fr_one() {
    Fr r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = cfr::ONE[i];
    return r;
}
VKZG_DEV Fr fr_const(const uint32_t (&c)[8]) {
    Fr r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = c[i];
    return r;
}
VKZG_DEV Fr fr_mul(const Fr &a, const Fr &b) {
    Fr r;
    mont_mul<8>(r.v, a.v, b.v, cfr::P, cfr::N0);
    return r;
}
VKZG_DEV Fr fr_sqr(const Fr &a) { return fr_mul(a, a); }
VKZG_DEV Fr fr_add(const Fr &a, const Fr &b) {
    Fr r;
    fr_add_limbs(r.v, a.v, b.v);
    return r;
}
VKZG_DEV Fr fr_sub(const Fr &a, const Fr &b) {
    Fr r;
    fr_sub_limbs(r.v, a.v, b.v);
    return r;
}
VKZG_DEV Fr fr_neg(const Fr &a) { return fr_sub(fr_zero(), a); }
VKZG_DEV bool fr_is_zero(const Fr &a) { return limbs_zero<8>(a.v); }

// Montgomery -> canonical (multiply by plain 1).
VKZG_DEV Fr fr_to_canonical(const Fr &a) {
    Fr one = fr_zero();
    one.v[0] = 1;
    return fr_mul(a, one);
}
// canonical (< r) -> Montgomery.
VKZG_DEV Fr fr_from_canonical(const Fr &a) { return fr_mul(a, fr_const(cfr::R2)); }
// Any 256-bit value -> Montgomery form of (value mod r).  2^256 < 3r, so at
// most two subtractions bring it below r first (k_cells_to_fr.comp:5-6 reduces
// non-canonical cell bytes mod r).
VKZG_DEV Fr fr_from_any(const Fr &a) {
    Fr t;
    fr_reduce_once(t.v, a.v);
    fr_reduce_once(t.v, t.v);
    return fr_from_canonical(t);
}

template <int NE>
VKZG_DEV Fr fr_pow_const(const Fr &a, const uint32_t (&e)[NE]) {
    Fr acc = fr_one();
#pragma unroll 1
    for (int i = NE - 1; i >= 0; i--) {
        const uint32_t limb = e[i];
#pragma unroll 1
        for (int b = 31; b >= 0; b--) {
            acc = fr_sqr(acc);
            if ((limb >> b) & 1u) acc = fr_mul(acc, a);
        }
    }
    return acc;
}
VKZG_DEV Fr fr_inv(const Fr &a) { return fr_pow_const<8>(a, cfr::P_MINUS_2); }

// 32 big-endian bytes (as 8 big-endian words) -> little-endian limbs
// A canonical a < r as sign and magnitude: r - a with bit 255 set when
// r - a < a, else a.  The magnitude is at most (r - 1) / 2 < 2^254; bit 255 is
// free since r < 2^255.
VKZG_DEV Fr fr_signed_magnitude(const Fr &a) {
    Fr n;
    uint32_t br = 0, lt = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint64_t d = (uint64_t)cfr::P[i] - a.v[i] - br;
        n.v[i] = (uint32_t)d;
        br = (uint32_t)(d >> 32) & 1u;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint64_t d = (uint64_t)n.v[i] - a.v[i] - lt;
        lt = (uint32_t)(d >> 32) & 1u;
    }
    if (!lt) return a;
    n.v[7] |= 0x80000000u;
    return n;
}
VKZG_DEV Fr fr_from_be_words(const uint32_t *w) {
    Fr r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = __builtin_bswap32(w[7 - i]);
    return r;
}
VKZG_DEV void fr_to_be_words(uint32_t *w, const Fr &a) {
#pragma unroll
    for (int i = 0; i < 8; i++) w[7 - i] = __builtin_bswap32(a.v[i]);
}

// ------------------------------------------------------------------ Fp
//
// Fp in registers: 14 limbs of 28 bits (one per 32-bit word), Montgomery
// form with radix R = 2^392, value kept in [0, 2p) ("almost reduced") with
// every limb below 2^28.
//  - Limb products are below 2^56, so one Montgomery column (at most 28
//    products plus the incoming carry) fits a single 64-bit accumulator: each
//    product is one v_mad_u64_u32 and no carry is propagated inside a column.
//  - 4p < R, so the Montgomery product of two values below 2p is below 2p
//    and needs no final subtraction.
//  - add/sub bring their result back to [0, 2p) with one conditional
//    correction by 2p; comparisons and encodings reduce to [0, p) first
//    (fp_canonical).

constexpr int kFpLimbs = 14;
constexpr uint32_t kFpMask = (1u << 28) - 1u;

VKZG_DEV Fp fp_zero() {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = 0;
    return r;
}
VKZG_DEV Fp fp_const(const uint32_t (&c)[kFpLimbs]) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = c[i];
    return r;
}
VKZG_DEV Fp fp_one() { return fp_const(cfp::ONE); }

// acc += x * y: the compiler emits one v_mad_u64_u32 (no carry-out needed,
// the accumulator cannot overflow).
VKZG_DEV void fp_mac(uint64_t &acc, uint32_t x, uint32_t y) { acc += (uint64_t)x * y; }

// Montgomery multiplication by finely integrated product scanning: column k
// of a*b and of m*p share the accumulator, m_k = low 28 bits * N0 clears the
// low limb of columns k < 14, and the accumulator shifts down 28 bits per
// column.  Inputs below 2p give a result below 2p (no final subtraction).
VKZG_DEV Fp fp_mul(const Fp &a, const Fp &b) {
    uint32_t m[kFpLimbs];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
#pragma unroll
        for (int j = 0; j < i; j++) {
            fp_mac(acc, a.v[j], b.v[i - j]);
            fp_mac(acc, m[j], cfp::P[i - j]);
        }
        fp_mac(acc, a.v[i], b.v[0]);
        m[i] = ((uint32_t)acc * cfp::N0) & kFpMask;
        fp_mac(acc, m[i], cfp::P[0]);
        acc >>= 28;
    }
    Fp r;
#pragma unroll
    for (int i = kFpLimbs; i < 2 * kFpLimbs; i++) {
#pragma unroll
        for (int j = i - kFpLimbs + 1; j < kFpLimbs; j++) {
            fp_mac(acc, a.v[j], b.v[i - j]);
            fp_mac(acc, m[j], cfp::P[i - j]);
        }
        r.v[i - kFpLimbs] = (uint32_t)acc & kFpMask;
        acc >>= 28;
    }
    return r;
}
// a^2: the off-diagonal products of a column are summed once and doubled
// (one v_lshl_add_u64), 105 limb products instead of 196.
VKZG_DEV Fp fp_sqr(const Fp &a) {
    uint32_t m[kFpLimbs];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
        uint64_t t = 0;
#pragma unroll
        for (int j = 0; 2 * j < i; j++) fp_mac(t, a.v[j], a.v[i - j]);
        acc += t << 1;
        if ((i & 1) == 0) fp_mac(acc, a.v[i / 2], a.v[i / 2]);
#pragma unroll
        for (int j = 0; j < i; j++) fp_mac(acc, m[j], cfp::P[i - j]);
        m[i] = ((uint32_t)acc * cfp::N0) & kFpMask;
        fp_mac(acc, m[i], cfp::P[0]);
        acc >>= 28;
    }
    Fp r;
#pragma unroll
    for (int i = kFpLimbs; i < 2 * kFpLimbs; i++) {
        uint64_t t = 0;
#pragma unroll
        for (int j = i - kFpLimbs + 1; 2 * j < i; j++) fp_mac(t, a.v[j], a.v[i - j]);
        acc += t << 1;
        if ((i & 1) == 0) fp_mac(acc, a.v[i / 2], a.v[i / 2]);
#pragma unroll
        for (int j = i - kFpLimbs + 1; j < kFpLimbs; j++) fp_mac(acc, m[j], cfp::P[i - j]);
        r.v[i - kFpLimbs] = (uint32_t)acc & kFpMask;
        acc >>= 28;
    }
    return r;
}

// Lazy forms, valid only as inputs of fp_mul / fp_sqr: limbs below 3 * 2^28
// and values below 6p.  A Montgomery column of two such inputs stays below
// 14 * 9 * 2^56 + 14 * 2^56 + carry < 2^63.2, and 36p < R keeps the product
// below 2p.  P4S is 4p with every limb but the top one in [2^28 - 1, 2^29),
// so subtracting a normalised value below 2p leaves no negative limb.
VKZG_DEV Fp fp_sub_lazy(const Fp &a, const Fp &b) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = a.v[i] + cfp::P4S[i] - b.v[i];
    return r;
}
// a + b and 3a limb by limb, for normalised inputs below 2p: limbs below
// 3 * 2^28 and values below 6p, the lazy form above.
VKZG_DEV Fp fp_add_lazy(const Fp &a, const Fp &b) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = a.v[i] + b.v[i];
    return r;
}
VKZG_DEV Fp fp_triple_lazy(const Fp &a) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = 3 * a.v[i];
    return r;
}
VKZG_DEV Fp fp_neg_lazy_if(const Fp &a, bool neg) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = neg ? cfp::P4S[i] - a.v[i] : a.v[i];
    return r;
}
// 4p - a limb by limb for a normalised a < 2p: limbs below 2^29, value in
// (2p, 4p], congruent to -a.
VKZG_DEV Fp fp_neg_lazy(const Fp &a) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = cfp::P4S[i] - a.v[i];
    return r;
}

// a*b + c*d with one Montgomery reduction: the two products and m*p share
// fp_mul's column accumulators, so the second product costs its 196 limb
// products and no reduction of its own.  Contract: a and b lazy (limbs below
// 3 * 2^28, values below 6p), c normalised (limbs below 2^28, value below
// 2p), d limbs below 2^29 and value below 4p.  A column then stays below
// 14 * (9 + 2 + 1) * 2^56 + carry < 2^63.4, and a*b + c*d < 44p^2 keeps the
// result below p + 44p^2/R < 2p.
VKZG_DEV Fp fp_mul_sum(const Fp &a, const Fp &b, const Fp &c, const Fp &d) {
    uint32_t m[kFpLimbs];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
#pragma unroll
        for (int j = 0; j < i; j++) {
            fp_mac(acc, a.v[j], b.v[i - j]);
            fp_mac(acc, c.v[j], d.v[i - j]);
            fp_mac(acc, m[j], cfp::P[i - j]);
        }
        fp_mac(acc, a.v[i], b.v[0]);
        fp_mac(acc, c.v[i], d.v[0]);
        m[i] = ((uint32_t)acc * cfp::N0) & kFpMask;
        fp_mac(acc, m[i], cfp::P[0]);
        acc >>= 28;
    }
    Fp r;
#pragma unroll
    for (int i = kFpLimbs; i < 2 * kFpLimbs; i++) {
#pragma unroll
        for (int j = i - kFpLimbs + 1; j < kFpLimbs; j++) {
            fp_mac(acc, a.v[j], b.v[i - j]);
            fp_mac(acc, c.v[j], d.v[i - j]);
            fp_mac(acc, m[j], cfp::P[i - j]);
        }
        r.v[i - kFpLimbs] = (uint32_t)acc & kFpMask;
        acc >>= 28;
    }
    return r;
}

// a + b in [0, 2p): s = a + b and t = s - 2p with their own signed carry
// chains; t when non-negative.
VKZG_DEV Fp fp_add(const Fp &a, const Fp &b) {
    Fp s, t;
    int32_t cs = 0, ct = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
        const int32_t x = (int32_t)(a.v[i] + b.v[i]) + cs;
        s.v[i] = (uint32_t)x & kFpMask;
        cs = x >> 28;
        const int32_t y = (int32_t)s.v[i] - (int32_t)cfp::P2[i] + ct;
        t.v[i] = (uint32_t)y & kFpMask;
        ct = y >> 28;
    }
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = ct < 0 ? s.v[i] : t.v[i];
    return r;
}
VKZG_DEV Fp fp_dbl(const Fp &a) { return fp_add(a, a); }

// a - b in [0, 2p): d = a - b (two's complement over 392 bits) and d + 2p;
// the latter when d is negative.
VKZG_DEV Fp fp_sub(const Fp &a, const Fp &b) {
    Fp d, e;
    int32_t cd = 0, ce = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
        const int32_t x = (int32_t)a.v[i] - (int32_t)b.v[i] + cd;
        d.v[i] = (uint32_t)x & kFpMask;
        cd = x >> 28;
        const int32_t y = (int32_t)d.v[i] + (int32_t)cfp::P2[i] + ce;
        e.v[i] = (uint32_t)y & kFpMask;
        ce = y >> 28;
    }
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = cd < 0 ? e.v[i] : d.v[i];
    return r;
}
VKZG_DEV Fp fp_neg(const Fp &a) { return fp_sub(fp_zero(), a); }

// [0, 2p) -> [0, p)
VKZG_DEV Fp fp_canonical(const Fp &a) {
    Fp t;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
        const int32_t y = (int32_t)a.v[i] - (int32_t)cfp::P[i] + c;
        t.v[i] = (uint32_t)y & kFpMask;
        c = y >> 28;
    }
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = c < 0 ? a.v[i] : t.v[i];
    return r;
}
// For a value in [0, 2p): zero iff its limbs are those of 0 or of p.
VKZG_DEV bool fp_is_zero(const Fp &a) {
    uint32_t z = 0, zp = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
        z |= a.v[i];
        zp |= a.v[i] ^ cfp::P[i];
    }
    return z == 0 || zp == 0;
}
VKZG_DEV bool fp_eq(const Fp &a, const Fp &b) {
    return limbs_eq<kFpLimbs>(fp_canonical(a).v, fp_canonical(b).v);
}
// c ? a : b limb by limb (v_cndmask), so a lane-varying choice between two
// register values never goes through an address.
VKZG_DEV Fp fp_select(bool c, const Fp &a, const Fp &b) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = c ? a.v[i] : b.v[i];
    return r;
}
VKZG_DEV Fp fp_cond_neg(const Fp &a, bool neg) {
    const Fp n = fp_neg(a);
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) r.v[i] = neg ? n.v[i] : a.v[i];
    return r;
}

// Montgomery -> canonical integer in [0, p).
VKZG_DEV Fp fp_to_canonical(const Fp &a) {
    Fp one = fp_zero();
    one.v[0] = 1;
    return fp_canonical(fp_mul(a, one));
}
// Integer below p (28-bit limbs) -> Montgomery form.
VKZG_DEV Fp fp_from_canonical(const Fp &a) { return fp_mul(a, fp_const(cfp::R2)); }

// 12 little-endian 32-bit words (a value below 2^384) <-> 28-bit limbs.
VKZG_DEV Fp fp_from_words(const uint32_t (&w)[12]) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
        const int bit = 28 * i, wi = bit >> 5, sh = bit & 31;
        uint32_t x = w[wi] >> sh;
        if (sh > 4 && wi + 1 < 12) x |= w[wi + 1] << (32 - sh);
        r.v[i] = x & kFpMask;
    }
    return r;
}
VKZG_DEV void fp_to_words(uint32_t (&w)[12], const Fp &a) {
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const int bit = 32 * k, li = bit / 28, sh = bit % 28;
        uint32_t x = a.v[li] >> sh;
        if (li + 1 < kFpLimbs) x |= a.v[li + 1] << (28 - sh);
        if (sh > 24 && li + 2 < kFpLimbs) x |= a.v[li + 2] << (56 - sh);
        w[k] = x;
    }
}
// a < p for a value in 28-bit limbs
VKZG_DEV bool fp_limbs_lt_p(const Fp &a) {
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) c = ((int32_t)a.v[i] - (int32_t)cfp::P[i] + c) >> 28;
    return c < 0;
}

template <int NE>
VKZG_DEV Fp fp_pow_const(const Fp &a, const uint32_t (&e)[NE]) {
    Fp acc = fp_one();
#pragma unroll 1
    for (int i = NE - 1; i >= 0; i--) {
        const uint32_t limb = e[i];
#pragma unroll 1
        for (int b = 31; b >= 0; b--) {
            acc = fp_sqr(acc);
            if ((limb >> b) & 1u) acc = fp_mul(acc, a);
        }
    }
    return acc;
}
VKZG_DEV Fp fp_inv_fermat(const Fp &a) { return fp_pow_const<12>(a, cfp::P_MINUS_2); }

// a^e for a constant exponent by a sliding window of 4 bits over the odd
// powers a, a^3, ..., a^15: the same squarings as fp_pow_const and about a
// fifth of its multiplications.
template <int NE>
VKZG_DEV Fp fp_pow_const_window(const Fp &a, const uint32_t (&e)[NE]) {
    Fp odd[8];
    odd[0] = a;
    const Fp a2 = fp_sqr(a);
#pragma unroll
    for (int k = 1; k < 8; k++) odd[k] = fp_mul(odd[k - 1], a2);
    Fp acc = fp_one();
    int i = NE * 32 - 1;
    while (i >= 0 && !((e[i / 32] >> (i % 32)) & 1u)) i--;
    bool started = false;
#pragma unroll 1
    while (i >= 0) {
        if (!((e[i / 32] >> (i % 32)) & 1u)) {
            acc = fp_sqr(acc);
            i--;
            continue;
        }
        // longest window of at most 4 bits ending in a set bit
        int j = i - 3 < 0 ? 0 : i - 3;
        while (!((e[j / 32] >> (j % 32)) & 1u)) j++;
        uint32_t w = 0;
        for (int k = i; k >= j; k--) w = (w << 1) | ((e[k / 32] >> (k % 32)) & 1u);
        if (started) {
            for (int k = i; k >= j; k--) acc = fp_sqr(acc);
            acc = fp_mul(acc, odd[w >> 1]);
        } else {
            acc = odd[w >> 1];
            started = true;
        }
        i = j - 1;
    }
    return acc;
}

// Inverse of a Montgomery value xR (0 -> 0): the binary GCD of inverse.h gives
// the plain inverse (xR)^-1 as a signed value in (-31p, 31p); adding 32p makes
// it non-negative and a Montgomery product by R^3 turns it into x^-1 R.
VKZG_DEV Fp fp_inv(const Fp &a) {
    if (fp_is_zero(a)) return fp_zero();
    int32_t v[kFpLimbs];
    if (!bgcd::inverse(v, a.v)) return fp_inv_fermat(a);
    Fp t;
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) {
        c += (int64_t)v[i] + cfp::P32[i];
        t.v[i] = (uint32_t)c & kFpMask;
        c >>= 28;
    }
    return fp_mul(t, fp_const(cfp::R3_M));
}

// canonical value > (p-1)/2 (the sign flag of the compressed encoding)
VKZG_DEV bool fp_is_lex_largest(const Fp &canonical) {
    // (p-1)/2 < x  <=>  HALF - x < 0
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++) c = ((int32_t)cfp::HALF[i] - (int32_t)canonical.v[i] + c) >> 28;
    return c < 0;
}

// The value lane k of this lane's quad holds (DPP quad_perm [k,k,k,k]), word
// by word; the quad's four lanes must all be active.
template <int K>
VKZG_DEV Fp fp_quad_bcast(const Fp &v) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++)
        r.v[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)v.v[i], K * 0x55, 0xf, 0xf, false);
    return r;
}

// The value lane k of this lane's pair (lanes 2i, 2i+1) holds (DPP quad_perm
// [0,0,2,2] or [1,1,3,3]); both lanes of the pair must be active.
template <int K>
VKZG_DEV Fp fp_pair_bcast(const Fp &v) {
    Fp r;
#pragma unroll
    for (int i = 0; i < kFpLimbs; i++)
        r.v[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)v.v[i], K ? 0xF5 : 0xA0, 0xf, 0xf, false);
    return r;
}

// Reference Montgomery product (operand scanning, plain C++) for the test
// hooks: same radix and limbs, fully reduced.
VKZG_DEV Fp fp_mul_reference(const Fp &a, const Fp &b) {
    uint64_t t[kFpLimbs + 2];
#pragma unroll
    for (int j = 0; j < kFpLimbs + 2; j++) t[j] = 0;
    for (int i = 0; i < kFpLimbs; i++) {
        for (int j = 0; j < kFpLimbs; j++) t[j] += (uint64_t)a.v[j] * b.v[i];
        const uint32_t m = ((uint32_t)t[0] * cfp::N0) & kFpMask;
        for (int j = 0; j < kFpLimbs; j++) t[j] += (uint64_t)m * cfp::P[j];
        // t[0] is now divisible by 2^28: shift the whole array one limb down
        uint64_t carry = t[0] >> 28;
        for (int j = 1; j <= kFpLimbs; j++) {
            const uint64_t v = t[j] + carry;
            t[j - 1] = v & kFpMask;
            carry = v >> 28;
        }
        t[kFpLimbs] = carry;
    }
    Fp r;
    for (int j = 0; j < kFpLimbs; j++) r.v[j] = (uint32_t)t[j];
    return fp_canonical(r);
}

}  // namespace vkzg_amd
