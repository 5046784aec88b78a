// Fp inversion by a binary GCD on 28-bit approximations (the "optimized
// binary GCD" of T. Pornin, 2020, eprint 2020/972), for values in 14 x 28-bit
// limbs.  Fermat's p - 2 power needs ~570 dependent multiplications; this
// runs a fixed 30 outer iterations of ~1,100 integer instructions each, with
// no data-dependent branches, so every lane of a wave follows the same path.
//
// Invariants (a, b >= 0, u and v signed, all mod p):  a = u*y,  b = v*y.
// Start a = y, b = p, u = 1, v = 0.  Each outer iteration
//  - approximates a and b by their low 28 bits and their top 30 bits,
//  - runs 28 binary-GCD steps on the approximations (if a is odd: swap when
//    a < b, then a -= b; then a /= 2), recording the update as a matrix
//    [[f0, g0], [f1, g1]] with |f| + |g| <= 2^28,
//  - applies it exactly: a <- (a f0 + b g0) / 2^28, b <- (a f1 + b g1) / 2^28
//    (the low 28 bits are exact, so the divisions are exact), negating a row
//    when its result is negative,
//  - and u <- (u f0 + v g0) / 2^28 mod p, v likewise, the division by 2^28
//    done Montgomery-style (add q p with q = -w p^-1 mod 2^28).
// 28 outer iterations reach a = 0, b = gcd = 1 for every y (2 * 381 - 1 binary
// steps suffice); two more are harmless.  Then v = y^-1 mod p, held as a
// signed value of magnitude below 31p (u and v grow by less than p per
// iteration and are never reduced).
//
// Signed multi-limb values: limbs 0..12 in [0, 2^28), limb 13 signed.
// This header is plain C++ apart from the function qualifier, so the host
// test in tests/c can compile it with g++.
#pragma once

#include <cstdint>

#include "constants.h"

#ifndef VKZG_HD
#define VKZG_HD __host__ __device__ inline
#endif

namespace vkzg_amd {
namespace bgcd {

constexpr int kLimbs = 14;
constexpr uint32_t kMask = (1u << 28) - 1u;
constexpr int kOuter = 30;

// r = (x f + y g) / 2^28 for non-negative x, y < 2^382 (exact division);
// returns true when the result is negative.
VKZG_HD bool lin_shift(uint32_t (&r)[kLimbs], const uint32_t (&x)[kLimbs], const uint32_t (&y)[kLimbs],
                       int32_t f, int32_t g) {
    int64_t c = (int64_t)x[0] * f + (int64_t)y[0] * g;
    c >>= 28;  // the low 28 bits are zero
    for (int i = 1; i < kLimbs; i++) {
        c += (int64_t)x[i] * f + (int64_t)y[i] * g;
        r[i - 1] = (uint32_t)c & kMask;
        c >>= 28;
    }
    r[kLimbs - 1] = (uint32_t)(int32_t)c;  // signed top limb
    return c < 0;
}

// r = -r for a value with a signed top limb; the result is non-negative.
VKZG_HD void negate(uint32_t (&r)[kLimbs]) {
    int64_t c = 0;
    for (int i = 0; i < kLimbs - 1; i++) {
        c -= (int64_t)r[i];
        r[i] = (uint32_t)c & kMask;
        c >>= 28;
    }
    r[kLimbs - 1] = (uint32_t)(int32_t)(c - (int64_t)(int32_t)r[kLimbs - 1]);
}

// r = (u f + v g + q p) / 2^28 with q = -(u f + v g) p^-1 mod 2^28, for signed
// u, v (limbs 0..12 unsigned, limb 13 signed).
VKZG_HD void lin_shift_mod(int32_t (&r)[kLimbs], const int32_t (&u)[kLimbs], const int32_t (&v)[kLimbs],
                           int32_t f, int32_t g) {
    int64_t c = (int64_t)u[0] * f + (int64_t)v[0] * g;
    const uint32_t q = ((uint32_t)c * cfp::N0) & kMask;
    c += (int64_t)q * cfp::P[0];
    c >>= 28;
    for (int i = 1; i < kLimbs; i++) {
        c += (int64_t)u[i] * f + (int64_t)v[i] * g + (int64_t)q * cfp::P[i];
        r[i - 1] = (int32_t)((uint32_t)c & kMask);
        c >>= 28;
    }
    r[kLimbs - 1] = (int32_t)c;
}

// Low 28 bits and bits [s, s + 30) of x, as a 58-bit approximation.
VKZG_HD uint64_t approx(const uint32_t (&x)[kLimbs], int li, int sh) {
    uint32_t x0 = 0, x1 = 0, x2 = 0;
    for (int i = 0; i < kLimbs; i++) {
        const bool at = i == li;
        x0 = at ? x[i] : x0;
        x1 = at ? (i + 1 < kLimbs ? x[i + 1] : 0u) : x1;
        x2 = at ? (i + 2 < kLimbs ? x[i + 2] : 0u) : x2;
    }
    uint32_t top = x0 >> sh;
    top |= x1 << (28 - sh);
    if (sh > 26) top |= x2 << (56 - sh);
    top &= (1u << 30) - 1u;
    return (uint64_t)(x[0] & kMask) | ((uint64_t)top << 28);
}

// v = y^-1 mod p (signed, |v| < 31p) for y in [1, 2^382) with y != 0 mod p.
// Returns false if the iteration bound was not enough (never observed).
VKZG_HD bool inverse(int32_t (&v)[kLimbs], const uint32_t (&y)[kLimbs]) {
    uint32_t a[kLimbs], b[kLimbs];
    int32_t u[kLimbs];
    for (int i = 0; i < kLimbs; i++) {
        a[i] = y[i];
        b[i] = cfp::P[i];
        u[i] = 0;
        v[i] = 0;
    }
    u[0] = 1;
    for (int it = 0; it < kOuter; it++) {
        // n = max(len(a), len(b), 58); approximation window s = n - 30
        int hi = 0;
        uint32_t topv = 0;
        for (int i = 0; i < kLimbs; i++) {
            const uint32_t t = a[i] | b[i];
            if (t) {
                hi = i;
                topv = t;
            }
        }
        int len = 28 * hi + (topv ? 32 - __builtin_clz(topv) : 0);
        if (len < 58) len = 58;
        const int s = len - 30;
        const int li = s / 28, sh = s - 28 * li;
        uint64_t ah = approx(a, li, sh), bh = approx(b, li, sh);
        int32_t f0 = 1, g0 = 0, f1 = 0, g1 = 1;
        for (int j = 0; j < 28; j++) {
            const bool odd = (ah & 1u) != 0;
            const bool sw = odd && ah < bh;
            const uint64_t ta = ah;
            ah = sw ? bh : ah;
            bh = sw ? ta : bh;
            const int32_t tf = f0, tg = g0;
            f0 = sw ? f1 : f0;
            f1 = sw ? tf : f1;
            g0 = sw ? g1 : g0;
            g1 = sw ? tg : g1;
            ah -= odd ? bh : 0u;
            f0 -= odd ? f1 : 0;
            g0 -= odd ? g1 : 0;
            ah >>= 1;
            f1 *= 2;
            g1 *= 2;
        }
        uint32_t na[kLimbs], nb[kLimbs];
        if (lin_shift(na, a, b, f0, g0)) {
            negate(na);
            f0 = -f0;
            g0 = -g0;
        }
        if (lin_shift(nb, a, b, f1, g1)) {
            negate(nb);
            f1 = -f1;
            g1 = -g1;
        }
        int32_t nu[kLimbs], nv[kLimbs];
        lin_shift_mod(nu, u, v, f0, g0);
        lin_shift_mod(nv, u, v, f1, g1);
        for (int i = 0; i < kLimbs; i++) {
            a[i] = na[i];
            b[i] = nb[i];
            u[i] = nu[i];
            v[i] = nv[i];
        }
    }
    uint32_t bad = b[0] ^ 1u;
    for (int i = 0; i < kLimbs; i++) bad |= a[i] | (i ? b[i] : 0u);
    return bad == 0;
}

}  // namespace bgcd
}  // namespace vkzg_amd
