// GLV decomposition of a G1 scalar, host and device: k = k1 + LAMBDA k2 for a
// canonical k < r, with LAMBDA = z^2 - 1 the eigenvalue of phi(x, y) =
// (beta x, y) on G1.  Since r = LAMBDA^2 + LAMBDA + 1, the plain division
// k1 = k mod LAMBDA, k2 = k div LAMBDA gives k1 < LAMBDA < 2^128 and
// k2 <= LAMBDA + 1 < 2^128: two 128-bit scalars instead of one 255-bit one,
// which halves the doubling chain of a multi-scalar multiplication.
//
// The quotient is found by Barrett reduction (Menezes, van Oorschot and
// Vanstone, Handbook of Applied Cryptography, Algorithm 14.42, base 2,
// k = 128): q1 = k >> 127, q3 = (q1 * MU) >> 129 with MU = floor(2^256 /
// LAMBDA), then at most two corrections.
#pragma once

#include <cstdint>

#include "constants.h"

#ifndef VKZG_HD
#define VKZG_HD __host__ __device__ inline
#endif

namespace vkzg_amd {
namespace glv {

// r = a - b over n words; returns the borrow.
template <int N>
VKZG_HD uint32_t sub_words(uint32_t *r, const uint32_t *a, const uint32_t *b) {
    uint64_t borrow = 0;
    for (int i = 0; i < N; i++) {
        const uint64_t d = (uint64_t)a[i] - b[i] - borrow;
        r[i] = (uint32_t)d;
        borrow = (d >> 63) & 1u;
    }
    return (uint32_t)borrow;
}

// k: 8 little-endian words, k < r.  k1, k2: 4 words each.
VKZG_HD void split(uint32_t *k1, uint32_t *k2, const uint32_t *k) {
    uint32_t q1[4];
    for (int i = 0; i < 4; i++) q1[i] = (k[3 + i] >> 31) | (k[4 + i] << 1);
    uint32_t q2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        uint64_t carry = 0;
        for (int j = 0; j < 5; j++) {
            const uint64_t t = (uint64_t)q1[i] * cfr::GLV_MU[j] + q2[i + j] + carry;
            q2[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        q2[i + 5] = (uint32_t)carry;
    }
    uint32_t q[4];
    for (int i = 0; i < 4; i++) q[i] = (q2[4 + i] >> 1) | (q2[5 + i] << 31);
    // rem = k - q * LAMBDA (below 3 LAMBDA, so 5 words suffice)
    uint32_t ql[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        uint64_t carry = 0;
        for (int j = 0; j < 4; j++) {
            const uint64_t t = (uint64_t)q[i] * cfr::GLV_LAMBDA[j] + ql[i + j] + carry;
            ql[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        ql[i + 4] = (uint32_t)carry;
    }
    uint32_t rem[8];
    sub_words<8>(rem, k, ql);
    const uint32_t lam5[5] = {cfr::GLV_LAMBDA[0], cfr::GLV_LAMBDA[1], cfr::GLV_LAMBDA[2],
                              cfr::GLV_LAMBDA[3], 0};
    for (int pass = 0; pass < 2; pass++) {
        uint32_t t[5];
        if (sub_words<5>(t, rem, lam5)) break;  // rem < LAMBDA
        for (int i = 0; i < 5; i++) rem[i] = t[i];
        for (int i = 0; i < 4; i++)
            if (++q[i]) break;
    }
    for (int i = 0; i < 4; i++) {
        k1[i] = rem[i];
        k2[i] = q[i];
    }
}

}  // namespace glv
}  // namespace vkzg_amd
