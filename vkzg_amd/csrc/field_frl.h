// Fr as 9 limbs of 29 bits in 32-bit words, Montgomery radix 2^261: the
// arithmetic of the radix-4 recovery transforms.
//
// - Products: a limb product is below 2^61 and a column of 9 + 9 products
//   plus the incoming carry stays below 2^64 (bounds below), so each product
//   is one v_mad_u64_u32 into a 64-bit accumulator with no carry-out
//   instruction: 162 of them and ~60 other instructions per product, against
//   128 v_mad_u64_u32 + 128 v_addc and the final subtraction of the 8 x 32-bit
//   form.  r = 1 mod 2^29, so m_i = -acc mod 2^29.
// - Sums and differences are lazy: limb by limb, no carries.  A difference
//   adds OFF4 = 4r (or OFF8 = 8r) written with limbs 0..7 in [2^29, 2^30)
//   (or [2^30, 2^31)), so no limb goes negative.  frl_norm propagates carries
//   without changing the value.
//
// Bounds (values as multiples of r; "normalised" = limbs below 2^29):
// - frl_mul(a, b): a's limbs below 3 * 2^30, b normalised.  A column is then
//   below 9 * 3 * 2^59 + 9 * 2^58 + 2^35 = 63 * 2^58 + 2^35 < 2^64.  With
//   a < A r and b < B r, A B < 70, the result is normalised and below
//   r + A B r^2 / 2^261 < 2r (r / 2^261 < 2^-6.14).
// - frl_sub4(a, b): b normalised with b < 2r (top limb at most OFF4's);
//   limbs of a - b grow by under 2^30, the value by 4r.  frl_sub8(a, b): b's
//   limbs below 2^30 and b < 4r; limbs grow by under 2^31, the value by 8r.
// The transforms' use of these (recover.hip, radix 4) keeps every product's
// first operand below 36r (64r for the data by 1/Z product, whose second
// operand is below 2r), and every limb below 3 * 2^30.
#pragma once

#include "field.h"

namespace vkzg_amd {

struct FrL {
    uint32_t v[9];
};

namespace cfrl {
constexpr uint32_t kMask = (1u << 29) - 1u;
constexpr uint32_t P[9] = {0x00000001u, 0x1ffffff8u, 0x1f96ffbfu, 0x1b4805ffu, 0x1d80553bu,
                           0x0c0404d0u, 0x1520cce7u, 0x0a6533afu, 0x0073eda7u};
constexpr uint32_t OFF4[9] = {0x20000004u, 0x3fffffdfu, 0x3e5bfefeu, 0x2d2017feu, 0x360154eeu,
                              0x30101342u, 0x3483339cu, 0x2994cebdu, 0x01cfb69cu};
constexpr uint32_t OFF8[9] = {0x40000008u, 0x5fffffbeu, 0x5cb7fdfdu, 0x5a402ffdu, 0x4c02a9dcu,
                              0x40202685u, 0x49066739u, 0x53299d7bu, 0x039f6d38u};
// Montgomery forms (x 2^261 mod r)
constexpr uint32_t SEVEN_POW64_M[9] = {0x127eef80u, 0x18bc375fu, 0x1bc9599bu, 0x15b992f4u, 0x1b44e988u,
                                       0x00b62cb2u, 0x1be8dd61u, 0x1aa5a869u, 0x00263153u};
constexpr uint32_t INV_SEVEN_POW8192_MINUS_1_M[9] = {0x0d4f86b8u, 0x0e2526d6u, 0x18645494u, 0x1116bbdbu,
                                                     0x014c330au, 0x1cbbdb75u, 0x1bb0daccu, 0x0f5f5d89u,
                                                     0x00305b29u};
// 64 / (7^8192 - 1): the three-kernel recovery's coset factor with the 64 of
// the cancelled per-cell DFT_64 pair
constexpr uint32_t SIXTY_FOUR_INV_SEVEN_POW8192_MINUS_1_M[9] = {
    0x13e1ade6u, 0x0949b66au, 0x03bf2b9cu, 0x005e5af1u, 0x14041a8eu,
    0x168e600au, 0x06e1e3b9u, 0x099022a0u, 0x0050a760u};
// 2^261 mod r: one in Montgomery form
constexpr uint32_t ONE_M[9] = {0x1fffffbau, 0x0000022fu, 0x1cb61180u, 0x0a4e5c00u, 0x0ee8b1a2u,
                               0x16e6aedfu, 0x1907f8bbu, 0x0853ddf7u, 0x004d043fu};
// 1: one product takes a Montgomery form to the plain value
constexpr uint32_t ONE_PLAIN[9] = {1u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
// 2^522 mod r: one product takes a plain value to Montgomery form
constexpr uint32_t R2[9] = {0x0a71b3c0u, 0x1d32207eu, 0x1663d999u, 0x1c5abc93u, 0x03b58c44u,
                            0x0be37438u, 0x0829f771u, 0x1660139eu, 0x0027fd91u};
// 2^266 mod r: one product takes x 2^256 (the 8 x 32-bit Montgomery form) to x 2^261
constexpr uint32_t FROM_R256[9] = {0x1ffff72bu, 0x000046a7u, 0x1f5f3540u, 0x0ce3021cu, 0x118f3661u,
                                   0x008176cbu, 0x054e487cu, 0x102e8190u, 0x001e092eu};
// 2^261 mod r in 8 x 32-bit limbs: fr_mul(x 2^256, TO_R261) = x 2^261
constexpr uint32_t TO_R261[8] = {0xffffffbau, 0x00000045u, 0x0072d846u, 0x1a25272eu,
                                 0x5dbeee8bu, 0xfe2eedcdu, 0x9eefbe41u, 0x4d043f42u};
}  // namespace cfrl

VKZG_DEV FrL frl_const(const uint32_t (&c)[9]) {
    FrL r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = c[i];
    return r;
}

VKZG_DEV FrL frl_mul(const FrL &a, const FrL &b) {
    uint32_t m[9];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
#pragma unroll
        for (int j = 0; j < i; j++) {
            fp_mac(acc, a.v[j], b.v[i - j]);
            fp_mac(acc, m[j], cfrl::P[i - j]);
        }
        fp_mac(acc, a.v[i], b.v[0]);
        m[i] = (0u - (uint32_t)acc) & cfrl::kMask;  // P[0] = 1
        acc += m[i];
        acc >>= 29;
    }
    FrL r;
#pragma unroll
    for (int i = 9; i < 18; i++) {
#pragma unroll
        for (int j = i - 8; j < 9; j++) {
            fp_mac(acc, a.v[j], b.v[i - j]);
            fp_mac(acc, m[j], cfrl::P[i - j]);
        }
        r.v[i - 9] = (uint32_t)acc & cfrl::kMask;
        acc >>= 29;
    }
    return r;
}
VKZG_DEV FrL frl_add(const FrL &a, const FrL &b) {
    FrL r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = a.v[i] + b.v[i];
    return r;
}
VKZG_DEV FrL frl_sub4(const FrL &a, const FrL &b) {
    FrL r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = a.v[i] + cfrl::OFF4[i] - b.v[i];
    return r;
}
VKZG_DEV FrL frl_sub8(const FrL &a, const FrL &b) {
    FrL r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = a.v[i] + cfrl::OFF8[i] - b.v[i];
    return r;
}
VKZG_DEV FrL frl_norm(const FrL &a) {
    FrL r = a;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        r.v[i + 1] += r.v[i] >> 29;
        r.v[i] &= cfrl::kMask;
    }
    return r;
}

// 8 little-endian 32-bit words (a value below 2^256) <-> normalised limbs.
VKZG_DEV FrL frl_unpack(const uint32_t (&w)[8]) {
    FrL r;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int bit = 29 * k, j = bit >> 5, s = bit & 31;
        const uint32_t lo = w[j], hi = j + 1 < 8 ? w[j + 1] : 0u;
        r.v[k] = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)s) & cfrl::kMask;
    }
    return r;
}
// a normalised and below 2^256
VKZG_DEV void frl_pack(uint32_t (&w)[8], const FrL &a) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int bit = 32 * j, k = bit / 29, s = bit % 29;
        uint32_t x = a.v[k] >> s;
        if (k + 1 < 9) x |= a.v[k + 1] << (29 - s);
        if (s > 26 && k + 2 < 9) x |= a.v[k + 2] << (58 - s);
        w[j] = x;
    }
}

VKZG_DEV FrL frl_shfl(const FrL &v, int src) {
    FrL r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = (uint32_t)__shfl((int)v.v[i], src);
    return r;
}
VKZG_DEV FrL frl_shfl_xor(const FrL &v, int m) {
    FrL r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.v[i] = (uint32_t)__shfl_xor((int)v.v[i], m);
    return r;
}

}  // namespace vkzg_amd
