// SHA-256 from FIPS 180-4: a portable compression function and one on the
// x86 SHA extensions (SHA256RNDS2 / SHA256MSG1 / SHA256MSG2), chosen at run
// time from CPUID.  Built with the host compiler (g++).
#include "sha256.h"

#include <cpuid.h>
#include <immintrin.h>

#include <cstring>

namespace vkzg_amd {
namespace sha256 {

namespace {

// FIPS 180-4 section 4.2.2: the first 32 bits of the fractional parts of the
// cube roots of the first 64 primes.
constexpr uint32_t K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

// Section 5.3.3: the first 32 bits of the fractional parts of the square
// roots of the first 8 primes.
constexpr uint32_t H0[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                            0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void compress_portable(uint32_t *h, const uint8_t *blocks, size_t nblocks) {
    for (size_t b = 0; b < nblocks; b++) {
        const uint8_t *p = blocks + 64 * b;
        uint32_t w[64];
        for (int t = 0; t < 16; t++)
            w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 | (uint32_t)p[4 * t + 2] << 8 |
                   p[4 * t + 3];
        for (int t = 16; t < 64; t++) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = h[0], bb = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int t = 0; t < 64; t++) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = hh + S1 + ch + K[t] + w[t];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t maj = (a & bb) ^ (a & c) ^ (bb & c);
            const uint32_t t2 = S0 + maj;
            hh = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = bb;
            bb = a;
            a = t1 + t2;
        }
        h[0] += a;
        h[1] += bb;
        h[2] += c;
        h[3] += d;
        h[4] += e;
        h[5] += f;
        h[6] += g;
        h[7] += hh;
    }
}

// The SHA extensions keep the state as two vectors (A, B, E, F) and
// (C, D, G, H); SHA256RNDS2 runs two rounds on them with the two message +
// constant words in the low half of its third operand.  The schedule
// W[t] = s1(W[t-2]) + W[t-7] + s0(W[t-15]) + W[t-16] is built four words at a
// time: MSG1 adds s0, an ALIGNR supplies W[t-7], MSG2 adds s1.
__attribute__((target("sha,sse4.1,ssse3"))) void compress_sha_ni(uint32_t *h, const uint8_t *blocks,
                                                                    size_t nblocks) {
    const __m128i bswap = _mm_set_epi64x(0x0c0d0e0f08090a0bLL, 0x0405060700010203LL);
    __m128i tmp = _mm_loadu_si128(reinterpret_cast<const __m128i *>(h));       // A B C D (lane 0 = A)
    __m128i st1 = _mm_loadu_si128(reinterpret_cast<const __m128i *>(h + 4));   // E F G H
    tmp = _mm_shuffle_epi32(tmp, 0xB1);                                         // B A D C
    st1 = _mm_shuffle_epi32(st1, 0x1B);                                         // H G F E
    __m128i st0 = _mm_alignr_epi8(tmp, st1, 8);                                 // F E B A = (A B E F)
    st1 = _mm_blend_epi16(st1, tmp, 0xF0);                                      // H G D C = (C D G H)
    for (size_t b = 0; b < nblocks; b++) {
        const uint8_t *p = blocks + 64 * b;
        const __m128i save0 = st0, save1 = st1;
        __m128i m[4];
        for (int g = 0; g < 4; g++)
            m[g] = _mm_shuffle_epi8(_mm_loadu_si128(reinterpret_cast<const __m128i *>(p + 16 * g)), bswap);
        for (int g = 0; g < 16; g++) {
            __m128i msg;
            if (g < 4) {
                msg = m[g];
            } else {
                // m[g & 3] holds W[4(g-4) ..]; build W[4g ..]
                __m128i t = _mm_sha256msg1_epu32(m[g & 3], m[(g + 1) & 3]);
                t = _mm_add_epi32(t, _mm_alignr_epi8(m[(g + 3) & 3], m[(g + 2) & 3], 4));
                msg = _mm_sha256msg2_epu32(t, m[(g + 3) & 3]);
                m[g & 3] = msg;
            }
            __m128i wk = _mm_add_epi32(msg, _mm_loadu_si128(reinterpret_cast<const __m128i *>(K + 4 * g)));
            st1 = _mm_sha256rnds2_epu32(st1, st0, wk);
            wk = _mm_shuffle_epi32(wk, 0x0E);
            st0 = _mm_sha256rnds2_epu32(st0, st1, wk);
        }
        st0 = _mm_add_epi32(st0, save0);
        st1 = _mm_add_epi32(st1, save1);
    }
    tmp = _mm_shuffle_epi32(st0, 0x1B);           // A B E F -> lanes F E B A reversed
    st1 = _mm_shuffle_epi32(st1, 0xB1);
    st0 = _mm_blend_epi16(tmp, st1, 0xF0);        // A B C D
    st1 = _mm_alignr_epi8(st1, tmp, 8);           // E F G H
    _mm_storeu_si128(reinterpret_cast<__m128i *>(h), st0);
    _mm_storeu_si128(reinterpret_cast<__m128i *>(h + 4), st1);
}

void compress(Ctx &c, const uint8_t *blocks, size_t nblocks) {
    if (c.hw) compress_sha_ni(c.h, blocks, nblocks);
    else compress_portable(c.h, blocks, nblocks);
}

}  // namespace

bool hw_available() {
    unsigned a = 0, b = 0, c = 0, d = 0;
    if (!__get_cpuid_count(7, 0, &a, &b, &c, &d)) return false;
    const bool sha = (b >> 29) & 1u;
    unsigned a1 = 0, b1 = 0, c1 = 0, d1 = 0;
    if (!__get_cpuid(1, &a1, &b1, &c1, &d1)) return false;
    const bool sse41 = (c1 >> 19) & 1u, ssse3 = (c1 >> 9) & 1u;
    return sha && sse41 && ssse3;
}

void init(Ctx &c, bool allow_hw) {
    std::memcpy(c.h, H0, sizeof(H0));
    c.buflen = 0;
    c.bytes = 0;
    static const bool hw = hw_available();
    c.hw = allow_hw && hw;
}

void update(Ctx &c, const void *data, size_t len) {
    const uint8_t *p = static_cast<const uint8_t *>(data);
    c.bytes += len;
    if (c.buflen) {
        const size_t take = len < 64 - c.buflen ? len : 64 - c.buflen;
        std::memcpy(c.buf + c.buflen, p, take);
        c.buflen += take;
        p += take;
        len -= take;
        if (c.buflen < 64) return;
        compress(c, c.buf, 1);
        c.buflen = 0;
    }
    if (len >= 64) {
        compress(c, p, len / 64);
        p += len & ~size_t(63);
        len &= 63;
    }
    std::memcpy(c.buf, p, len);
    c.buflen = len;
}

void finish(Ctx &c, uint8_t out[32]) {
    const uint64_t bits = c.bytes * 8;
    uint8_t pad[72] = {0x80};
    const size_t padlen = (c.buflen < 56 ? 56 : 120) - c.buflen;
    for (int i = 0; i < 8; i++) pad[padlen + i] = (uint8_t)(bits >> (56 - 8 * i));
    update(c, pad, padlen + 8);
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)(c.h[i] >> 24);
        out[4 * i + 1] = (uint8_t)(c.h[i] >> 16);
        out[4 * i + 2] = (uint8_t)(c.h[i] >> 8);
        out[4 * i + 3] = (uint8_t)c.h[i];
    }
}

}  // namespace sha256
}  // namespace vkzg_amd
