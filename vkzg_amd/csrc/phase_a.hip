// FK20 phase A: 128 fixed-base MSMs of length 64 per blob (see kernels.h).
#include "proof_common.h"

namespace vkzg_amd {

namespace {
// c-bit window `w` of a canonical 256-bit scalar.
VKZG_DEV uint32_t window_bits(const Fr &s, uint32_t w, uint32_t c) {
    const uint32_t off = w * c;
    const uint32_t li = off >> 5;
    const uint32_t sh = off & 31u;
    uint32_t lo = li < 8 ? s.v[li] >> sh : 0u;
    if (sh && li + 1 < 8) lo |= s.v[li + 1] << (32u - sh);
    return lo & ((1u << c) - 1u);
}

constexpr int kMsmLanes = 16;  // lanes per phase-A MSM (64 terms / 16 = 4 each)

// One MSM recomputed with the checked point operations: reached only when the
// fast path flagged a doubling or cancellation case for this MSM.
__device__ __attribute__((noinline)) void phase_a_slow(G1X *out, const Fr *scal, const G1A *tab_j,
                                                       uint32_t c, uint32_t W) {
    const uint32_t H = 1u << (c - 1);
    G1X acc = g1x_identity();
    for (uint32_t i = 0; i < 64; i++) {
        const Fr s = scal[i];
        const G1A *tab = tab_j + (size_t)i * W * H;
        uint32_t carry = 0;
        for (uint32_t w = 0; w < W; w++) {
            const uint32_t raw = window_bits(s, w, c) + carry;
            carry = raw > H ? 1u : 0u;
            const uint32_t mag = carry ? (2u * H - raw) : raw;
            if (mag) acc = g1x_madd(acc, g1a_neg_if(tab[(size_t)w * H + (mag - 1)], carry != 0));
        }
    }
    *out = acc;
}

// u[blob][j] = sum_{i<64} scal[blob][j][i] * xext[j][i], with signed c-bit
// digits v in [-(2^(c-1)-1), 2^(c-1)] (the recoding of k_phase_a_sort.comp:
// 31-40 with a c-bit window) and table[j][i][w][|v|-1] = |v| * 2^(cw) * xext[j][i].
__global__ void __launch_bounds__(256, 2)
    k_phase_a(G1X *u, const Fr *scal, const G1A *table, uint32_t c, uint32_t W, uint32_t nmsm) {
    const uint32_t msm = blockIdx.x * (256 / kMsmLanes) + threadIdx.x / kMsmLanes;
    const uint32_t t = threadIdx.x % kMsmLanes;
    const bool live = msm < nmsm;
    const uint32_t j = msm % 128;
    const uint32_t H = 1u << (c - 1);
    G1X acc = g1x_identity();
    bool exc = false;
    if (live) {
#pragma unroll 1
        for (uint32_t q = 0; q < 64 / kMsmLanes; q++) {
            const uint32_t i = t * (64 / kMsmLanes) + q;
            const Fr s = scal[(size_t)msm * 64 + i];
            const G1A *tab = table + ((size_t)(j * 64 + i) * W) * H;
            uint32_t carry = 0;
#pragma unroll 1
            for (uint32_t w = 0; w < W; w++) {
                uint32_t raw = window_bits(s, w, c) + carry;
                carry = raw > H ? 1u : 0u;
                const bool neg = carry != 0;
                const uint32_t mag = neg ? (2u * H - raw) : raw;
                if (mag == 0) continue;
                const G1A pt = g1a_neg_if(tab[(size_t)w * H + (mag - 1)], neg);
                acc = g1x_madd_fast(acc, pt, exc);
            }
        }
    }
    // Sum the 16 partial results of this MSM.
#pragma unroll 1
    for (int m = 1; m < kMsmLanes; m <<= 1) acc = g1x_add_fast(acc, g1x_shfl_xor(acc, m), exc);
    const uint64_t flagged = __ballot(exc);
    const uint32_t group = (threadIdx.x & 63u) / kMsmLanes;
    const bool redo = ((flagged >> (group * kMsmLanes)) & ((1ull << kMsmLanes) - 1)) != 0;
    if (live && t == 0) {
        if (redo) phase_a_slow(&u[msm], scal + (size_t)msm * 64, table + (size_t)j * 64 * W * H, c, W);
        else u[msm] = acc;
    }
}

}  // namespace

hipError_t launch_phase_a(G1X *u, const Fr *scal, const G1A *table, uint32_t c, uint32_t W,
                          uint32_t nblobs, hipStream_t s) {
    const uint32_t nmsm = nblobs * 128;
    hipLaunchKernelGGL(k_phase_a, dim3(blocks_for(nmsm, 256 / kMsmLanes)), dim3(256), 0, s, u, scal,
                       table, c, W, nmsm);
    return hipGetLastError();
}

}  // namespace vkzg_amd
