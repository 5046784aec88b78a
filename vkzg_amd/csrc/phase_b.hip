// FK20 phase B buckets (split circulant, fixed scalars, variable points).
#include "proof_common.h"

namespace vkzg_amd {

namespace {
// One bucket recomputed with the checked point operations (see k_phase_b).
__device__ __attribute__((noinline)) void phase_b_slow(G1X *out, const G1A *L, const G1A *Phi,
                                                       const uint32_t *items, uint32_t lo,
                                                       uint32_t hi, uint32_t half, uint32_t aa) {
    G1X acc = g1x_identity();
    for (uint32_t it = lo; it < hi; it++) {
        const uint32_t packed = items[it];
        const uint32_t e = packed & 0xffu, d = (packed >> 8) & 0xffu;
        bool neg = ((packed >> 16) & 1u) != 0;
        if (half && e > aa) neg = !neg;
        const uint32_t src = ((aa - e) & 63u) * kLadPos;
        const G1A pt = d < kLadPos ? L[src + d] : Phi[src + d - kLadPos];
        acc = g1x_madd(acc, g1a_neg_if(pt, neg));
    }
    *out = acc;
}

// Lane (blob, a, k): bucket k of output a.  Output a < 64 is the cyclic half
// over L+, a >= 64 the negacyclic half over L- with a sign flip when a tap
// wraps (k_phase_b_split.comp:25-61).  An item (e, d, sign) adds
// +-2^(8d) L[aa-e] for d < 17 and +-2^(8(d-17)) phi(L[aa-e]) above (the two
// GLV halves of the scalar).  The item list depends only on (half, k), so it
// is uniform across a wave.
__global__ void __launch_bounds__(64, 2)
    k_phase_b(G1X *buckets, const G1A *lad_aff, const G1A *lad_phi, const uint32_t *items,
              const uint32_t *offsets) {
    const uint32_t k = blockIdx.x;                       // bucket, 0..127
    const uint32_t ba = blockIdx.y * 64 + threadIdx.x;   // blob * 128 + a
    const uint32_t blob = ba / 128, a = ba % 128;
    const uint32_t half = a >> 6, aa = a & 63u;
    const uint32_t lo = offsets[half * 129 + k], hi = offsets[half * 129 + k + 1];
    const G1A *L = lad_aff + (size_t)(blob * 2 + half) * 64 * kLadPos;
    const G1A *Phi = lad_phi + (size_t)(blob * 2 + half) * 64 * kLadPos;
    G1X acc = g1x_identity();
    bool exc = false, started = false;
#pragma unroll 1
    for (uint32_t it = lo; it < hi; it++) {
        const uint32_t packed = items[it];
        const uint32_t e = packed & 0xffu;
        const uint32_t d = (packed >> 8) & 0xffu;
        bool neg = ((packed >> 16) & 1u) != 0;
        if (half && e > aa) neg = !neg;
        const uint32_t src = ((aa - e) & 63u) * kLadPos;
        const G1A pt = d < kLadPos ? L[src + d] : Phi[src + d - kLadPos];
        if (started) {
            acc = g1x_madd_acc(acc, pt, neg, exc);
        } else {
            acc = g1x_start(pt, neg, exc);
            started = true;
        }
    }
    if (exc) phase_b_slow(&buckets[(size_t)ba * 128 + k], L, Phi, items, lo, hi, half, aa);
    else buckets[(size_t)ba * 128 + k] = acc;
}

}  // namespace

hipError_t launch_phase_b(G1X *buckets, const G1A *lad_aff, const G1A *lad_phi, const uint32_t *items,
                          const uint32_t *offsets, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_phase_b, dim3(128, nblobs * 2), dim3(64), 0, s, buckets, lad_aff, lad_phi,
                       items, offsets);
    return hipGetLastError();
}

}  // namespace vkzg_amd
