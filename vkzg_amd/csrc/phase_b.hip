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
//
// Persistent: every wave of the `grid` blocks takes work items from the
// counter q[0] (cleared by k_glv_phi, the kernel before) until the queue is
// drained, and leaves after one failed take; waves share no barrier.  Item i
// is bucket k = i % 128 of the 64-lane group i / 128 of (blob, a).  A flagged
// bucket is recomputed by k_phase_b_fix, which keeps this kernel free of
// calls.
__global__ void __launch_bounds__(256, 2)
    k_phase_b(G1X *buckets, uint8_t *flags, const G1A *lad_aff, const G1A *lad_phi, const uint32_t *items,
              const uint32_t *offsets, uint32_t ngroups, uint32_t *q, uint32_t force_checked) {
    const uint32_t lane = threadIdx.x % 64;
    const uint32_t nitems = 128u * ngroups;
#pragma unroll 1
    for (;;) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(q, 1u);
        item = __shfl(item, 0);
        if (item >= nitems) break;
        const uint32_t k = item % 128;  // bucket
        const uint32_t ba = (item / 128) * 64 + lane;  // blob * 128 + a
        const uint32_t blob = ba / 128, a = ba % 128;
        const uint32_t half = a >> 6, aa = a & 63u;
        const uint32_t lo = offsets[half * 129 + k], hi = offsets[half * 129 + k + 1];
        const G1A *L = lad_aff + (size_t)(blob * 2 + half) * 64 * kLadPos;
        const G1A *Phi = lad_phi + (size_t)(blob * 2 + half) * 64 * kLadPos;
        G1X acc = g1x_identity();
        bool exc = force_checked != 0, started = false;
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
        buckets[(size_t)ba * 128 + k] = acc;
        flags[(size_t)ba * 128 + k] = exc ? 1 : 0;
    }
}

// Recompute, with the checked point operations, the buckets k_phase_b flagged
// (a doubling or cancellation inside the bucket; random data never has one).
__global__ void __launch_bounds__(256) k_phase_b_fix(G1X *buckets, const uint8_t *flags, const G1A *lad_aff, const G1A *lad_phi,
                              const uint32_t *items, const uint32_t *offsets, uint32_t nbuckets) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nbuckets || !flags[idx]) return;
    const uint32_t ba = idx / 128, k = idx % 128;
    const uint32_t blob = ba / 128, a = ba % 128;
    const uint32_t half = a >> 6, aa = a & 63u;
    const uint32_t lo = offsets[half * 129 + k], hi = offsets[half * 129 + k + 1];
    const G1A *L = lad_aff + (size_t)(blob * 2 + half) * 64 * kLadPos;
    const G1A *Phi = lad_phi + (size_t)(blob * 2 + half) * 64 * kLadPos;
    phase_b_slow(&buckets[idx], L, Phi, items, lo, hi, half, aa);
}

}  // namespace

hipError_t launch_phase_b(G1X *buckets, uint8_t *flags, const G1A *lad_aff, const G1A *lad_phi,
                          const uint32_t *items, const uint32_t *offsets, uint32_t nblobs, uint32_t grid,
                          uint32_t *queue, uint32_t force_checked, hipStream_t s) {
    const uint32_t ngroups = nblobs * 2;  // 64-lane groups of (blob, a)
    const uint32_t nitems = 128u * ngroups;
    if (grid > (nitems + 3) / 4) grid = (nitems + 3) / 4;
    hipLaunchKernelGGL(k_phase_b, dim3(grid), dim3(256), 0, s, buckets, flags, lad_aff, lad_phi, items,
                       offsets, ngroups, queue, force_checked);
    const uint32_t nbuckets = nblobs * 128 * 128;
    hipLaunchKernelGGL(k_phase_b_fix, dim3(blocks_for(nbuckets, 256)), dim3(256), 0, s, buckets, flags,
                       lad_aff, lad_phi, items, offsets, nbuckets);
    return hipGetLastError();
}

}  // namespace vkzg_amd
