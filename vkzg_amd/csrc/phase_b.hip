// FK20 phase B buckets (split circulant, fixed scalars, variable points).
#include "proof_common.h"

namespace vkzg_amd {

namespace {
// Output a (< 128) of a blob belongs to segment 0 (a < 32: cyclic over the
// 32 points Fc), 1 (32 <= a < 64: negacyclic over the 32 points Fn) or 2
// (a >= 64: negacyclic over the 64 points F-); the ladder slots of the
// three are 0..31, 32..63 and 64..127 of the blob's ladder (k_ladder), laid
// out [d][slot]: the 64 lanes of a group read 64 consecutive points.  An item (e, d, sign) of segment s
// adds +-2^(8d) F[a - e] for d < 17 and +-2^(8(d-17)) phi(F[a - e]) above
// (the two GLV halves of the tap), with the index cyclic in the segment and
// the sign flipped when a negacyclic tap wraps (k_phase_b_split.comp:25-61).
struct PhaseBOut {
    uint32_t seg, aa, base;
    VKZG_DEV explicit PhaseBOut(uint32_t a)
        : seg(a < 32 ? 0u : (a < 64 ? 1u : 2u)), aa(a < 64 ? (a & 31u) : (a & 63u)), base(a < 64 ? (a & 32u) : 64u) {}
    // ladder slot and sign of tap e
    VKZG_DEV uint32_t slot(uint32_t e) const { return base + ((aa - e) & (seg == 2 ? 63u : 31u)); }
    VKZG_DEV bool wraps(uint32_t e) const { return seg != 0 && e > aa; }
};

// One bucket recomputed with the checked point operations (see k_phase_b).
__device__ __attribute__((noinline)) void phase_b_slow(G1X *out, const G1A *L, const G1A *Phi,
                                                       const uint32_t *items, uint32_t lo,
                                                       uint32_t hi, uint32_t a) {
    const PhaseBOut o(a);
    G1X acc = g1x_identity();
    for (uint32_t it = lo; it < hi; it++) {
        const uint32_t packed = items[it];
        const uint32_t e = packed & 0xffu, d = (packed >> 8) & 0xffu;
        bool neg = ((packed >> 16) & 1u) != 0;
        if (o.wraps(e)) neg = !neg;
        const uint32_t src = o.slot(e);
        const G1A pt = d < kLadPos ? L[d * 128 + src] : Phi[(d - kLadPos) * 128 + src];
        acc = g1x_madd(acc, g1a_neg_if(pt, neg));
    }
    *out = acc;
}

// Lane (blob, a, k): bucket k of output a.  The item list depends only on
// (segment, k), so work groups are formed per segment: a 64-lane group of
// segment 0 or 1 holds the 32 outputs of two blobs (pair bp: blobs 2bp and
// 2bp + 1), one of segment 2 the 64 outputs of one blob.  Groups
// [0, npairs) are segment 0, [npairs, 2 npairs) segment 1, the rest
// segment 2.
//
// Persistent: every wave of the `grid` blocks takes work items from the
// counter q[0] (cleared by k_glv_phi, the kernel before) until the queue is
// drained, and leaves after one failed take; waves share no barrier.  Item i
// is bucket k = i % 128 of group i / 128.  A flagged bucket is recomputed by
// k_phase_b_fix, which keeps this kernel free of calls.
__global__ void __launch_bounds__(256, 2)
    k_phase_b(G1X *buckets, uint8_t *flags, const G1A *lad_aff, const G1A *lad_phi, const uint32_t *items,
              const uint32_t *offsets, uint32_t nblobs, uint32_t *q, uint32_t force_checked) {
    const uint32_t lane = threadIdx.x % 64;
    const uint32_t npairs = (nblobs + 1) / 2;
    const uint32_t nitems = 128u * (2 * npairs + nblobs);
#pragma unroll 1
    for (;;) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(q, 1u);
        item = __shfl(item, 0);
        if (item >= nitems) break;
        const uint32_t k = item % 128;  // bucket
        const uint32_t grp = item / 128;
        uint32_t blob, a;
        if (grp < 2 * npairs) {
            blob = 2 * (grp % npairs) + (lane >> 5);
            a = (grp / npairs) * 32 + (lane & 31u);
        } else {
            blob = grp - 2 * npairs;
            a = 64 + lane;
        }
        const bool live = blob < nblobs;  // the odd blob count's last pair
        if (!live) blob = nblobs - 1;
        const PhaseBOut o(a);
        const uint32_t lo = offsets[o.seg * 129 + k], hi = offsets[o.seg * 129 + k + 1];
        const G1A *L = lad_aff + (size_t)blob * kLadPos * 128;
        const G1A *Phi = lad_phi + (size_t)blob * kLadPos * 128;
        G1X acc = g1x_identity();
        bool exc = force_checked != 0, started = false;
#pragma unroll 1
        for (uint32_t it = lo; it < hi; it++) {
            const uint32_t packed = items[it];
            const uint32_t e = packed & 0xffu;
            const uint32_t d = (packed >> 8) & 0xffu;
            bool neg = ((packed >> 16) & 1u) != 0;
            if (o.wraps(e)) neg = !neg;
            const uint32_t src = o.slot(e);
            const G1A pt = d < kLadPos ? L[d * 128 + src] : Phi[(d - kLadPos) * 128 + src];
            if (started) {
                acc = g1x_madd_acc(acc, pt, neg, exc);
            } else {
                acc = g1x_start(pt, neg, exc);
                started = true;
            }
        }
        if (live) {
            const size_t ba = (size_t)blob * 128 + a;
            buckets[ba * 128 + k] = acc;
            flags[ba * 128 + k] = exc ? 1 : 0;
        }
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
    const PhaseBOut o(a);
    const uint32_t lo = offsets[o.seg * 129 + k], hi = offsets[o.seg * 129 + k + 1];
    const G1A *L = lad_aff + (size_t)blob * kLadPos * 128;
    const G1A *Phi = lad_phi + (size_t)blob * kLadPos * 128;
    phase_b_slow(&buckets[idx], L, Phi, items, lo, hi, a);
}

}  // namespace

hipError_t launch_phase_b(G1X *buckets, uint8_t *flags, const G1A *lad_aff, const G1A *lad_phi,
                          const uint32_t *items, const uint32_t *offsets, uint32_t nblobs, uint32_t grid,
                          uint32_t *queue, uint32_t force_checked, hipStream_t s) {
    // 64-lane groups: two per pair of blobs (segments 0 and 1), one per blob
    const uint32_t ngroups = 2 * ((nblobs + 1) / 2) + nblobs;
    const uint32_t nitems = 128u * ngroups;
    if (grid > (nitems + 3) / 4) grid = (nitems + 3) / 4;
    hipLaunchKernelGGL(k_phase_b, dim3(grid), dim3(256), 0, s, buckets, flags, lad_aff, lad_phi, items,
                       offsets, nblobs, queue, force_checked);
    const uint32_t nbuckets = nblobs * 128 * 128;
    hipLaunchKernelGGL(k_phase_b_fix, dim3(blocks_for(nbuckets, 256)), dim3(256), 0, s, buckets, flags,
                       lad_aff, lad_phi, items, offsets, nbuckets);
    return hipGetLastError();
}

}  // namespace vkzg_amd
