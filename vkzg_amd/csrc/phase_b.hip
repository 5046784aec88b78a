// FK20 phase B buckets (split circulant, fixed scalars, variable points).
#include "proof_common.h"

namespace vkzg_amd {

namespace {
// Output a (< 128) of a blob belongs to segment 0 (a < 32: cyclic over the
// 32 points Fc), 1 (32 <= a < 64: negacyclic over the 32 points Fn) or 2
// (a >= 64: negacyclic over the 64 points F-); the ladder slots of the
// three are 0..31, 32..63 and 64..127 of the blob's ladder (k_ladder), laid
// out [d][slot]: the 64 lanes of a group read 64 consecutive points.  An item (e, d, sign) of segment s
// adds +-2^(8d) F[a - e] for d < kLadPos and +-2^(8(d-kLadPos)) phi(F[a - e]) above
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

// The same for ladder points in XYZZ form (the small-batch path).
__device__ __attribute__((noinline)) void phase_b_slow_xyzz(G1X *out, const G1X *L, const G1X *Phi,
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
        const G1X pt = d < kLadPos ? L[d * 128 + src] : Phi[(d - kLadPos) * 128 + src];
        acc = g1x_add(acc, neg ? g1x_neg(pt) : pt);
    }
    *out = acc;
}

// Lane (blob, a, k): bucket k of output a.  The item list depends only on
// (segment, k), so work groups are formed per segment: a 64-lane group of
// segment 0 or 1 holds the 32 outputs of two blobs (blobs 2p and 2p + 1 of
// pair p), one of segment 2 the 64 outputs of one blob.  A work item is one
// (group, bucket).
//
// Work order: the blobs go in bands of kBand; within a band, class by class
// (a class is a (segment, bucket) pair, longest item list first: the host's
// order in offsets[]), every group of the band takes the class's bucket.
// Bucket lengths vary several-fold (magnitude 1 also takes the top
// positions' carries), and a long bucket taken last would run alone.  Class
// c starts at item start(c) = n2(c) bb + (c - n2(c)) pb of its band, with
// bb the band's blobs, pb its pairs and n2(c) the segment-2 classes before c.
//
// Persistent: every wave of the `grid` blocks takes work items from the
// counter q[0] (cleared by k_glv_phi, the kernel before) until the queue is
// drained, and leaves after one failed take; waves share no barrier.  A
// flagged bucket is recomputed by k_phase_b_fix, which keeps this kernel
// free of calls.
//
// Q = 4 (batches too small to fill the SIMDs, where the longest bucket's
// chain sets the time): a quad of lanes shares each bucket
// (g1x_madd_acc_quad), and a work item is a quarter of a group, 16 of its 64
// buckets.  XZ (with Q = 4): the ladder points come in XYZZ form
// (k_ladder_xyzz) and the quad adds them whole (g1x_add_quad, also four
// rounds), so the batch needs no normalisation.
constexpr uint32_t kBand = 32;       // blobs per band (even)
constexpr uint32_t kClasses = 3 * kBuckets;
constexpr uint32_t top_pow2(uint32_t n) { return n < 2 ? 1u : 2u * top_pow2(n / 2); }
constexpr uint32_t kSearchTop = top_pow2(kClasses);  // largest power of two <= kClasses

template <uint32_t Q, bool XZ>
__global__ void __launch_bounds__(256, 2)
    k_phase_b(G1X *buckets, uint8_t *flags, const void *lad_pts, const void *lad_phi, const uint32_t *items,
              const uint32_t *offsets, uint32_t nblobs, uint32_t *q, uint32_t force_checked) {
    const uint32_t plane = threadIdx.x % 64;
    const uint32_t *cls = offsets + 3 * (kBuckets + 1);
    const uint32_t *n2 = cls + kClasses;
    const uint32_t band_items = kBuckets * 2 * kBand;  // a full band: kBand / 2 pairs x 2 + kBand
    const uint32_t nfull = nblobs / kBand, rest = nblobs % kBand;
    const uint32_t nitems = (nfull * band_items + kBuckets * (2 * ((rest + 1) / 2) + rest)) * Q;
#pragma unroll 1
    for (;;) {
        uint32_t qitem = 0;
        if (plane == 0) qitem = atomicAdd(q, 1u);
        qitem = __shfl(qitem, 0);
        if (qitem >= nitems) break;
        // the group's lane (bucket) this lane works on
        const uint32_t item = qitem / Q, lane = (qitem % Q) * (64 / Q) + plane / Q;
        const uint32_t band = item / band_items, j = item % band_items;
        const uint32_t b0 = band * kBand;
        const uint32_t bb = band < nfull ? kBand : rest, pb = (bb + 1) / 2;
        // the last class c with start(c) <= j
        uint32_t c = 0;
#pragma unroll 1
        for (uint32_t step = kSearchTop; step; step >>= 1) {
            const uint32_t t = c + step;
            if (t < kClasses && n2[t] * bb + (t - n2[t]) * pb <= j) c = t;
        }
        const uint32_t g = j - (n2[c] * bb + (c - n2[c]) * pb);
        const uint32_t seg = cls[c] >> 16, k = cls[c] & 0xffffu;
        uint32_t blob, a;
        if (seg == 2) {
            blob = b0 + g;
            a = 64 + lane;
        } else {
            blob = b0 + 2 * g + (lane >> 5);
            a = seg * 32 + (lane & 31u);
        }
        const bool live = blob < nblobs;  // the odd blob count's last pair
        if (!live) blob = nblobs - 1;
        const PhaseBOut o(a);
        const uint32_t lo = offsets[o.seg * (kBuckets + 1) + k], hi = offsets[o.seg * (kBuckets + 1) + k + 1];
        G1X acc = g1x_identity();
        bool exc = force_checked != 0, started = false;
        if constexpr (XZ) {
            const G1X *L = static_cast<const G1X *>(lad_pts) + (size_t)blob * kLadPos * 128;
            const G1X *Phi = static_cast<const G1X *>(lad_phi) + (size_t)blob * kLadPos * 128;
#pragma unroll 1
            for (uint32_t it = lo; it < hi; it++) {
                const uint32_t packed = items[it];
                const uint32_t e = packed & 0xffu;
                const uint32_t d = (packed >> 8) & 0xffu;
                bool neg = ((packed >> 16) & 1u) != 0;
                if (o.wraps(e)) neg = !neg;
                const uint32_t src = o.slot(e);
                G1X pt = d < kLadPos ? L[d * 128 + src] : Phi[(d - kLadPos) * 128 + src];
                if (neg) pt = g1x_neg(pt);
                if (started) {
                    acc = g1x_add_quad(acc, pt, plane & 3u, exc);
                } else {
                    acc = pt;
                    started = true;
                }
            }
        } else {
            const G1A *L = static_cast<const G1A *>(lad_pts) + (size_t)blob * kLadPos * 128;
            const G1A *Phi = static_cast<const G1A *>(lad_phi) + (size_t)blob * kLadPos * 128;
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
                    if constexpr (Q == 4) acc = g1x_madd_acc_quad(acc, pt, neg, plane & 3u, exc);
                    else acc = g1x_madd_acc(acc, pt, neg, exc);
                } else {
                    acc = g1x_start(pt, neg, exc);
                    started = true;
                }
            }
        }
        if (live && plane % Q == 0) {
            const size_t ba = (size_t)blob * 128 + a;
            buckets[ba * kBuckets + k] = acc;
            flags[ba * kBuckets + k] = exc ? 1 : 0;
        }
    }
}

// Recompute, with the checked point operations, the buckets k_phase_b flagged
// (a doubling or cancellation inside the bucket; random data never has one).
template <bool XZ>
__global__ void __launch_bounds__(256) k_phase_b_fix(G1X *buckets, const uint8_t *flags, const void *lad_pts,
                                                     const void *lad_phi, const uint32_t *items,
                                                     const uint32_t *offsets, uint32_t nbuckets) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nbuckets || !flags[idx]) return;
    const uint32_t ba = idx / kBuckets, k = idx % kBuckets;
    const uint32_t blob = ba / 128, a = ba % 128;
    const PhaseBOut o(a);
    const uint32_t lo = offsets[o.seg * (kBuckets + 1) + k], hi = offsets[o.seg * (kBuckets + 1) + k + 1];
    const size_t base = (size_t)blob * kLadPos * 128;
    if constexpr (XZ)
        phase_b_slow_xyzz(&buckets[idx], static_cast<const G1X *>(lad_pts) + base,
                          static_cast<const G1X *>(lad_phi) + base, items, lo, hi, a);
    else
        phase_b_slow(&buckets[idx], static_cast<const G1A *>(lad_pts) + base,
                     static_cast<const G1A *>(lad_phi) + base, items, lo, hi, a);
}

}  // namespace

hipError_t launch_phase_b(G1X *buckets, uint8_t *flags, const void *lad_pts, const void *lad_phi,
                          const uint32_t *items, const uint32_t *offsets, uint32_t nblobs, uint32_t grid,
                          bool quad, bool xyzz, uint32_t *queue, uint32_t force_checked, hipStream_t s) {
    if (xyzz && !quad) return hipErrorInvalidValue;
    // 64-lane groups: two per pair of blobs of a band (segments 0 and 1), one per blob
    const uint32_t rest = nblobs % kBand;
    const uint32_t ngroups = (nblobs / kBand) * kBuckets * 2 * kBand + kBuckets * (2 * ((rest + 1) / 2) + rest);
    const uint32_t nitems = ngroups * (quad ? 4u : 1u);  // one wave each
    if (grid > (nitems + 3) / 4) grid = (nitems + 3) / 4;
    if (xyzz)
        hipLaunchKernelGGL((k_phase_b<4, true>), dim3(grid), dim3(256), 0, s, buckets, flags, lad_pts, lad_phi,
                           items, offsets, nblobs, queue, force_checked);
    else if (quad)
        hipLaunchKernelGGL((k_phase_b<4, false>), dim3(grid), dim3(256), 0, s, buckets, flags, lad_pts, lad_phi,
                           items, offsets, nblobs, queue, force_checked);
    else
        hipLaunchKernelGGL((k_phase_b<1, false>), dim3(grid), dim3(256), 0, s, buckets, flags, lad_pts, lad_phi,
                           items, offsets, nblobs, queue, force_checked);
    const uint32_t nbuckets = nblobs * 128 * kBuckets;
    if (xyzz)
        hipLaunchKernelGGL(k_phase_b_fix<true>, dim3(blocks_for(nbuckets, 256)), dim3(256), 0, s, buckets, flags,
                           lad_pts, lad_phi, items, offsets, nbuckets);
    else
        hipLaunchKernelGGL(k_phase_b_fix<false>, dim3(blocks_for(nbuckets, 256)), dim3(256), 0, s, buckets, flags,
                           lad_pts, lad_phi, items, offsets, nbuckets);
    return hipGetLastError();
}

}  // namespace vkzg_amd
