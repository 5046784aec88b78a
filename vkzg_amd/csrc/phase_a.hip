// FK20 phase A: 128 fixed-base MSMs of length 64 per blob (see kernels.h).
#include "proof_common.h"

namespace vkzg_amd {

namespace {
// The `width` bits of s from bit `off` (bits above 255 read as zero).
VKZG_DEV uint32_t window_bits(const Fr &s, uint32_t off, uint32_t width) {
    const uint32_t li = off >> 5;
    const uint32_t sh = off & 31u;
    uint32_t lo = li < 8 ? s.v[li] >> sh : 0u;
    if (sh && li + 1 < 8) lo |= s.v[li + 1] << (32u - sh);
    return lo & ((1u << width) - 1u);
}

// A phase-A scalar's magnitude (bit 255 cleared) and its sign.
VKZG_DEV Fr scalar_magnitude(const Fr &s, bool &neg) {
    Fr m = s;
    neg = (s.v[7] >> 31) != 0;
    m.v[7] &= 0x7fffffffu;
    return m;
}

constexpr int kMsmLanes = 8;  // lanes per tile (64 terms / 8 = 8 each)
constexpr int kTilesPerItem = 256 / kMsmLanes;  // tiles one 256-thread block takes per queue item

// Windows [w0, w1) of the 64 terms of one MSM, recomputed with the checked
// point operations: reached only when the fast path flagged a doubling or
// cancellation case in this tile.
__device__ __attribute__((noinline)) void phase_a_slow(G1X *out, const Fr *scal, const G1AT *tab_j,
                                                       WindowPlan plan, uint32_t w0, uint32_t w1) {
    G1X acc = g1x_identity();
    for (uint32_t i = 0; i < 64; i++) {
        bool sneg;
        const Fr s = scalar_magnitude(scal[i], sneg);
        const G1AT *tab = tab_j + (size_t)i * plan.entries();
        uint32_t carry = 0;
        for (uint32_t w = 0; w < w1; w++) {
            const uint32_t Hw = plan.H() << plan.wide(w);
            const uint32_t raw = window_bits(s, plan.bit(w), plan.c + plan.wide(w)) + carry;
            carry = raw > Hw ? 1u : 0u;
            const uint32_t mag = carry ? (2u * Hw - raw) : raw;
            if (w >= w0 && mag)
                acc = g1x_madd(acc, g1a_neg_if(tab[plan.entry(w) + (mag - 1)].p, (carry != 0) != sneg));
        }
    }
    *out = acc;
}

// Persistent phase A.  The work is a queue of tiles (msm, g): windows
// [g W / G, (g+1) W / G) of MSM msm = (blob * splits + m) * 128 + j, where
//   u[msm] = sum_{i<64} scal[msm][i] * xext[j][i]
// with the scalar as sign and magnitude (k_fk20_scalars) and the magnitude in
// signed digits v in [-(2^(b-1)-1), 2^(b-1)] of each window's width b (the
// recoding of k_phase_a_sort.comp:31-40 with the plan's windows), each digit's
// sign flipped for a negative scalar, and
// table[j][i][entry(w) + |v| - 1] = |v| * 2^bit(w) * xext[j][i].  Tile (msm, g) writes its
// partial sum to part[msm * G + g]; k_sum_groups adds the G partials.  Tiles
// are numbered group-major (tile = g * nmsm + msm): nmsm is a multiple of 128,
// so the 8 tiles of a wave share g and run the same windows -- mixed groups
// in one wave would make it step through the union of their windows.  A
// block takes 32 consecutive tiles per queue item (8 lanes per tile, 8 terms
// per lane) from the counter q[0] until the queue is drained; every block
// exits after one failed take, so the grid always drains.  q[0] is cleared by
// the kernel that runs before this one.
__global__ void __launch_bounds__(256, 2)
    k_phase_a(G1X *part, const Fr *scal, const G1AT *table, WindowPlan plan, uint32_t nmsm,
              uint32_t groups, uint32_t *q, uint32_t force_checked) {
    __shared__ uint32_t item_sh;
    const uint32_t t = threadIdx.x % kMsmLanes;
    const uint32_t W = plan.W, H = plan.H(), narrow = plan.W - plan.k;
    const uint32_t ntiles = nmsm * groups;
    const uint32_t nitems = (ntiles + kTilesPerItem - 1) / kTilesPerItem;
#pragma unroll 1
    for (;;) {
        if (threadIdx.x == 0) item_sh = atomicAdd(q, 1u);
        __syncthreads();
        const uint32_t item = item_sh;
        __syncthreads();
        if (item >= nitems) break;
        const uint32_t tile = item * kTilesPerItem + threadIdx.x / kMsmLanes;
        const bool live = tile < ntiles;
        const uint32_t g = live ? tile / nmsm : 0u;
        const uint32_t msm = live ? tile % nmsm : 0u;
        const uint32_t j = msm % 128;
        const uint32_t w0 = g * W / groups, w1 = (g + 1) * W / groups;
        G1X acc = g1x_identity();
        bool exc = force_checked != 0, started = false;
        if (live) {
#pragma unroll 1
            for (uint32_t qq = 0; qq < 64 / kMsmLanes; qq++) {
                const uint32_t i = t * (64 / kMsmLanes) + qq;
                bool sneg;
                const Fr s = scalar_magnitude(scal[(size_t)msm * 64 + i], sneg);
                const G1AT *tab = table + (size_t)(j * 64 + i) * plan.entries();
                uint32_t carry = 0;
#pragma unroll 1
                for (uint32_t w = 0; w < w1; w++) {
                    // windows at or above `narrow` are one bit wider
                    const uint32_t wide = w >= narrow ? 1u : 0u, below = w > narrow ? w - narrow : 0u;
                    const uint32_t Hw = H << wide;
                    const uint32_t raw = window_bits(s, w * plan.c + below, plan.c + wide) + carry;
                    carry = raw > Hw ? 1u : 0u;
                    const bool neg = (carry != 0) != sneg;
                    const uint32_t mag = carry ? (2u * Hw - raw) : raw;
                    if (w < w0 || mag == 0) continue;
                    const G1A pt = tab[(w + below) * H + (mag - 1)].p;
                    if (started) {
                        acc = g1x_madd_acc(acc, pt, neg, exc);
                    } else {
                        acc = g1x_start(pt, neg, exc);
                        started = true;
                    }
                }
            }
        }
        // Sum the 8 lanes' partial results of this tile.
#pragma unroll 1
        for (int m = 1; m < kMsmLanes; m <<= 1) acc = g1x_add_fast(acc, g1x_shfl_xor(acc, m), exc);
        const uint64_t flagged = __ballot(exc);
        const uint32_t sub = (threadIdx.x & 63u) / kMsmLanes;
        const bool redo = ((flagged >> (sub * kMsmLanes)) & ((1ull << kMsmLanes) - 1)) != 0;
        if (live && t == 0) {
            G1X *out = &part[(size_t)msm * groups + g];
            if (redo)
                phase_a_slow(out, scal + (size_t)msm * 64, table + (size_t)j * 64 * plan.entries(), plan, w0, w1);
            else
                *out = acc;
        }
    }
}

}  // namespace

hipError_t launch_phase_a(G1X *part, const Fr *scal, const G1AT *table, WindowPlan plan,
                          uint32_t nmsm, uint32_t groups, uint32_t grid, uint32_t *queue,
                          uint32_t force_checked, hipStream_t s) {
    const uint32_t nitems = (nmsm * groups + kTilesPerItem - 1) / kTilesPerItem;
    if (grid > nitems) grid = nitems;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_phase_a, dim3(grid), dim3(256), 0, s, part, scal, table, plan, nmsm, groups,
                       queue, force_checked);
    return hipGetLastError();
}

}  // namespace vkzg_amd
