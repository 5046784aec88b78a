// FK20 phase B inputs: doubling ladders of the split-circulant folds of u.
#include "proof_common.h"

namespace vkzg_amd {

namespace {
// g1j_dbl (dbl-2009-l) shared by the two lanes of a pair, which both hold p.
// Its seven products fall into four dependent rounds of at most two -- {X^2,
// Y^2}, {B^2, (X+B)^2}, {E^2} (computed by both lanes) and {E(D - X3), YZ} --
// and DPP exchanges hand each round's results to both lanes: a doubling waits
// on three squarings and one product instead of five and two.  The formulas
// and operand order are g1j_dbl's; the sums that only feed products (X + B,
// E = 3A, D - X3) stay lazy, so coordinates are the same field elements (the
// proofs, canonical, are unchanged).
VKZG_DEV G1J g1j_dbl_pair(const G1J &p, uint32_t q) {
    if (fp_is_zero(p.z)) return p;  // uniform in the pair
    const Fp r1 = fp_sqr(fp_select(q != 0, p.y, p.x));
    const Fp A = fp_pair_bcast<0>(r1), B = fp_pair_bcast<1>(r1);
    const Fp r2 = fp_sqr(fp_select(q != 0, fp_add_lazy(p.x, B), B));
    const Fp C = fp_pair_bcast<0>(r2), T = fp_pair_bcast<1>(r2);
    const Fp E = fp_triple_lazy(A);  // feeds products only
    const Fp F = fp_sqr(E);
    const Fp D = fp_dbl(fp_sub(fp_sub(T, A), C));
    G1J r;
    r.x = fp_sub(F, fp_dbl(D));
    const Fp C8 = fp_dbl(fp_dbl(fp_dbl(C)));
    const Fp r4 = fp_mul(fp_select(q != 0, p.y, E), fp_select(q != 0, p.z, fp_sub_lazy(D, r.x)));
    r.y = fp_sub(fp_pair_bcast<0>(r4), C8);
    r.z = fp_dbl(fp_pair_bcast<1>(r4));
    return r;
}

// The same doubling on the four lanes of a quad: E^2 needs only A, so it
// joins round two on a third lane, and a doubling waits on two squarings and
// one product -- {X^2, Y^2}, {B^2, (X+B)^2, E^2}, {E(D - X3), YZ} -- at twice
// the lanes per chain.  Operands and lazy sums are g1j_dbl_pair's, so the
// results are the same field elements.  For batches whose pair ladder leaves
// SIMDs idle: there the chain's length, not the issue rate, sets the time.
VKZG_DEV G1J g1j_dbl_quad(const G1J &p, uint32_t q) {
    if (fp_is_zero(p.z)) return p;  // uniform in the quad
    const bool odd = (q & 1u) != 0;
    const Fp r1 = fp_sqr(fp_select(odd, p.y, p.x));
    const Fp A = fp_quad_bcast<0>(r1), B = fp_quad_bcast<1>(r1);
    const Fp E = fp_triple_lazy(A);  // feeds products only
    const Fp r2 = fp_sqr(fp_select(q == 2, E, fp_select(q == 0, B, fp_add_lazy(p.x, B))));
    const Fp C = fp_quad_bcast<0>(r2), T = fp_quad_bcast<1>(r2), F = fp_quad_bcast<2>(r2);
    const Fp D = fp_dbl(fp_sub(fp_sub(T, A), C));
    G1J r;
    r.x = fp_sub(F, fp_dbl(D));
    const Fp C8 = fp_dbl(fp_dbl(fp_dbl(C)));
    const Fp r4 = fp_mul(fp_select(odd, p.y, E), fp_select(odd, p.z, fp_sub_lazy(D, r.x)));
    r.y = fp_sub(fp_quad_bcast<0>(r4), C8);
    r.z = fp_dbl(fp_quad_bcast<1>(r4));
    return r;
}

// The phase-B points are linear in u, and so are the doublings, so the
// split-circulant folds (k_fold_ladder.comp) are taken before the ladder:
//   F+[i] = u[i] + u[i+64], F-[i] = u[i] - u[i+64] (i < 64),
//   Fc[i] = F+[i] + F+[i+32], Fn[i] = F+[i] - F+[i+32] (i < 32),
// and chain j of a blob doubles f[j] = Fc[j] (j < 32), Fn[j-32] (j < 64),
// F-[j-64] (j >= 64): lad[blob][d][j] = 2^(kDigitBits d) * f[j], d < kLadPos.
// With `splits` = M segments, pair (blob, m, j) starts from the same folds of
// phase A's 2^(kDigitBits start_m) u (u[(blob*M + m)*128 + k], phase A's MSMs with
// their window groups summed), start_m = lad_start(M, m), and writes
// positions [start_m, start_{m+1}).  LANES (2 or 4) lanes share a chain
// (g1j_dbl_pair or g1j_dbl_quad) and exit together.
template <int LANES>
__global__ __launch_bounds__(256) void k_ladder(G1J *lad, const G1X *u, uint32_t splits, uint32_t nblobs) {
    // A long dependent chain: give these waves issue priority over the
    // throughput kernels of another sub-batch sharing the SIMD.
    __builtin_amdgcn_s_setprio(3);
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = tid / LANES, q = tid % LANES;
    if (g >= nblobs * 128 * splits) return;
    const uint32_t j = g % 128, m = (g / 128) % splits, blob = g / (128 * splits);
    const size_t base = (size_t)(blob * splits + m) * 128;
    G1X f;
    if (j < 64) {
        const uint32_t i = j & 31u;
        const G1X lo = g1x_add(u[base + i], u[base + i + 64]);
        const G1X hi = g1x_add(u[base + i + 32], u[base + i + 96]);
        f = g1x_add(lo, j < 32 ? hi : g1x_neg(hi));
    } else {
        const uint32_t i = j - 64;
        f = g1x_add(u[base + i], g1x_neg(u[base + i + 64]));
    }
    G1J p = g1j_from_xyzz(f);
    const uint32_t d0 = lad_start(splits, m), d1 = m + 1 == splits ? kLadPos : lad_start(splits, m + 1);
    G1J *dst = lad + (size_t)blob * kLadPos * 128 + j;
#pragma unroll 1
    for (uint32_t d = d0; d < d1; d++) {
        if (q == 0) {
            dst[d * 128].x = p.x;
            dst[d * 128].z = p.z;
        } else {
            dst[d * 128].y = p.y;
        }
        if (d + 1 == d1) break;
#pragma unroll 1
        for (uint32_t k = 0; k < kDigitBits; k++) p = LANES == 4 ? g1j_dbl_quad(p, q) : g1j_dbl_pair(p, q);
    }
}

// The window-group partials of phase A summed per MSM:
// usum[msm] = sum_g part[msm * G + g].  Eight partial sums per MSM: sum t
// adds partials t, t + 8 and t + 16, then a three-level tree.  Q = 1: a lane
// per sum, with the checked additions; Q = 4 (small batches, where this
// chain of additions is the kernel's time): a quad per sum sharing each
// addition (g1x_add_quad, four rounds of products instead of fourteen), any
// flagged doubling or cancellation redone checked by the quad's lane 0.
// With G > 1 the ladder reads these sums, so its prologue (a dependent chain
// per group of lanes) adds no partials of its own.  nmsm is a multiple of
// 128, so the grid holds whole MSMs and no lane leaves before the shuffles.
VKZG_DEV_COLD void sum_groups_checked(G1X *out, const G1X *part, uint32_t groups) {
    G1X acc = g1x_identity();
    for (uint32_t g = 0; g < groups; g++) acc = g1x_add(acc, part[g]);
    *out = acc;
}

template <uint32_t Q>
__global__ __launch_bounds__(256) void k_sum_groups(G1X *usum, const G1X *part, uint32_t groups, uint32_t nmsm,
                                                    uint32_t force_checked) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t msm = tid / (8 * Q), t = (tid / Q) % 8, qd = tid % Q;
    if (msm >= nmsm) return;
    G1X acc = g1x_identity();
    if constexpr (Q == 4) {
        bool exc = force_checked != 0;
#pragma unroll 1
        for (uint32_t g = t; g < groups; g += 8) acc = g1x_add_quad(acc, part[(size_t)msm * groups + g], qd, exc);
#pragma unroll 1
        for (uint32_t m = 1; m < 8; m <<= 1) acc = g1x_add_quad(acc, g1x_shfl_xor(acc, m * Q), qd, exc);
        // a flag anywhere in the MSM's 32 lanes: lane 0 redoes the sum
        const uint64_t flagged = __ballot(exc);
        const uint32_t sub = (threadIdx.x & 63u) / 32u;
        if (((flagged >> (32u * sub)) & 0xffffffffull) != 0) {
            if (t == 0 && qd == 0) sum_groups_checked(&usum[msm], part + (size_t)msm * groups, groups);
            return;
        }
    } else {
#pragma unroll 1
        for (uint32_t g = t; g < groups; g += 8) acc = g1x_add(acc, part[(size_t)msm * groups + g]);
#pragma unroll 1
        for (uint32_t m = 1; m < 8; m <<= 1) acc = g1x_add(acc, g1x_shfl_xor(acc, m));
    }
    if (t == 0 && qd == 0) usum[msm] = acc;
}

// phi[k] = (beta x, y) of pts[k]: the endomorphism [lambda], lambda = z^2 - 1
// (the identity (0, 0) maps to itself).
__global__ void k_glv_phi(G1A *phi, const G1A *pts, uint64_t count, uint32_t *queue_b, uint32_t *queue_r) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {  // work counters of phase B and the bucket reduction
        queue_b[0] = 0u;
        queue_r[0] = 0u;
    }
    if (k >= count) return;
    G1A q = pts[k];
    q.x = fp_mul(q.x, fp_const(cfp::BETA_M));
    phi[k] = q;
}

// Small batches: the ladder points in XYZZ form for phase B's quad
// additions instead of normalised ones -- ZZ = Z^2, ZZZ = Z^3 keep X and Y
// (Z = 0 gives the XYZZ identity) -- and their images under the
// endomorphism, (beta X, Y, ZZ, ZZZ).  Two products per point, no inversion.
// Clears the phase-B and bucket-reduction work counters, as k_glv_phi does.
__global__ void k_ladder_xyzz(G1X *pts, G1X *phi, const G1J *lad, uint64_t count, uint32_t *queue_b,
                              uint32_t *queue_r) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        queue_b[0] = 0u;
        queue_r[0] = 0u;
    }
    if (k >= count) return;
    const G1J p = lad[k];
    G1X r;
    r.x = p.x;
    r.y = p.y;
    r.zz = fp_sqr(p.z);
    r.zzz = fp_mul(r.zz, p.z);
    pts[k] = r;
    r.x = fp_mul(p.x, fp_const(cfp::BETA_M));
    phi[k] = r;
}

}  // namespace

hipError_t launch_ladder_xyzz(G1X *pts, G1X *phi, const G1J *lad, uint64_t count, uint32_t *queue_b,
                              uint32_t *queue_r, hipStream_t s) {
    hipLaunchKernelGGL(k_ladder_xyzz, dim3(blocks_for(count, 256)), dim3(256), 0, s, pts, phi, lad, count,
                       queue_b, queue_r);
    return hipGetLastError();
}

hipError_t launch_ladder(G1J *lad, const G1X *u, uint32_t splits, uint32_t nblobs, uint32_t lanes,
                         hipStream_t s) {
    // Workgroups of four waves: the dispatcher spreads a workgroup's waves over
    // a CU's four SIMDs, where one-wave workgroups doubled up on some SIMDs
    // while others idled (1.6 ms instead of 0.9 ms for 256 blobs; each wave's
    // chain is issue-bound, so a shared SIMD halves its rate).  Whole groups
    // of LANES lanes fill each 256-thread block.
    const uint64_t threads = nblobs * 128ull * splits * lanes;
    if (lanes == 4)
        hipLaunchKernelGGL(k_ladder<4>, dim3(blocks_for(threads, 256)), dim3(256), 0, s, lad, u, splits, nblobs);
    else if (lanes == 2)
        hipLaunchKernelGGL(k_ladder<2>, dim3(blocks_for(threads, 256)), dim3(256), 0, s, lad, u, splits, nblobs);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_sum_groups(G1X *usum, const G1X *part, uint32_t groups, uint32_t nmsm, bool quad,
                             uint32_t force_checked, hipStream_t s) {
    if (nmsm % 32 != 0 || groups > 24) return hipErrorInvalidValue;
    if (quad)
        hipLaunchKernelGGL(k_sum_groups<4>, dim3(blocks_for(nmsm * 32ull, 256)), dim3(256), 0, s, usum, part,
                           groups, nmsm, force_checked);
    else
        hipLaunchKernelGGL(k_sum_groups<1>, dim3(blocks_for(nmsm * 8ull, 256)), dim3(256), 0, s, usum, part,
                           groups, nmsm, force_checked);
    return hipGetLastError();
}

hipError_t launch_glv_phi(G1A *phi, const G1A *pts, uint64_t count, uint32_t *queue_b, uint32_t *queue_r,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_glv_phi, dim3(blocks_for(count, 256)), dim3(256), 0, s, phi, pts, count, queue_b,
                       queue_r);
    return hipGetLastError();
}

}  // namespace vkzg_amd
