// FK20 phase B inputs: doubling ladders of u[j] and the split-circulant folds.
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

// lad[blob][d][j] = 2^(8d) * u[blob][j], d < kLadPos.  With `splits` = M
// segments, pair (blob, m, j) starts from phase A's 2^(8 start_m) u (the
// sum of its `groups` window-group partials part[((blob*M + m)*128 + j)*G + g])
// and writes positions [start_m, start_{m+1}): chains of 128 / M doublings.
__global__ __launch_bounds__(256) void k_ladder(G1J *lad, const G1X *part, uint32_t groups, uint32_t splits, uint32_t nblobs) {
    // A long dependent chain: give these waves issue priority over the
    // throughput kernels of another sub-batch sharing the SIMD.
    __builtin_amdgcn_s_setprio(3);
    // two lanes per ladder (g1j_dbl_pair); pairs exit together
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = tid >> 1, q = tid & 1u;
    if (g >= nblobs * 128 * splits) return;
    const uint32_t j = g % 128, m = (g / 128) % splits, blob = g / (128 * splits);
    G1X u = part[(size_t)g * groups];
#pragma unroll 1
    for (uint32_t k = 1; k < groups; k++) u = g1x_add(u, part[(size_t)g * groups + k]);
    G1J p = g1j_from_xyzz(u);
    const uint32_t seg = 16 / splits;  // positions per segment; the last one also takes d = 16
    const uint32_t d0 = m * seg, d1 = m + 1 == splits ? kLadPos : d0 + seg;
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
        for (int k = 0; k < 8; k++) p = g1j_dbl_pair(p, q);
    }
}

// The split-circulant folds of the ladder, position by position
// (k_fold_ladder.comp), with the cyclic half split once more:
//   F+[j] = L[j] + L[j+64], F-[j] = L[j] - L[j+64] (j < 64),
//   Fc[i] = F+[i] + F+[i+32], Fn[i] = F+[i] - F+[i+32] (i < 32),
// stored fold[blob][0][i][d] = Fc[i], fold[blob][0][32+i][d] = Fn[i],
// fold[blob][1][j][d] = F-[j] (the three phase-B segments' points).
__global__ void __launch_bounds__(64) k_fold(G1J *fold, const G1J *lad, uint32_t nblobs) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblobs * 32 * kLadPos) return;
    const uint32_t blob = g / (32 * kLadPos);
    const uint32_t r = g % (32 * kLadPos);
    const uint32_t i = r / kLadPos, d = r % kLadPos;
    const G1J *L = lad + (size_t)blob * kLadPos * 128 + d * 128;
    G1J *F = fold + (size_t)blob * 2 * 64 * kLadPos + d;
    G1J p0;
    {
        const G1J a = L[i], b = L[i + 64];
        p0 = g1j_add(a, b);
        F[(64 + i) * kLadPos] = g1j_add(a, g1j_neg(b));
    }
    G1J p1;
    {
        const G1J a = L[i + 32], b = L[i + 96];
        p1 = g1j_add(a, b);
        F[(96 + i) * kLadPos] = g1j_add(a, g1j_neg(b));
    }
    F[i * kLadPos] = g1j_add(p0, p1);
    F[(32 + i) * kLadPos] = g1j_add(p0, g1j_neg(p1));
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

}  // namespace

hipError_t launch_ladder(G1J *lad, const G1X *part, uint32_t groups, uint32_t splits, uint32_t nblobs,
                         hipStream_t s) {
    // Workgroups of four waves: the dispatcher spreads a workgroup's waves over
    // a CU's four SIMDs, where one-wave workgroups doubled up on some SIMDs
    // while others idled (1.6 ms instead of 0.9 ms for 256 blobs; each wave's
    // chain is issue-bound, so a shared SIMD halves its rate).
    hipLaunchKernelGGL(k_ladder, dim3(blocks_for(nblobs * 128ull * splits * 2, 256)), dim3(256), 0, s, lad,
                       part, groups, splits, nblobs);
    return hipGetLastError();
}

hipError_t launch_fold(G1J *fold, const G1J *lad, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_fold, dim3(blocks_for(nblobs * 32ull * kLadPos, 64)), dim3(64), 0, s, fold,
                       lad, nblobs);
    return hipGetLastError();
}

hipError_t launch_glv_phi(G1A *phi, const G1A *pts, uint64_t count, uint32_t *queue_b, uint32_t *queue_r,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_glv_phi, dim3(blocks_for(count, 256)), dim3(256), 0, s, phi, pts, count, queue_b,
                       queue_r);
    return hipGetLastError();
}

}  // namespace vkzg_amd
