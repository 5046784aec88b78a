// FK20 phase B inputs: doubling ladders of u[j] and the split-circulant folds.
#include "proof_common.h"

namespace vkzg_amd {

namespace {
// lad[blob][d][j] = 2^(8d) * u[blob][j], d < 32, where u[blob][j] is the
// sum of phase A's `groups` window-group partials part[(blob*128 + j)*groups + g].
__global__ void k_ladder(G1J *lad, const G1X *part, uint32_t groups, uint32_t nblobs) {
    // A long dependent chain: give these waves issue priority over the
    // throughput kernels of another sub-batch sharing the SIMD.
    __builtin_amdgcn_s_setprio(3);
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblobs * 128) return;
    const uint32_t blob = g / 128, j = g % 128;
    G1X u = part[(size_t)g * groups];
#pragma unroll 1
    for (uint32_t k = 1; k < groups; k++) u = g1x_add(u, part[(size_t)g * groups + k]);
    G1J p = g1j_from_xyzz(u);
    G1J *dst = lad + (size_t)blob * 32 * 128 + j;
#pragma unroll 1
    for (int d = 0; d < 32; d++) {
        dst[d * 128] = p;
        if (d == 31) break;
#pragma unroll 1
        for (int k = 0; k < 8; k++) p = g1j_dbl(p);
    }
}

// fold[blob][h][j][d] = L[j][d] + (h ? -1 : 1) * L[j+64][d]  (k_fold_ladder.comp)
__global__ void k_fold(G1J *fold, const G1J *lad, uint32_t nblobs) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblobs * 64 * 32) return;
    const uint32_t blob = g / (64 * 32);
    const uint32_t r = g % (64 * 32);
    const uint32_t j = r / 32, d = r % 32;
    const G1J *L = lad + (size_t)blob * 32 * 128;
    const G1J a = L[d * 128 + j];
    const G1J b = L[d * 128 + j + 64];
    G1J *F = fold + (size_t)blob * 2 * 64 * 32;
    F[(0 * 64 + j) * 32 + d] = g1j_add(a, b);
    F[(1 * 64 + j) * 32 + d] = g1j_add(a, g1j_neg(b));
}

}  // namespace

hipError_t launch_ladder(G1J *lad, const G1X *part, uint32_t groups, uint32_t nblobs,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_ladder, dim3(blocks_for(nblobs * 128ull, 64)), dim3(64), 0, s, lad, part,
                       groups, nblobs);
    return hipGetLastError();
}

hipError_t launch_fold(G1J *fold, const G1J *lad, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_fold, dim3(blocks_for(nblobs * 2048ull, 64)), dim3(64), 0, s, fold, lad,
                       nblobs);
    return hipGetLastError();
}

}  // namespace vkzg_amd
