// Prover creation: the FK20 setup in the Fourier domain (x_ext_fft).
#include "proof_common.h"

namespace vkzg_amd {

namespace {

// MSB-first double-and-add by a canonical scalar (setup only).
VKZG_DEV G1J g1j_mul_canonical(const G1J &p, const Fr &k) {
    G1J acc = g1j_identity();
#pragma unroll 1
    for (int bit = 255; bit >= 0; bit--) {
        acc = g1j_dbl(acc);
        if ((k.v[bit >> 5] >> (bit & 31)) & 1u) acc = g1j_add(acc, p);
    }
    return acc;
}

// x_i[t] = srs[4096-64-1-i-64t] for t < 63, identity for t >= 63;
// x_ext_fft[.][i] = G1-FFT_128(x_i).  One block of 64 threads per i.
__global__ void __launch_bounds__(64) k_xext_fft(G1J *out, const G1A *srs, const Fr *roots) {
    __shared__ G1J pts[128];
    const int i = blockIdx.x;
    const int tid = threadIdx.x;
    for (int t = tid; t < 128; t += 64) {
        G1J v = g1j_identity();
        if (t < 63) v = g1j_from_affine(srs[4096 - 64 - 1 - i - 64 * t]);
        pts[brev_bits(t, 7)] = v;
    }
    __syncthreads();
#pragma unroll 1
    for (int s = 1; s <= 7; s++) {
        const int half = 1 << (s - 1);
        const int j = tid & (half - 1);
        const int i0 = ((tid >> (s - 1)) << s) + j;
        const int i1 = i0 + half;
        const G1J u = pts[i0];
        G1J v = pts[i1];
        if (j) v = g1j_mul_canonical(v, fr_to_canonical(roots[j << (13 - s)]));
        __syncthreads();
        pts[i0] = g1j_add(u, v);
        pts[i1] = g1j_add(u, g1j_neg(v));
        __syncthreads();
    }
    for (int k = tid; k < 128; k += 64) out[k * 64 + i] = pts[k];
}

}  // namespace

hipError_t launch_xext_fft(G1J *out, const G1A *srs, const Fr *roots, hipStream_t s) {
    hipLaunchKernelGGL(k_xext_fft, dim3(64), dim3(64), 0, s, out, srs, roots);
    return hipGetLastError();
}

}  // namespace vkzg_amd
