// Prover-creation kernels: root tables, setup decompression, the FK20
// Fourier-domain setup x_ext_fft and the phase-A fixed-base table, all built
// on the device and kept resident in HBM.  They replace the reference's
// precomputed-tables file and its upload path (src/precomputed_tables.cpp:
// 34-73, src/vkzg.cpp:669-697, k_pow_seq.comp).

#include "proof_common.h"

namespace vkzg_amd {

namespace {

VKZG_DEV Fr fr_pow_small(const Fr &base, uint32_t k) {
    Fr acc = fr_one();
    Fr b = base;
#pragma unroll 1
    while (k) {
        if (k & 1u) acc = fr_mul(acc, b);
        b = fr_sqr(b);
        k >>= 1;
    }
    return acc;
}

__global__ void k_roots(Fr *roots, Fr *roots_inv, Fr *shift_fwd, Fr *shift_inv) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 8192) return;
    roots[k] = fr_pow_small(fr_const(cfr::OMEGA8192_M), k);
    roots_inv[k] = fr_pow_small(fr_const(cfr::OMEGA8192_INV_M), k);
    const Fr inv_n = fr_const(cfr::INV_8192_M);
    shift_fwd[k] = fr_mul(fr_pow_small(fr_const(cfr::SEVEN_M), k), inv_n);
    shift_inv[k] = fr_mul(fr_pow_small(fr_const(cfr::INV_SEVEN_M), k), inv_n);
}

// 48-byte compressed G1 -> affine.  Sets err[0] on a malformed encoding.
__global__ void __launch_bounds__(64) k_decompress(G1A *out, const uint8_t *in, uint32_t n, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    G1A r;
    const bool ok = g1a_decompress(r, reinterpret_cast<const uint32_t *>(in + 48ull * i));
    out[i] = r;
    if (!ok) atomicOr(err, 1u);
}

// Split-circulant phase-B scalars, canonical: out[e] = K+[e], out[64+e] = K-[e]
// with K[0] = 1/2, K[odd e] = 1/(64(1 - omega_128^e)), K[even e != 0] = 0 and
// K+-[e] = (K[e] +- K[e+64]) / 2 (src/layout_defs.h:30-45, k_phase_b_split.comp).
VKZG_DEV Fr circulant_tap(uint32_t e, const Fr *roots) {
    if (e == 0) return fr_inv(fr_add(fr_one(), fr_one()));
    if ((e & 1u) == 0) return fr_zero();
    Fr sixty_four = fr_one();
#pragma unroll 1
    for (int k = 0; k < 6; k++) sixty_four = fr_add(sixty_four, sixty_four);
    const Fr denom = fr_mul(sixty_four, fr_sub(fr_one(), roots[(64u * e) & 8191u]));
    return fr_inv(denom);
}

__global__ void k_phase_b_scalars(Fr *out, const Fr *roots) {
    const uint32_t e = threadIdx.x;
    if (e >= 64) return;
    const Fr a = circulant_tap(e, roots);
    const Fr b = circulant_tap(e + 64, roots);
    const Fr half = fr_inv(fr_add(fr_one(), fr_one()));
    out[e] = fr_to_canonical(fr_mul(fr_add(a, b), half));
    out[64 + e] = fr_to_canonical(fr_mul(fr_sub(a, b), half));
}

// Test hooks: canonical little-endian limbs in and out.
__global__ void k_debug_fp_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t aw[12], bw[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
        aw[k] = a[12 * i + k];
        bw[k] = b[12 * i + k];
    }
    const Fp xm = fp_from_canonical(fp_from_words(aw)), ym = fp_from_canonical(fp_from_words(bw));
    const Fp prod = fp_mul(xm, ym);
    const Fp r = fp_to_canonical(prod);
    // the scheduled product must equal the plain operand-scanning one modulo p
    const bool agree = limbs_eq<kFpLimbs>(fp_canonical(prod).v, fp_mul_reference(xm, ym).v);
    uint32_t rw[12];
    fp_to_words(rw, r);
#pragma unroll
    for (int k = 0; k < 12; k++) out[12 * i + k] = agree ? rw[k] : 0xffffffffu;
}

// fp_inv on canonical inputs (b is ignored); a nonzero result must also
// agree with the Fermat inverse, else all ones.
__global__ void k_debug_fp_inv(uint32_t *out, const uint32_t *a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t aw[12];
#pragma unroll
    for (int k = 0; k < 12; k++) aw[k] = a[12 * i + k];
    const Fp xm = fp_from_canonical(fp_from_words(aw));
    const Fp inv = fp_inv(xm);
    const bool agree = fp_eq(inv, fp_inv_fermat(xm));
    uint32_t rw[12];
    fp_to_words(rw, fp_to_canonical(inv));
#pragma unroll
    for (int k = 0; k < 12; k++) out[12 * i + k] = agree ? rw[k] : 0xffffffffu;
}

__global__ void k_debug_fr_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Fr x, y;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        x.v[k] = a[8 * i + k];
        y.v[k] = b[8 * i + k];
    }
    const Fr xm = fr_from_any(x), ym = fr_from_any(y);
    const Fr r = fr_to_canonical(fr_mul(xm, ym));
    Fr c;
    mont_mul_cios<8>(c.v, xm.v, ym.v, cfr::P, cfr::N0);
    const bool agree = limbs_eq<8>(c.v, fr_mul(xm, ym).v);
#pragma unroll
    for (int k = 0; k < 8; k++) out[8 * i + k] = agree ? r.v[k] : 0xffffffffu;
}

// n dependent Montgomery products per lane on lane-varying operands (the
// issue-rate measurement of vkzg_amd_debug_fp_mul_rate); each lane stores its
// result into a 1024-entry buffer so the chain is kept.
__global__ void k_debug_fp_mul_chain(Fp *out, uint32_t n) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    Fp a = fp_one(), b = fp_const(cfp::BETA_M);
    a.v[0] += gid & 0xffffu;
    b.v[1] ^= threadIdx.x & 7u;
#pragma unroll 1
    for (uint32_t i = 0; i < n; i++) a = fp_mul(a, b);
    out[gid & 1023u] = a;
}

// Issue-rate measurement of v_mad_u64_u32 itself (vkzg_amd_debug_mad_rate):
// each lane runs 8 independent accumulator chains, 16 products per chain per
// iteration, against a wave-uniform multiplier.
__global__ void __launch_bounds__(64) k_debug_mad_chain(uint64_t *out, uint32_t n) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t acc[8];
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        acc[k] = gid + k;
        x[k] = 0x9E3779B9u * (gid + 17u * k) | 1u;
    }
    const uint32_t y = 0x85EBCA6Bu + n;
#pragma unroll 1
    for (uint32_t i = 0; i < n; i++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
#pragma unroll
            for (int k = 0; k < 8; k++)
                asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(acc[k]) : "v"(x[k]), "s"(y) : "vcc");
        }
    }
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) sum ^= acc[k];
    out[gid & 1023u] = sum;
}

}  // namespace

hipError_t launch_debug_fp_mul_chain(Fp *out, uint32_t blocks, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_debug_fp_mul_chain, dim3(blocks), dim3(64), 0, s, out, n);
    return hipGetLastError();
}

hipError_t launch_debug_mad_chain(uint64_t *out, uint32_t blocks, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_debug_mad_chain, dim3(blocks), dim3(64), 0, s, out, n);
    return hipGetLastError();
}

hipError_t launch_roots(Fr *roots, Fr *roots_inv, Fr *shift_fwd, Fr *shift_inv, hipStream_t s) {
    hipLaunchKernelGGL(k_roots, dim3(8192 / 256), dim3(256), 0, s, roots, roots_inv, shift_fwd,
                       shift_inv);
    return hipGetLastError();
}

hipError_t launch_decompress(G1A *out, const uint8_t *in, uint32_t n, uint32_t *err, hipStream_t s) {
    hipLaunchKernelGGL(k_decompress, dim3(blocks_for(n, 64)), dim3(64), 0, s, out, in, n, err);
    return hipGetLastError();
}

hipError_t launch_phase_b_scalars(Fr *out, const Fr *roots, hipStream_t s) {
    hipLaunchKernelGGL(k_phase_b_scalars, dim3(1), dim3(64), 0, s, out, roots);
    return hipGetLastError();
}

hipError_t launch_debug_fp_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, uint32_t n,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_debug_fp_mul, dim3(blocks_for(n, 64)), dim3(64), 0, s, out, a, b, n);
    return hipGetLastError();
}

hipError_t launch_debug_fp_inv(uint32_t *out, const uint32_t *a, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_debug_fp_inv, dim3(blocks_for(n, 64)), dim3(64), 0, s, out, a, n);
    return hipGetLastError();
}

hipError_t launch_debug_fr_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, uint32_t n,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_debug_fr_mul, dim3(blocks_for(n, 64)), dim3(64), 0, s, out, a, b, n);
    return hipGetLastError();
}

}  // namespace vkzg_amd
