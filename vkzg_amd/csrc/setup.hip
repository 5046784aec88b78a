// Prover-creation kernels: root tables, setup decompression, the FK20
// Fourier-domain setup x_ext_fft and the phase-A fixed-base table, all built
// on the device and kept resident in HBM.  They replace the reference's
// precomputed-tables file and its upload path (src/precomputed_tables.cpp:
// 34-73, src/vkzg.cpp:669-697, k_pow_seq.comp).
#include "curve.h"
#include "kernels.h"
#include "ntt.h"

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
__global__ void k_decompress(G1A *out, const uint8_t *in, uint32_t n, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(in + 48ull * i);
    uint32_t w[12];
#pragma unroll
    for (int k = 0; k < 12; k++) w[k] = __builtin_bswap32(src[k]);
    const uint32_t b0 = w[0] >> 24;
    bool bad = (b0 & 0x80u) == 0;
    if (b0 & 0x40u) {
        uint32_t rest = w[0] & 0x3fffffffu;
#pragma unroll
        for (int k = 1; k < 12; k++) rest |= w[k];
        if (rest) bad = true;
        out[i] = g1a_identity();
        if (bad) atomicOr(err, 1u);
        return;
    }
    Fp x;
#pragma unroll
    for (int k = 0; k < 12; k++) x.v[k] = w[11 - k];
    x.v[11] &= 0x1fffffffu;
    if (!limbs_lt<12>(x.v, cfp::P)) bad = true;
    const Fp xm = fp_from_canonical(x);
    const Fp rhs = fp_add(fp_mul(fp_sqr(xm), xm), fp_const(cfp::FOUR_M));
    Fp y = fp_pow_const<12>(rhs, cfp::SQRT_EXP);
    if (!fp_eq(fp_sqr(y), rhs)) bad = true;
    const bool want_large = (b0 & 0x20u) != 0;
    if (fp_is_lex_largest(fp_to_canonical(y)) != want_large) y = fp_neg(y);
    G1A r;
    r.x = xm;
    r.y = y;
    out[i] = r;
    if (bad) atomicOr(err, 1u);
}

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

// Batched Jacobian -> affine (Montgomery's trick, k_normalize.comp:29-56).
// Thread t converts the points t, t+T, t+2T, ... (T = thread count) so that
// neighbouring lanes touch neighbouring points.
__global__ void k_jac_to_affine(G1A *out, const G1J *in, Fp *prefix, uint64_t count, uint32_t chunk,
                                uint64_t nthreads) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nthreads) return;
    Fp acc = fp_one();
    uint32_t m = 0;
#pragma unroll 1
    for (uint64_t k = t; m < chunk && k < count; k += nthreads, m++) {
        prefix[k] = acc;
        const Fp z = in[k].z;
        if (!fp_is_zero(z)) acc = fp_mul(acc, z);
    }
    Fp inv = fp_inv(acc);
#pragma unroll 1
    while (m-- > 0) {
        const uint64_t k = t + (uint64_t)m * nthreads;
        const G1J p = in[k];
        if (fp_is_zero(p.z)) {
            out[k] = g1a_identity();
            continue;
        }
        const Fp zi = fp_mul(inv, prefix[k]);
        inv = fp_mul(inv, p.z);
        const Fp zi2 = fp_sqr(zi);
        G1A r;
        r.x = fp_mul(p.x, zi2);
        r.y = fp_mul(p.y, fp_mul(zi2, zi));
        out[k] = r;
    }
}

// out[w][b] = 2^(c*w) * xext[b] for the 8192 bases.
__global__ void k_table_chains(G1J *out, const G1A *xext, uint32_t c, uint32_t W) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= 8192) return;
    G1J p = g1j_from_affine(xext[b]);
#pragma unroll 1
    for (uint32_t w = 0; w < W; w++) {
        out[(size_t)w * 8192 + b] = p;
#pragma unroll 1
        for (uint32_t k = 0; k < c; k++) p = g1j_dbl(p);
    }
}

// Thread g = (b*W + w)*nseg + seg writes table[b][w][d-1] for d in
// [seg*S+1, seg*S+S]: a running sum of the affine base P = pw[w][b],
// normalised with one inversion per thread.
__global__ void k_table_multiples(G1A *table, const G1A *pw, uint32_t c, uint32_t W,
                                  uint64_t thread_begin, uint32_t thread_count, G1J *scratch,
                                  Fp *prefix) {
    const uint32_t lt = blockIdx.x * blockDim.x + threadIdx.x;
    if (lt >= thread_count) return;
    const uint64_t g = thread_begin + lt;
    const uint32_t H = 1u << (c - 1);
    const uint32_t S = H < 64u ? H : 64u;
    const uint32_t nseg = H / S;
    const uint32_t seg = (uint32_t)(g % nseg);
    const uint64_t bw = g / nseg;  // b*W + w
    const uint32_t w = (uint32_t)(bw % W);
    const uint32_t b = (uint32_t)(bw / W);
    const G1A P = pw[(size_t)w * 8192 + b];
    const uint32_t d0 = seg * S + 1;

    G1J q = g1j_identity();
#pragma unroll 1
    for (int bit = 31 - __builtin_clz(d0); bit >= 0; bit--) {
        q = g1j_dbl(q);
        if ((d0 >> bit) & 1u) q = g1j_madd(q, P);
    }
    Fp acc = fp_one();
#pragma unroll 1
    for (uint32_t m = 0; m < S; m++) {
        const size_t slot = (size_t)m * thread_count + lt;
        scratch[slot] = q;
        prefix[slot] = acc;
        if (!fp_is_zero(q.z)) acc = fp_mul(acc, q.z);
        q = g1j_madd(q, P);
    }
    Fp inv = fp_inv(acc);
    G1A *dst = table + bw * H + (d0 - 1);
#pragma unroll 1
    for (int m = (int)S - 1; m >= 0; m--) {
        const size_t slot = (size_t)m * thread_count + lt;
        const G1J p = scratch[slot];
        if (fp_is_zero(p.z)) {
            dst[m] = g1a_identity();
            continue;
        }
        const Fp zi = fp_mul(inv, prefix[slot]);
        inv = fp_mul(inv, p.z);
        const Fp zi2 = fp_sqr(zi);
        G1A r;
        r.x = fp_mul(p.x, zi2);
        r.y = fp_mul(p.y, fp_mul(zi2, zi));
        dst[m] = r;
    }
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
    Fp x, y;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        x.v[k] = a[12 * i + k];
        y.v[k] = b[12 * i + k];
    }
    const Fp r = fp_to_canonical(fp_mul(fp_from_canonical(x), fp_from_canonical(y)));
#pragma unroll
    for (int k = 0; k < 12; k++) out[12 * i + k] = r.v[k];
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
    const Fr r = fr_to_canonical(fr_mul(fr_from_any(x), fr_from_any(y)));
#pragma unroll
    for (int k = 0; k < 8; k++) out[8 * i + k] = r.v[k];
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_roots(Fr *roots, Fr *roots_inv, Fr *shift_fwd, Fr *shift_inv, hipStream_t s) {
    hipLaunchKernelGGL(k_roots, dim3(8192 / 256), dim3(256), 0, s, roots, roots_inv, shift_fwd,
                       shift_inv);
    return hipGetLastError();
}

hipError_t launch_decompress(G1A *out, const uint8_t *in, uint32_t n, uint32_t *err, hipStream_t s) {
    hipLaunchKernelGGL(k_decompress, dim3(blocks_for(n, 64)), dim3(64), 0, s, out, in, n, err);
    return hipGetLastError();
}

hipError_t launch_xext_fft(G1J *out, const G1A *srs, const Fr *roots, hipStream_t s) {
    hipLaunchKernelGGL(k_xext_fft, dim3(64), dim3(64), 0, s, out, srs, roots);
    return hipGetLastError();
}

hipError_t launch_jac_to_affine(G1A *out, const G1J *in, Fp *prefix, uint64_t count, uint32_t chunk,
                                hipStream_t s) {
    const uint64_t nthreads = (count + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_jac_to_affine, dim3(blocks_for(nthreads, 64)), dim3(64), 0, s, out, in,
                       prefix, count, chunk, nthreads);
    return hipGetLastError();
}

hipError_t launch_table_chains(G1J *out, const G1A *xext, uint32_t c, uint32_t W, hipStream_t s) {
    hipLaunchKernelGGL(k_table_chains, dim3(8192 / 64), dim3(64), 0, s, out, xext, c, W);
    return hipGetLastError();
}

hipError_t launch_table_multiples(G1A *table, const G1A *pw, uint32_t c, uint32_t W,
                                  uint64_t thread_begin, uint32_t thread_count, G1J *scratch,
                                  Fp *prefix, hipStream_t s) {
    hipLaunchKernelGGL(k_table_multiples, dim3(blocks_for(thread_count, 64)), dim3(64), 0, s, table,
                       pw, c, W, thread_begin, thread_count, scratch, prefix);
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

hipError_t launch_debug_fr_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, uint32_t n,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_debug_fr_mul, dim3(blocks_for(n, 64)), dim3(64), 0, s, out, a, b, n);
    return hipGetLastError();
}

}  // namespace vkzg_amd
