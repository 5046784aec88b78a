// FK20 cell-proof kernels (the reference's computeBatch, src/vkzg.cpp:899-1039).
//
//   k_fk20_scalars   blob bytes -> canonical check -> iNTT_4096 -> circulant
//                    columns -> NTT_128 -> phase-A scalars (1-8 blocks per
//                    blob, each with the whole polynomial resident in LDS)
//   k_phase_a        u[j] = sum_i scal[j][i] * xext[j][i]: fixed-base windowed
//                    MSM against the resident multiples table, 8 lanes per tile
//   k_sum_groups     the window-group partials of each MSM summed
//   k_ladder         2^(8d) times the split-circulant folds of u
//   k_phase_b        phase-B buckets, one lane per (blob, output, bucket)
//   k_bucket_reduce  sum_k k * B[k-1] per output
//   k_finalize       C+ +- C-, affine, compressed, bit-reversed order

#include "proof_common.h"

namespace vkzg_amd {

namespace {
constexpr int kScalarThreads = 512;
constexpr int kScalarWaves = kScalarThreads / 64;
constexpr int kPolyPitch = 4096;                 // LDS pitch of the polynomial (SoA limbs)
constexpr int kColPitch = kScalarWaves * 128;    // LDS pitch of the per-wave circulant columns

// `parts` blocks per blob (1, 2, 4 or 8), 8 waves each (2 per SIMD).  LDS:
// 8 x 4096 words of polynomial + 8 x 1024 words of circulant columns =
// 160 KiB, all of a CU.  Every part parses the blob and runs the inverse
// NTT_4096 in its own LDS; part p then takes the circulant columns of rounds
// [p R, (p + 1) R), R = 8 / parts.  Small batches leave most CUs idle, so
// they split a blob over more blocks: a shorter chain for the same CUs'
// time.  Part 0 alone owns the blob's error flag and its coefficients.
__global__ void __launch_bounds__(kScalarThreads)
    k_fk20_scalars(Fr *scal, const uint8_t *blobs, uint64_t blob_stride, uint32_t *err,
                   const Fr *roots, const Fr *roots_inv, uint32_t *queue, Fr *coeffs,
                   uint32_t splits, uint32_t parts) {
    __shared__ uint32_t poly[8 * kPolyPitch];
    __shared__ uint32_t cols[8 * kColPitch];
    const uint32_t blob = blockIdx.x / parts, part = blockIdx.x % parts;
    const int tid = threadIdx.x;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(blobs + blob_stride * blob);

    // 1. Parse, check canonical, enter Montgomery form (k_blob_to_fr.comp:23-37).
    //    blob[i] = p(omega_4096^brev12(i)), i.e. the bit-reversed evaluation
    //    order a decimation-in-time transform takes as input.
    bool bad = false;
#pragma unroll 1
    for (int i = tid; i < 4096; i += kScalarThreads) {
        uint32_t w[8];
#pragma unroll
        for (int k = 0; k < 8; k++) w[k] = src[8 * i + k];
        const Fr v = fr_from_be_words(w);
        if (!limbs_lt<8>(v.v, cfr::P)) bad = true;
        lds_store_fr(poly, kPolyPitch, i, fr_from_canonical(v));
    }
    // Every blob's flag is cleared here and set by any lane that saw a
    // non-canonical element, so err needs no clearing before the launch (a
    // barrier orders the clear before the sets; __syncthreads_or would take
    // LDS beyond the 160 KiB the arrays use).
    if (part == 0 && tid == 0) err[blob] = 0u;
    __syncthreads();
    if (part == 0 && bad) atomicOr(&err[blob], 1u);
    // Clear the tile counter of the persistent phase-A launch that follows.
    if (blockIdx.x == 0 && tid == 0) queue[0] = 0u;

    // 2. Inverse NTT_4096 -> monomial coefficients, scaled by 1/4096.
    lds_ntt_dit<12>(poly, kPolyPitch, 1, roots_inv, tid, kScalarThreads);
    const Fr inv_n = fr_const(cfr::INV_4096_M);
#pragma unroll 1
    for (int i = tid; i < 4096; i += kScalarThreads)
        lds_store_fr(poly, kPolyPitch, i, fr_mul(lds_load_fr(poly, kPolyPitch, i), inv_n));
    __syncthreads();
    // Optionally the monomial coefficients, zero-padded to 8192, for the cell
    // extension (compute_cells_and_kzg_proofs).
    if (coeffs && part == 0) {
        Fr *cdst = coeffs + (size_t)blob * 8192;
#pragma unroll 1
        for (int i = tid; i < 8192; i += kScalarThreads)
            cdst[i] = i < 4096 ? lds_load_fr(poly, kPolyPitch, i) : fr_zero();
    }

    // 3. 64 circulant columns, 8 at a time (one per wave): tc_i[0] = c[4095-i],
    //    tc_i[128-m] = c[4095-i-64m] (m = 1..62), zero elsewhere; NTT_128;
    //    scal[j][i] = canonical(result[j]) (k_build_circulant.comp:30-72).
    const int wave = tid >> 6;
    const int lane = tid & 63;
    uint32_t *col = cols + wave * 128;
    Fr *out = scal + (size_t)blob * splits * 128 * 64;
    // Scalar sets of a split ladder: segment m starts at 2^(kDigitBits
    // lad_start(M, m)) u; the multipliers by doubling one (under 8 kLadPos Fr
    // additions per thread, once per blob).  Indexed by constants only, so
    // they stay in registers.
    Fr mult[kMaxLadSplits];
    mult[0] = fr_one();
#pragma unroll
    for (uint32_t m = 1; m < kMaxLadSplits; m++) {
        Fr x = mult[m - 1];
        if (m < splits) {
#pragma unroll 1
            for (uint32_t k = kDigitBits * lad_start(splits, m - 1); k < kDigitBits * lad_start(splits, m); k++)
                x = fr_add(x, x);
        }
        mult[m] = x;
    }
    const int rounds = 64 / kScalarWaves / (int)parts;
#pragma unroll 1
    for (int round = (int)part * rounds; round < ((int)part + 1) * rounds; round++) {
        const int i = round * kScalarWaves + wave;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int t = lane + 64 * h;
            Fr v = fr_zero();
            if (t == 0) v = lds_load_fr(poly, kPolyPitch, 4095 - i);
            else if (t >= 66) v = lds_load_fr(poly, kPolyPitch, 4095 - i - 64 * (128 - t));
            lds_store_fr(col, kColPitch, (int)brev_bits(t, 7), v);
        }
        __syncthreads();
        // NTT_128 of this wave's column; all waves step together.
#pragma unroll 1
        for (int s = 1; s <= 7; s++) {
            const int half = 1 << (s - 1);
            const int j = lane & (half - 1);
            const int i0 = ((lane >> (s - 1)) << s) + j;
            const int i1 = i0 + half;
            const Fr u = lds_load_fr(col, kColPitch, i0);
            const Fr vv = fr_mul(lds_load_fr(col, kColPitch, i1), roots[j << (13 - s)]);
            lds_store_fr(col, kColPitch, i0, fr_add(u, vv));
            lds_store_fr(col, kColPitch, i1, fr_sub(u, vv));
            __syncthreads();
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int j = lane + 64 * h;
            const Fr v = lds_load_fr(col, kColPitch, j);
            out[j * 64 + i] = fr_signed_magnitude(fr_to_canonical(v));
#pragma unroll
            for (uint32_t m = 1; m < kMaxLadSplits; m++)
                if (m < splits)
                    out[(size_t)m * 8192 + j * 64 + i] = fr_signed_magnitude(fr_to_canonical(fr_mul(v, mult[m])));
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_fk20_scalars(Fr *scal, const uint8_t *blobs, uint64_t blob_stride, uint32_t *err,
                               const Fr *roots, const Fr *roots_inv, uint32_t *queue, Fr *coeffs,
                               uint32_t splits, uint32_t nblobs, hipStream_t s) {
    // blocks of a blob: up to 8 while the batch fills fewer than 256 blocks
    uint32_t parts = 1;
    while (parts < 8 && nblobs * parts * 2 <= 256) parts *= 2;
    hipLaunchKernelGGL(k_fk20_scalars, dim3(nblobs * parts), dim3(kScalarThreads), 0, s, scal, blobs,
                       blob_stride, err, roots, roots_inv, queue, coeffs, splits, parts);
    return hipGetLastError();
}

}  // namespace vkzg_amd
