// Cell-recovery kernels (the reference's recoverBatch, src/vkzg.cpp:1086-1179).
//
// The reference computes P = coset_iNTT(coset_NTT(iNTT(E * NTT(Z))) / coset_NTT(Z))
// and evaluates NTT(P).  Z(X) = z(X^64) with z of degree <= 64 in X^64, so both
// of Z's 8192-point evaluations take only 128 distinct values:
//   Z(omega^k)     = z(omega_128^(k mod 128))            (zero on missing cells)
//   Z(7 omega^k)   = z(7^64 omega_128^(k mod 128))
// k_rec_vanishing computes those 2 x 128 values (and the 128 inverses) from
// the product form of z, which removes two of the six transforms and the
// per-element Fermat inversion of k_ew_div.comp.  The remaining four
// transforms run as two four-step passes each (8192 = 64 x 128) with every
// pointwise product fused into a pass:
//   pass A/B inverse   : load cells * Z(omega^k)          -> * 7^k / 8192
//   pass A/B forward   :                                  -> * 1 / Z(7 omega^k)
//   pass A/B inverse   :                                  -> * 7^-k / 8192
//   pass A/B forward   :                                  -> canonical cell bytes
#include "kernels.h"
#include "ntt.h"

namespace vkzg_amd {

namespace {

// Per blob: zd[k] = z(omega_128^k), izc[k] = 1 / z(7^64 omega_128^k),
// z(y) = prod over missing wire cells c of (y - omega_128^brev7(c))
// (k_vanishing_poly.comp, src/vkzg.cpp:1043-1050).
// On the coset every y satisfies y^128 = 7^8192, so the product over all 128
// cells is y^128 - 1 = 7^8192 - 1, a constant K, and
//   1 / z(y) = K^-1 * prod over present cells c of (y - omega_128^brev7(c)),
// which needs no inversion.  Block (blob, side, half) computes 64 of one
// side's 128 values; each value is split over 8 threads (16 cells each, so a
// chain is at most 16 products) whose partial products meet in an LDS tree.
// A wave holds one side and one cell range, so its skips are uniform, and the
// 2 x 2 blocks per blob spread the chains over the CUs (the work is issue-bound).
__global__ void __launch_bounds__(512)
    k_rec_vanishing(Fr *zd, Fr *izc, const uint8_t *present, const Fr *roots) {
    __builtin_amdgcn_s_setprio(3);  // short dependent chains: issue first
    __shared__ Fr part[8][64];
    const uint32_t blob = blockIdx.x;
    const bool coset = blockIdx.y != 0;
    const uint32_t l = threadIdx.x & 63u;
    const uint32_t k = blockIdx.z * 64u + l;
    const uint32_t q = threadIdx.x >> 6;  // cells [16 q, 16 q + 16)
    Fr x = roots[(64u * k) & 8191u];
    if (coset) x = fr_mul(x, fr_const(cfr::SEVEN_POW64_M));
    const uint8_t *pm = present + 128ull * blob;
    // domain side: factors of missing cells; coset side: factors of present cells
    const uint8_t want = coset ? 1 : 0;
    Fr prod = fr_one();
#pragma unroll 1
    for (uint32_t c = 16u * q; c < 16u * q + 16u; c++) {
        if ((pm[c] != 0) != (want != 0)) continue;
        const Fr root = roots[(64u * brev_bits(c, 7)) & 8191u];
        prod = fr_mul(prod, fr_sub(x, root));
    }
    part[q][l] = prod;
    __syncthreads();
#pragma unroll 1
    for (uint32_t w = 4; w > 0; w >>= 1) {
        if (q < w) part[q][l] = fr_mul(part[q][l], part[q + w][l]);
        __syncthreads();
    }
    if (q != 0) return;
    prod = part[0][l];
    if (coset) izc[blob * 128 + k] = fr_mul(prod, fr_const(cfr::INV_SEVEN_POW8192_MINUS_1_M));
    else zd[blob * 128 + k] = prod;
}

// Threads of a pass block: one per butterfly of a stage, at most 256.
template <int COLS, int N>
constexpr int pass_threads() {
    return COLS * (N / 2) < 256 ? COLS * (N / 2) : 256;
}

// Pass A: for each n1 < 64, DFT_128 over n2 of x[n1 + 64 n2], times
// omega^(+-n1 k2), stored at y[n1 * 128 + k2].  Block = COLS consecutive n1.
template <int IN_MODE, bool INVERSE, int COLS>
__global__ void __launch_bounds__((pass_threads<COLS, 128>()))
    k_rec_pass_a(Fr *out, const void *in, const Fr *zd, const Fr *roots_dir) {
    constexpr int n = 128;
    constexpr int nthreads = pass_threads<COLS, n>();
    __shared__ uint32_t lds[8 * COLS * n];
    constexpr int pitch = COLS * n;
    const uint32_t blob = blockIdx.y;
    const uint32_t n1_0 = blockIdx.x * COLS;
    const int tid = threadIdx.x;
#pragma unroll 1
    for (int e = tid; e < COLS * n; e += nthreads) {
        const int c = e % COLS, n2 = e / COLS;
        const uint32_t idx = n1_0 + c + 64u * n2;  // natural evaluation index
        Fr v;
        if (IN_MODE == 0) {
            // wire element brev13(idx) of the blob's cells, reduced mod r, times z
            const uint32_t *cw = reinterpret_cast<const uint32_t *>(in) +
                                 ((size_t)blob * 8192 + brev_bits(idx, 13)) * 8;
            uint32_t w[8];
#pragma unroll
            for (int k = 0; k < 8; k++) w[k] = cw[k];
            v = fr_mul(fr_from_any(fr_from_be_words(w)), zd[blob * 128 + (idx & 127u)]);
        } else {
            v = reinterpret_cast<const Fr *>(in)[(size_t)blob * 8192 + idx];
        }
        lds_store_fr(lds, pitch, c * n + (int)brev_bits(n2, 7), v);
    }
    __syncthreads();
    lds_ntt_dit<7>(lds, pitch, COLS, roots_dir, tid, nthreads);
#pragma unroll 1
    for (int e = tid; e < COLS * n; e += nthreads) {
        const int k2 = e % n, c = e / n;
        const uint32_t n1 = n1_0 + c;
        const Fr v = fr_mul(lds_load_fr(lds, pitch, c * n + k2), roots_dir[(n1 * k2) & 8191u]);
        out[(size_t)blob * 8192 + n1 * 128 + k2] = v;
    }
}

// Pass B: for each k2 < 128, DFT_64 over n1 of y[n1 * 128 + k2], giving
// X[k2 + 128 k1].  Block = COLS consecutive k2.
template <int OUT_MODE, bool INVERSE, int COLS>
__global__ void __launch_bounds__((pass_threads<COLS, 64>()))
    k_rec_pass_b(void *out, const Fr *in, const Fr *mul_table, const Fr *izc, const Fr *roots_dir) {
    constexpr int n = 64;
    constexpr int nthreads = pass_threads<COLS, n>();
    __shared__ uint32_t lds[8 * COLS * n];
    constexpr int pitch = COLS * n;
    const uint32_t blob = blockIdx.y;
    const uint32_t k2_0 = blockIdx.x * COLS;
    const int tid = threadIdx.x;
#pragma unroll 1
    for (int e = tid; e < COLS * n; e += nthreads) {
        const int c = e % COLS, n1 = e / COLS;
        const Fr v = in[(size_t)blob * 8192 + n1 * 128 + k2_0 + c];
        lds_store_fr(lds, pitch, c * n + (int)brev_bits(n1, 6), v);
    }
    __syncthreads();
    lds_ntt_dit<6>(lds, pitch, COLS, roots_dir, tid, nthreads);
#pragma unroll 1
    for (int e = tid; e < COLS * n; e += nthreads) {
        const int c = e % COLS, k1 = e / COLS;
        const uint32_t k = k2_0 + c + 128u * k1;  // natural output index
        Fr v = lds_load_fr(lds, pitch, c * n + k1);
        if (OUT_MODE == 0) {
            reinterpret_cast<Fr *>(out)[(size_t)blob * 8192 + k] = fr_mul(v, mul_table[k]);
        } else if (OUT_MODE == 1) {
            reinterpret_cast<Fr *>(out)[(size_t)blob * 8192 + k] =
                fr_mul(v, izc[blob * 128 + (k & 127u)]);
        } else {
            uint32_t w[8];
            fr_to_be_words(w, fr_to_canonical(v));
            uint32_t *dst = reinterpret_cast<uint32_t *>(out) +
                            ((size_t)blob * 8192 + brev_bits(k, 13)) * 8;
#pragma unroll
            for (int q = 0; q < 8; q++) dst[q] = w[q];
        }
    }
}

template <int IN_MODE, bool INV, int COLS>
void launch_a(Fr *out, const void *in, const Fr *zd, const Fr *r, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL((k_rec_pass_a<IN_MODE, INV, COLS>), dim3(64 / COLS, nblobs),
                       dim3(pass_threads<COLS, 128>()), 0, s, out, in, zd, r);
}

template <int OUT_MODE, bool INV, int COLS>
void launch_b(void *out, const Fr *in, const Fr *mul_table, const Fr *izc, const Fr *r, uint32_t nblobs,
              hipStream_t s) {
    hipLaunchKernelGGL((k_rec_pass_b<OUT_MODE, INV, COLS>), dim3(128 / COLS, nblobs),
                       dim3(pass_threads<COLS, 64>()), 0, s, out, in, mul_table, izc, r);
}

template <int COLS>
void pass_a_cols(Fr *out, const void *in, int in_mode, const Fr *zd, bool inverse, const Fr *r,
                 uint32_t nblobs, hipStream_t s) {
    if (in_mode == 0 && inverse) launch_a<0, true, COLS>(out, in, zd, r, nblobs, s);
    else if (in_mode == 0) launch_a<0, false, COLS>(out, in, zd, r, nblobs, s);
    else if (inverse) launch_a<1, true, COLS>(out, in, zd, r, nblobs, s);
    else launch_a<1, false, COLS>(out, in, zd, r, nblobs, s);
}

template <int COLS>
void pass_b_cols(void *out, const Fr *in, int out_mode, const Fr *mul_table, const Fr *izc, const Fr *r,
                 uint32_t nblobs, hipStream_t s) {
    if (out_mode == 0) launch_b<0, true, COLS>(out, in, mul_table, izc, r, nblobs, s);
    else if (out_mode == 1) launch_b<1, false, COLS>(out, in, mul_table, izc, r, nblobs, s);
    else launch_b<2, false, COLS>(out, in, mul_table, izc, r, nblobs, s);
}

}  // namespace

hipError_t launch_rec_vanishing(Fr *zd, Fr *izc, const uint8_t *present, const Fr *roots,
                                uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_rec_vanishing, dim3(nblobs, 2, 2), dim3(512), 0, s, zd, izc, present, roots);
    return hipGetLastError();
}

hipError_t launch_rec_pass_a(Fr *out, const void *in, int in_mode, const Fr *zd, bool inverse,
                             const Fr *roots, const Fr *roots_inv, uint32_t nblobs, uint32_t cols,
                             hipStream_t s) {
    const Fr *r = inverse ? roots_inv : roots;
    switch (cols) {
        case 4: pass_a_cols<4>(out, in, in_mode, zd, inverse, r, nblobs, s); break;
        case 16: pass_a_cols<16>(out, in, in_mode, zd, inverse, r, nblobs, s); break;
        case 2: pass_a_cols<2>(out, in, in_mode, zd, inverse, r, nblobs, s); break;
        default: pass_a_cols<8>(out, in, in_mode, zd, inverse, r, nblobs, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_rec_pass_b(void *out, const Fr *in, int out_mode, const Fr *mul_table,
                             const Fr *izc, bool inverse, const Fr *roots, const Fr *roots_inv,
                             uint32_t nblobs, uint32_t cols, hipStream_t s) {
    const Fr *r = inverse ? roots_inv : roots;
    switch (cols) {
        case 4: pass_b_cols<4>(out, in, out_mode, mul_table, izc, r, nblobs, s); break;
        case 16: pass_b_cols<16>(out, in, out_mode, mul_table, izc, r, nblobs, s); break;
        case 2: pass_b_cols<2>(out, in, out_mode, mul_table, izc, r, nblobs, s); break;
        default: pass_b_cols<8>(out, in, out_mode, mul_table, izc, r, nblobs, s); break;
    }
    return hipGetLastError();
}

}  // namespace vkzg_amd
