// FK20 cell-proof kernels (the reference's computeBatch, src/vkzg.cpp:899-1039).
//
//   k_fk20_scalars   blob bytes -> canonical check -> iNTT_4096 -> circulant
//                    columns -> NTT_128 -> phase-A scalars (one block per blob,
//                    the whole polynomial resident in LDS)
//   k_phase_a        u[j] = sum_i scal[j][i] * xext[j][i]: fixed-base windowed
//                    MSM against the resident multiples table, 16 lanes per MSM
//   k_ladder/k_fold  2^(8d) * u[j] and the split-circulant inputs L+-
//   k_phase_b        phase-B buckets, one lane per (blob, output, bucket)
//   k_bucket_reduce  sum_k k * B[k-1] per output
//   k_finalize       C+ +- C-, affine, compressed, bit-reversed order
#include "curve.h"
#include "kernels.h"
#include "ntt.h"

namespace vkzg_amd {

namespace {

constexpr int kScalarThreads = 256;
constexpr int kPolyPitch = 4096;          // LDS pitch of the polynomial (SoA limbs)
constexpr int kColPitch = 4 * 128;        // LDS pitch of the 4 per-wave circulant columns

// One block per blob.  LDS: 8 x 4096 words of polynomial + 8 x 512 words of
// circulant columns = 144 KiB (gfx950 has 160 KiB per CU).
__global__ void __launch_bounds__(kScalarThreads)
    k_fk20_scalars(Fr *scal, const uint8_t *blobs, uint32_t *err, const Fr *roots,
                   const Fr *roots_inv) {
    __shared__ uint32_t poly[8 * kPolyPitch];
    __shared__ uint32_t cols[8 * kColPitch];
    const uint32_t blob = blockIdx.x;
    const int tid = threadIdx.x;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(blobs + 131072ull * blob);

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
    if (bad) atomicOr(&err[blob], 1u);
    __syncthreads();

    // 2. Inverse NTT_4096 -> monomial coefficients, scaled by 1/4096.
    lds_ntt_dit<12>(poly, kPolyPitch, 1, roots_inv, tid, kScalarThreads);
    const Fr inv_n = fr_const(cfr::INV_4096_M);
#pragma unroll 1
    for (int i = tid; i < 4096; i += kScalarThreads)
        lds_store_fr(poly, kPolyPitch, i, fr_mul(lds_load_fr(poly, kPolyPitch, i), inv_n));
    __syncthreads();

    // 3. 64 circulant columns, 4 at a time (one per wave): tc_i[0] = c[4095-i],
    //    tc_i[128-m] = c[4095-i-64m] (m = 1..62), zero elsewhere; NTT_128;
    //    scal[j][i] = canonical(result[j]) (k_build_circulant.comp:30-72).
    const int wave = tid >> 6;
    const int lane = tid & 63;
    uint32_t *col = cols + wave * 128;
    Fr *out = scal + (size_t)blob * 128 * 64;
#pragma unroll 1
    for (int round = 0; round < 16; round++) {
        const int i = round * 4 + wave;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int t = lane + 64 * h;
            Fr v = fr_zero();
            if (t == 0) v = lds_load_fr(poly, kPolyPitch, 4095 - i);
            else if (t >= 66) v = lds_load_fr(poly, kPolyPitch, 4095 - i - 64 * (128 - t));
            lds_store_fr(col, kColPitch, (int)brev_bits(t, 7), v);
        }
        __syncthreads();
        // NTT_128 of this wave's column; all four waves step together.
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
            out[j * 64 + i] = fr_to_canonical(lds_load_fr(col, kColPitch, j));
        }
        __syncthreads();
    }
}

// c-bit window `w` of a canonical 256-bit scalar.
VKZG_DEV uint32_t window_bits(const Fr &s, uint32_t w, uint32_t c) {
    const uint32_t off = w * c;
    const uint32_t li = off >> 5;
    const uint32_t sh = off & 31u;
    uint32_t lo = li < 8 ? s.v[li] >> sh : 0u;
    if (sh && li + 1 < 8) lo |= s.v[li + 1] << (32u - sh);
    return lo & ((1u << c) - 1u);
}

constexpr int kMsmLanes = 16;  // lanes per phase-A MSM (64 terms / 16 = 4 each)

// u[blob][j] = sum_{i<64} scal[blob][j][i] * xext[j][i], with signed c-bit
// digits v in [-(2^(c-1)-1), 2^(c-1)] (the recoding of k_phase_a_sort.comp:
// 31-40 with a c-bit window) and table[j][i][w][|v|-1] = |v| * 2^(cw) * xext[j][i].
__global__ void __launch_bounds__(256)
    k_phase_a(G1X *u, const Fr *scal, const G1A *table, uint32_t c, uint32_t W, uint32_t nmsm) {
    const uint32_t msm = blockIdx.x * (256 / kMsmLanes) + threadIdx.x / kMsmLanes;
    const uint32_t t = threadIdx.x % kMsmLanes;
    const bool live = msm < nmsm;
    const uint32_t j = msm % 128;
    const uint32_t H = 1u << (c - 1);
    G1X acc = g1x_identity();
    if (live) {
#pragma unroll 1
        for (uint32_t q = 0; q < 64 / kMsmLanes; q++) {
            const uint32_t i = t * (64 / kMsmLanes) + q;
            const Fr s = scal[(size_t)msm * 64 + i];
            const G1A *tab = table + ((size_t)(j * 64 + i) * W) * H;
            uint32_t carry = 0;
#pragma unroll 1
            for (uint32_t w = 0; w < W; w++) {
                uint32_t raw = window_bits(s, w, c) + carry;
                carry = raw > H ? 1u : 0u;
                const bool neg = carry != 0;
                const uint32_t mag = neg ? (2u * H - raw) : raw;
                if (mag == 0) continue;
                const G1A pt = g1a_neg_if(tab[(size_t)w * H + (mag - 1)], neg);
                acc = g1x_madd(acc, pt);
            }
        }
    }
    // Sum the 16 partial results of this MSM.
#pragma unroll 1
    for (int m = 1; m < kMsmLanes; m <<= 1) {
        G1X o;
#pragma unroll
        for (int k = 0; k < 12; k++) {
            o.x.v[k] = __shfl_xor(acc.x.v[k], m);
            o.y.v[k] = __shfl_xor(acc.y.v[k], m);
            o.zz.v[k] = __shfl_xor(acc.zz.v[k], m);
            o.zzz.v[k] = __shfl_xor(acc.zzz.v[k], m);
        }
        acc = g1x_add(acc, o);
    }
    if (live && t == 0) u[msm] = acc;
}

// lad[blob][d][j] = 2^(8d) * u[blob][j], d < 32.
__global__ void k_ladder(G1J *lad, const G1X *u, uint32_t nblobs) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblobs * 128) return;
    const uint32_t blob = g / 128, j = g % 128;
    G1J p = g1j_from_xyzz(u[g]);
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

// Lane (blob, a, k): bucket k of output a.  Output a < 64 is the cyclic half
// over L+, a >= 64 the negacyclic half over L- with a sign flip when a tap
// wraps (k_phase_b_split.comp:25-61).  The item list depends only on (half,
// k), so it is uniform across a wave.
__global__ void __launch_bounds__(64)
    k_phase_b(G1X *buckets, const G1A *lad_aff, const uint32_t *items, const uint32_t *offsets) {
    const uint32_t k = blockIdx.x;                       // bucket, 0..127
    const uint32_t ba = blockIdx.y * 64 + threadIdx.x;   // blob * 128 + a
    const uint32_t blob = ba / 128, a = ba % 128;
    const uint32_t half = a >> 6, aa = a & 63u;
    const uint32_t lo = offsets[half * 129 + k], hi = offsets[half * 129 + k + 1];
    const G1A *L = lad_aff + (size_t)(blob * 2 + half) * 64 * 32;
    G1X acc = g1x_identity();
#pragma unroll 1
    for (uint32_t it = lo; it < hi; it++) {
        const uint32_t packed = items[it];
        const uint32_t e = packed & 0xffu;
        const uint32_t d = (packed >> 8) & 0xffu;
        bool neg = ((packed >> 16) & 1u) != 0;
        if (half && e > aa) neg = !neg;
        const uint32_t src = (aa - e) & 63u;
        acc = g1x_madd(acc, g1a_neg_if(L[src * 32 + d], neg));
    }
    buckets[(size_t)ba * 128 + k] = acc;
}

VKZG_DEV G1X g1x_shfl_xor(const G1X &p, int m) {
    G1X o;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        o.x.v[k] = __shfl_xor(p.x.v[k], m);
        o.y.v[k] = __shfl_xor(p.y.v[k], m);
        o.zz.v[k] = __shfl_xor(p.zz.v[k], m);
        o.zzz.v[k] = __shfl_xor(p.zzz.v[k], m);
    }
    return o;
}

constexpr int kReduceLanes = 8;

// cvals[blob][a] = sum_{k=1..128} k * buckets[blob][a][k-1]
// (field.glsl.inc:478-517).  Lane l of a team owns buckets [16l, 16l+16):
// rs_l = sum_i (i+1) B[16l+i] by running sums, s_l = sum_i B[16l+i]; the
// output is sum_l (rs_l + 16 l * s_l).
__global__ void __launch_bounds__(256) k_bucket_reduce(G1X *cvals, const G1X *buckets, uint32_t nout) {
    const uint32_t o = blockIdx.x * (256 / kReduceLanes) + threadIdx.x / kReduceLanes;
    const uint32_t l = threadIdx.x % kReduceLanes;
    constexpr uint32_t per = 128 / kReduceLanes;
    G1X s = g1x_identity(), rs = g1x_identity();
    if (o < nout) {
        const G1X *B = buckets + (size_t)o * 128 + l * per;
#pragma unroll 1
        for (int i = per - 1; i >= 0; i--) {
            s = g1x_add(s, B[i]);
            rs = g1x_add(rs, s);
        }
    }
    // t = (16 l) * s by double-and-add over the bits of 16 l.
    const uint32_t mult = per * l;
    G1X t = g1x_identity();
    if (mult) {
#pragma unroll 1
        for (int bit = 31 - __builtin_clz(mult); bit >= 0; bit--) {
            t = g1x_dbl(t);
            if ((mult >> bit) & 1u) t = g1x_add(t, s);
        }
    }
    G1X v = g1x_add(rs, t);
#pragma unroll 1
    for (int m = 1; m < kReduceLanes; m <<= 1) v = g1x_add(v, g1x_shfl_xor(v, m));
    if (o < nout && l == 0) cvals[o] = v;
}

// XYZZ -> affine for two points with one inversion.
VKZG_DEV void g1x_pair_to_affine(G1A &ra, G1A &rb, const G1X &a, const G1X &b) {
    const bool ia = fp_is_zero(a.zz), ib = fp_is_zero(b.zz);
    const Fp za = ia ? fp_one() : a.zzz;
    const Fp zb = ib ? fp_one() : b.zzz;
    const Fp inv = fp_inv(fp_mul(za, zb));
    const Fp iza = fp_mul(inv, zb);  // 1/zzz_a
    const Fp izb = fp_mul(inv, za);  // 1/zzz_b
    ra = g1a_identity();
    rb = g1a_identity();
    if (!ia) {
        ra.y = fp_mul(a.y, iza);
        ra.x = fp_mul(a.x, fp_sqr(fp_mul(a.zz, iza)));  // 1/zz = (zz/zzz)^2
    }
    if (!ib) {
        rb.y = fp_mul(b.y, izb);
        rb.x = fp_mul(b.x, fp_sqr(fp_mul(b.zz, izb)));
    }
}

// out[a] = C+[a] + C-[a], out[a+64] = C+[a] - C-[a] (k_combine_split.comp),
// compressed and written at brev7(k) (k_compress_proofs.comp).
__global__ void k_finalize(uint8_t *proofs, const G1X *cvals, uint32_t nblobs) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblobs * 64) return;
    const uint32_t blob = g / 64, a = g % 64;
    const G1X cp = cvals[blob * 128 + a];
    const G1X cm = cvals[blob * 128 + a + 64];
    G1A p0, p1;
    g1x_pair_to_affine(p0, p1, g1x_add(cp, cm), g1x_add(cp, g1x_neg(cm)));
    uint32_t *dst = reinterpret_cast<uint32_t *>(proofs + 6144ull * blob);
    uint32_t w[12];
    g1a_compress_words(w, p0);
    const uint32_t k0 = brev_bits(a, 7), k1 = brev_bits(a + 64, 7);
#pragma unroll
    for (int i = 0; i < 12; i++) dst[k0 * 12 + i] = w[i];
    g1a_compress_words(w, p1);
#pragma unroll
    for (int i = 0; i < 12; i++) dst[k1 * 12 + i] = w[i];
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_fk20_scalars(Fr *scal, const uint8_t *blobs, uint32_t *err, const Fr *roots,
                               const Fr *roots_inv, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_fk20_scalars, dim3(nblobs), dim3(kScalarThreads), 0, s, scal, blobs, err,
                       roots, roots_inv);
    return hipGetLastError();
}

hipError_t launch_phase_a(G1X *u, const Fr *scal, const G1A *table, uint32_t c, uint32_t W,
                          uint32_t nblobs, hipStream_t s) {
    const uint32_t nmsm = nblobs * 128;
    hipLaunchKernelGGL(k_phase_a, dim3(blocks_for(nmsm, 256 / kMsmLanes)), dim3(256), 0, s, u, scal,
                       table, c, W, nmsm);
    return hipGetLastError();
}

hipError_t launch_ladder(G1J *lad, const G1X *u, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_ladder, dim3(blocks_for(nblobs * 128ull, 64)), dim3(64), 0, s, lad, u,
                       nblobs);
    return hipGetLastError();
}

hipError_t launch_fold(G1J *fold, const G1J *lad, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_fold, dim3(blocks_for(nblobs * 2048ull, 64)), dim3(64), 0, s, fold, lad,
                       nblobs);
    return hipGetLastError();
}

hipError_t launch_phase_b(G1X *buckets, const G1A *lad_aff, const uint32_t *items,
                          const uint32_t *offsets, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_phase_b, dim3(128, nblobs * 2), dim3(64), 0, s, buckets, lad_aff, items,
                       offsets);
    return hipGetLastError();
}

hipError_t launch_bucket_reduce(G1X *cvals, const G1X *buckets, uint32_t nblobs, hipStream_t s) {
    const uint32_t nout = nblobs * 128;
    hipLaunchKernelGGL(k_bucket_reduce, dim3(blocks_for(nout, 256 / kReduceLanes)), dim3(256), 0, s,
                       cvals, buckets, nout);
    return hipGetLastError();
}

hipError_t launch_finalize(uint8_t *proofs, const G1X *cvals, uint32_t nblobs, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(nblobs * 64ull, 64)), dim3(64), 0, s, proofs,
                       cvals, nblobs);
    return hipGetLastError();
}

}  // namespace vkzg_amd
