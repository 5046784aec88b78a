// Batch cell-proof verification and commitments on the device: the kernels
// around the MSM (msm.hip) and the pairing (pairing.hip).  The check is the
// consensus specs' verify_cell_kzg_proof_batch: with Fiat-Shamir weights r^k,
//   e(sum_k r^k pi_k, [tau^64]Q) = e(RL, Q),
//   RL = sum_j w_j C_j - [I(tau)] + sum_k r^k z_k pi_k,
// where w_j sums r^k over the cells of commitment j, z_k = h_c^64 for the
// coset shift h_c = omega_8192^brev7(c) of cell index c, and
// I = sum_k r^k I_k with I_k the degree < 64 interpolant of cell k on its
// coset (element j at h_c omega_64^brev6(j)).  The reference has no verifier
// (include/vkzg.h:14-18); cells are laid out as its recovery path writes them
// (src/vkzg.cpp:1086-1179, k_fr_to_cells.comp).
#include "verify.h"

#include "ntt.h"
#include "proof_common.h"

namespace vkzg_amd {

namespace {

VKZG_DEV bool g1x_eq(const G1X &a, const G1X &b) {
    const bool ia = g1x_is_identity(a), ib = g1x_is_identity(b);
    if (ia || ib) return ia && ib;
    return fp_eq(fp_mul(a.x, b.zz), fp_mul(b.x, a.zz)) && fp_eq(fp_mul(a.y, b.zzz), fp_mul(b.y, a.zzz));
}

// P in G1  <=>  phi(P) == [LAMBDA]P with LAMBDA = z^2 - 1, i.e.
// [z^2]P == phi(P) + P.  Sound for every point of E(Fp): phi^2 + phi + 1 = 0,
// so (phi + LAMBDA + 1)(phi - LAMBDA) = -(LAMBDA^2 + LAMBDA + 1) = -r, and
// (phi - LAMBDA)Q = 0 forces [r]Q = 0.  [z^2]P costs two 64-bit chains.
// The checked additions as calls: inlined into the two chains below they
// need more registers than a 64-thread kernel gets without spilling.
__device__ __attribute__((noinline)) G1X g1x_madd_call(const G1X &a, const G1A &b) { return g1x_madd(a, b); }
__device__ __attribute__((noinline)) G1X g1x_add_call(const G1X &a, const G1X &b) { return g1x_add(a, b); }

VKZG_DEV bool g1_in_subgroup(const G1A &p) {
    if (g1a_is_identity(p)) return true;
    G1X a = g1x_from_affine(p);
#pragma unroll 1
    for (int i = 62; i >= 0; i--) {
        a = g1x_dbl(a);
        if ((kZAbs >> i) & 1u) a = g1x_madd_call(a, p);
    }
    G1X b = a;
#pragma unroll 1
    for (int i = 62; i >= 0; i--) {
        b = g1x_dbl(b);
        if ((kZAbs >> i) & 1u) b = g1x_add_call(b, a);
    }
    G1A phi = p;
    phi.x = fp_mul(p.x, fp_const(cfp::BETA_M));
    return g1x_eq(b, g1x_madd(g1x_from_affine(p), phi));
}

// pts[i] = decompress(bytes[i]) for i < n; pts[n + i] = srs[i] for i < nsrs.
__global__ void __launch_bounds__(64) k_vfy_points(G1A *pts, const uint8_t *bytes, uint32_t n, const G1A *srs, uint32_t nsrs,
                             uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        G1A r;
        if (!g1a_decompress(r, reinterpret_cast<const uint32_t *>(bytes + 48ull * i))) atomicOr(err, kVfyBadPoint);
        pts[i] = r;
    } else if (i < n + nsrs) {
        pts[i] = srs[i - n];
    }
}

__global__ void __launch_bounds__(64) k_vfy_subgroup(const G1A *pts, uint32_t n, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!g1_in_subgroup(pts[i])) atomicOr(err, kVfyNotInGroup);
}

// ev[k * 64 + m] = value of cell k at h_c omega_64^m (Montgomery): cell
// element j sits at m = brev6(j).  Non-canonical elements set kVfyBadCell.
__global__ void k_vfy_cells(Fr *ev, const uint8_t *cells, uint32_t n, uint32_t *err) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * 64) return;
    const uint32_t k = g / 64, j = g % 64;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(cells + 2048ull * k + 32ull * j);
    uint32_t w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) w[q] = src[q];
    const Fr v = fr_from_be_words(w);
    if (!limbs_lt<8>(v.v, cfr::P)) atomicOr(err, kVfyBadCell);
    ev[(size_t)k * 64 + brev_bits(j, 6)] = fr_from_canonical(v);
}

// rp[k] = r^k (Montgomery) for the canonical challenge r.
__global__ void k_vfy_rpow(Fr *rp, const Fr *r_canon, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Fr r = fr_from_canonical(*r_canon);
    Fr acc = fr_one(), b = r;
    uint32_t e = k;
#pragma unroll 1
    while (e) {
        if (e & 1u) acc = fr_mul(acc, b);
        b = fr_sqr(b);
        e >>= 1;
    }
    rp[k] = acc;
}

// Block q handles chunk (c, lo, hi) of the cells sorted by cell index
// (order[lo..hi) all have index c): E[m] = sum r^k ev_k[m], then the
// coefficients of the interpolant: a_i = h_c^-i / 64 sum_m E[m] omega_64^-im.
__global__ void __launch_bounds__(64) k_vfy_interp(Fr *part, const uint32_t *chunks, const uint32_t *order,
                                                   const Fr *ev, const Fr *rp, const Fr *roots_inv) {
    __shared__ Fr E[64];
    const uint32_t q = blockIdx.x, t = threadIdx.x;
    const uint32_t c = chunks[3 * q], lo = chunks[3 * q + 1], hi = chunks[3 * q + 2];
    Fr acc = fr_zero();
#pragma unroll 1
    for (uint32_t i = lo; i < hi; i++) {
        const uint32_t k = order[i];
        acc = fr_add(acc, fr_mul(rp[k], ev[(size_t)k * 64 + t]));
    }
    E[t] = acc;
    __syncthreads();
    Fr s = fr_zero();
#pragma unroll 1
    for (uint32_t m = 0; m < 64; m++) s = fr_add(s, fr_mul(E[m], roots_inv[128u * ((t * m) & 63u)]));
    const uint32_t hc = brev_bits(c, 7);
    const Fr scale = fr_mul(roots_inv[(hc * t) & 8191u], fr_const(cfr::INV_64_M));
    part[(size_t)q * 64 + t] = fr_mul(s, scale);
}

__global__ void k_vfy_agg(Fr *agg, const Fr *part, uint32_t nchunks) {
    const uint32_t i = threadIdx.x;
    if (i >= 64) return;
    Fr acc = fr_zero();
#pragma unroll 1
    for (uint32_t q = 0; q < nchunks; q++) acc = fr_add(acc, part[(size_t)q * 64 + i]);
    agg[i] = acc;
}

// MSM entries.  Points: [0, u) commitments, [u, u + n) proofs, [u + n, u + n
// + 64) setup [tau^i]G1.  MSM 0: sum r^k pi_k.  MSM 1: RL.
__global__ void k_vfy_entries(Fr *sc, uint32_t *pidx, uint8_t *mid, const Fr *rp, const uint32_t *cgrp_off,
                              const uint32_t *cgrp, const uint32_t *cell_idx, const Fr *roots, const Fr *agg,
                              uint32_t n, uint32_t u) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t ne = 2 * n + u + 64;
    if (e >= ne) return;
    Fr v;
    uint32_t pt, m = 1;
    if (e < n) {
        v = rp[e];
        pt = u + e;
        m = 0;
    } else if (e < n + u) {
        const uint32_t j = e - n;
        v = fr_zero();
#pragma unroll 1
        for (uint32_t i = cgrp_off[j]; i < cgrp_off[j + 1]; i++) v = fr_add(v, rp[cgrp[i]]);
        pt = j;
    } else if (e < 2 * n + u) {
        const uint32_t k = e - n - u;
        const uint32_t hc = brev_bits(cell_idx[k], 7);
        v = fr_mul(rp[k], roots[(64u * hc) & 8191u]);  // r^k h_c^64
        pt = u + k;
    } else {
        const uint32_t i = e - 2 * n - u;
        v = fr_neg(agg[i]);
        pt = u + n + i;
    }
    sc[e] = fr_to_canonical(v);
    pidx[e] = pt;
    mid[e] = (uint8_t)m;
}

// Commitment scalars of nb <= 256 blobs: sc[b * 4096 + i] = coefficient i of
// blob b's polynomial (canonical), pidx = i, mid = b; err[b] = 1 for a blob with a
// non-canonical element (as k_fk20_scalars, k_blob_to_fr.comp:23-37).
constexpr int kCommitThreads = 256;
__global__ void __launch_bounds__(kCommitThreads) k_commit_scalars(Fr *sc, uint32_t *pidx, uint8_t *mid,
                                                                  uint32_t *err, const uint8_t *blobs,
                                                                  const Fr *roots_inv) {
    __shared__ uint32_t poly[8 * 4096];
    const uint32_t blob = blockIdx.x;
    const int tid = threadIdx.x;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(blobs + (size_t)131072 * blob);
    bool bad = false;
#pragma unroll 1
    for (int i = tid; i < 4096; i += kCommitThreads) {
        uint32_t w[8];
#pragma unroll
        for (int k = 0; k < 8; k++) w[k] = src[8 * i + k];
        const Fr v = fr_from_be_words(w);
        if (!limbs_lt<8>(v.v, cfr::P)) bad = true;
        lds_store_fr(poly, 4096, i, fr_from_canonical(v));
    }
    const int any_bad = __syncthreads_or(bad ? 1 : 0);
    if (tid == 0) err[blob] = any_bad ? 1u : 0u;
    lds_ntt_dit<12>(poly, 4096, 1, roots_inv, tid, kCommitThreads);
    const Fr inv_n = fr_const(cfr::INV_4096_M);
#pragma unroll 1
    for (int i = tid; i < 4096; i += kCommitThreads) {
        const size_t e = (size_t)blob * 4096 + i;
        sc[e] = fr_to_canonical(fr_mul(lds_load_fr(poly, 4096, i), inv_n));
        pidx[e] = (uint32_t)i;
        mid[e] = (uint8_t)blob;  // one MSM per blob: at most 256 blobs per launch
    }
}

__global__ void k_compress_points(uint8_t *out, const G1A *pts, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[12];
    g1a_compress_words(w, pts[i]);
    uint32_t *dst = reinterpret_cast<uint32_t *>(out + 48ull * i);
#pragma unroll
    for (int k = 0; k < 12; k++) dst[k] = w[k];
}

}  // namespace

hipError_t launch_vfy_points(G1A *pts, const uint8_t *bytes, uint32_t n, const G1A *srs, uint32_t nsrs,
                             uint32_t *err, hipStream_t s) {
    hipLaunchKernelGGL(k_vfy_points, dim3(blocks_for(n + nsrs, 64)), dim3(64), 0, s, pts, bytes, n, srs, nsrs,
                       err);
    return hipGetLastError();
}
hipError_t launch_vfy_subgroup(const G1A *pts, uint32_t n, uint32_t *err, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_vfy_subgroup, dim3(blocks_for(n, 64)), dim3(64), 0, s, pts, n, err);
    return hipGetLastError();
}
hipError_t launch_vfy_cells(Fr *ev, const uint8_t *cells, uint32_t n, uint32_t *err, hipStream_t s) {
    hipLaunchKernelGGL(k_vfy_cells, dim3(blocks_for((uint64_t)n * 64, 256)), dim3(256), 0, s, ev, cells, n, err);
    return hipGetLastError();
}
hipError_t launch_vfy_weights(Fr *rp, const Fr *r_canon, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_vfy_rpow, dim3(blocks_for(n, 256)), dim3(256), 0, s, rp, r_canon, n);
    return hipGetLastError();
}
hipError_t launch_vfy_interp(Fr *part, Fr *agg, const uint32_t *chunks, uint32_t nchunks, const uint32_t *order,
                             const Fr *ev, const Fr *rp, const Fr *roots_inv, hipStream_t s) {
    hipLaunchKernelGGL(k_vfy_interp, dim3(nchunks), dim3(64), 0, s, part, chunks, order, ev, rp, roots_inv);
    hipLaunchKernelGGL(k_vfy_agg, dim3(1), dim3(64), 0, s, agg, part, nchunks);
    return hipGetLastError();
}
hipError_t launch_vfy_entries(Fr *sc, uint32_t *pidx, uint8_t *mid, const Fr *rp, const uint32_t *cgrp_off,
                              const uint32_t *cgrp, const uint32_t *cell_idx, const Fr *roots, const Fr *agg,
                              uint32_t n, uint32_t u, hipStream_t s) {
    hipLaunchKernelGGL(k_vfy_entries, dim3(blocks_for(2ull * n + u + 64, 256)), dim3(256), 0, s, sc, pidx, mid, rp,
                       cgrp_off, cgrp, cell_idx, roots, agg, n, u);
    return hipGetLastError();
}
hipError_t launch_commit_scalars(Fr *sc, uint32_t *pidx, uint8_t *mid, uint32_t *err, const uint8_t *blobs,
                                 uint32_t nb, const Fr *roots_inv, hipStream_t s) {
    hipLaunchKernelGGL(k_commit_scalars, dim3(nb), dim3(kCommitThreads), 0, s, sc, pidx, mid, err, blobs,
                       roots_inv);
    return hipGetLastError();
}
hipError_t launch_compress_points(uint8_t *out, const G1A *pts, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_compress_points, dim3(blocks_for(n, 64)), dim3(64), 0, s, out, pts, n);
    return hipGetLastError();
}

}  // namespace vkzg_amd
