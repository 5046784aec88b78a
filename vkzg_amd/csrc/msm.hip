// Variable-base multi-scalar multiplication on G1 (Pippenger's bucket method
// with the GLV split), for the sums of batch cell-proof verification and for
// commitments.  No reference counterpart: the reference only has the FK20
// MSMs with fixed scalars or fixed points (k_phase_a.comp, k_phase_b_split.comp).
//
// Several independent sums ("MSMs") run in one launch sequence.  Entry e adds
// [scalar_e] P_{pidx_e} to MSM mid_e.  Each scalar is split k = k1 + LAMBDA k2
// (glv.h), so entry e becomes two half-entries: (P, k1) and (phi(P), k2) with
// 128-bit scalars.  Then, for windows of c bits and signed digits in
// [-2^(c-1), 2^(c-1)]:
//   count    histogram of (msm, window, |digit|) keys     (atomics)
//   scan     bucket offsets (one block)
//   scatter  point references sorted by key               (atomics)
//   buckets  one lane per bucket sums its points          (XYZZ, checked)
//   windows  one block per (msm, window): sum_b (b+1) B_b by running sums
//            over 128 segments, then a tree in LDS
//   combine  one lane per MSM: Horner over windows (c doublings each) and the
//            conversion to affine.
// The point order inside a bucket depends on atomics, but the sum is the same
// group element, so results are deterministic after normalisation.
#include "msm.h"

#include "glv.h"
#include "proof_common.h"

namespace vkzg_amd {

namespace {

constexpr uint32_t kWinThreads = 128;
constexpr uint32_t kPtMask = (1u << 30) - 1;

// Signed digit w of a 128-bit scalar (4 words) with window c; carry in/out.
VKZG_DEV int32_t digit_at(const uint32_t *s, uint32_t w, uint32_t c, uint32_t &carry) {
    const uint32_t bit = w * c;
    uint32_t v = 0;
    if (bit < 128) {
        const uint32_t wi = bit >> 5, sh = bit & 31u;
        uint64_t x = s[wi];
        if (wi + 1 < 4) x |= (uint64_t)s[wi + 1] << 32;
        v = (uint32_t)(x >> sh) & ((1u << c) - 1u);
    }
    v += carry;
    const uint32_t half = 1u << (c - 1);
    if (v > half) {
        carry = 1;
        return (int32_t)v - (int32_t)(1u << c);
    }
    carry = 0;
    return (int32_t)v;
}

__global__ void k_msm_split(uint32_t *hs, const Fr *sc, uint32_t ne) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    uint32_t k1[4], k2[4];
    glv::split(k1, k2, sc[e].v);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        hs[8 * e + i] = k1[i];
        hs[8 * e + 4 + i] = k2[i];
    }
}

__global__ void k_msm_zero(uint32_t *a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void k_msm_count(uint32_t *cnt, const uint32_t *hs, const uint8_t *mid, uint32_t nh, uint32_t c,
                            uint32_t W) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const uint32_t *s = hs + 4 * h;
    const uint32_t base = (uint32_t)mid[h >> 1] * W;
    const uint32_t bk = 1u << (c - 1);
    uint32_t carry = 0;
#pragma unroll 1
    for (uint32_t w = 0; w < W; w++) {
        const int32_t d = digit_at(s, w, c, carry);
        if (d != 0) atomicAdd(&cnt[(base + w) * bk + (uint32_t)(d < 0 ? -d : d) - 1], 1u);
    }
}

// Exclusive scan of cnt[0..K) into off[0..K] (one block of 1024 threads);
// cur = off (the scatter cursors).
__global__ void __launch_bounds__(1024) k_msm_scan(uint32_t *off, uint32_t *cur, const uint32_t *cnt,
                                                   uint32_t K) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (K + 1023) / 1024;
    const uint32_t lo = t * per, hi = lo + per < K ? lo + per : K;
    uint32_t sum = 0;
    for (uint32_t k = lo; k < hi; k++) sum += cnt[k];
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;  // exclusive prefix of this thread's range
    for (uint32_t k = lo; k < hi; k++) {
        off[k] = run;
        cur[k] = run;
        run += cnt[k];
    }
    if (t == 1023) off[K] = part[1023];
}

__global__ void k_msm_scatter(uint32_t *list, uint32_t *cur, const uint32_t *hs, const uint8_t *mid,
                              const uint32_t *pidx, uint32_t nh, uint32_t c, uint32_t W) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const uint32_t *s = hs + 4 * h;
    const uint32_t base = (uint32_t)mid[h >> 1] * W;
    const uint32_t bk = 1u << (c - 1);
    const uint32_t ref = pidx[h >> 1] | ((h & 1u) << 30);
    uint32_t carry = 0;
#pragma unroll 1
    for (uint32_t w = 0; w < W; w++) {
        const int32_t d = digit_at(s, w, c, carry);
        if (d == 0) continue;
        const uint32_t key = (base + w) * bk + (uint32_t)(d < 0 ? -d : d) - 1;
        const uint32_t pos = atomicAdd(&cur[key], 1u);
        list[pos] = ref | (d < 0 ? 0x80000000u : 0u);
    }
}

__global__ void k_msm_buckets(G1X *buckets, const uint32_t *off, const uint32_t *list, const G1A *pts,
                              uint32_t K) {
    const uint32_t key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= K) return;
    G1X acc = g1x_identity();
    const uint32_t lo = off[key], hi = off[key + 1];
#pragma unroll 1
    for (uint32_t q = lo; q < hi; q++) {
        const uint32_t ref = list[q];
        G1A p = pts[ref & kPtMask];
        if (ref & (1u << 30)) p.x = fp_mul(p.x, fp_const(cfp::BETA_M));
        p = g1a_neg_if(p, (ref >> 31) != 0);
        acc = g1x_madd(acc, p);
    }
    buckets[key] = acc;
}

// [k] p for k < 2^16 (double-and-add, checked operations)
VKZG_DEV G1X g1x_mul_small(const G1X &p, uint32_t k) {
    G1X acc = g1x_identity();
#pragma unroll 1
    for (int b = 15; b >= 0; b--) {
        acc = g1x_dbl(acc);
        if ((k >> b) & 1u) acc = g1x_add(acc, p);
    }
    return acc;
}

// Block (msm, window): sum_{b < bk} (b + 1) B_b.  Thread t owns buckets
// [t seg, (t + 1) seg): running sums from the top give R = sum B_b and
// T = sum (b - lo + 1) B_b, and its share is T + lo R.
__global__ void __launch_bounds__(kWinThreads) k_msm_windows(G1X *wsum, const G1X *buckets, uint32_t bk) {
    __shared__ G1X sh[kWinThreads];
    const uint32_t t = threadIdx.x;
    const G1X *B = buckets + (size_t)blockIdx.x * bk;
    const uint32_t seg = bk > kWinThreads ? bk / kWinThreads : 1;
    const uint32_t lo = t * seg;
    G1X share = g1x_identity();
    if (lo < bk) {
        G1X R = g1x_identity(), T = g1x_identity();
#pragma unroll 1
        for (uint32_t b = lo + seg; b-- > lo;) {
            R = g1x_add(R, B[b]);
            T = g1x_add(T, R);
        }
        share = lo ? g1x_add(T, g1x_mul_small(R, lo)) : T;
    }
    sh[t] = share;
    __syncthreads();
    for (uint32_t stride = kWinThreads / 2; stride > 0; stride >>= 1) {
        if (t < stride) sh[t] = g1x_add(sh[t], sh[t + stride]);
        __syncthreads();
    }
    if (t == 0) wsum[blockIdx.x] = sh[0];
}

// out[m] = sum_w 2^(c w) wsum[m W + w], affine ((0, 0) for the identity).
__global__ void __launch_bounds__(64) k_msm_combine(G1A *out, const G1X *wsum, uint32_t nmsm, uint32_t W,
                                                    uint32_t c) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmsm) return;
    G1X acc = wsum[(size_t)m * W + W - 1];
#pragma unroll 1
    for (uint32_t w = W - 1; w-- > 0;) {
#pragma unroll 1
        for (uint32_t i = 0; i < c; i++) acc = g1x_dbl(acc);
        acc = g1x_add(acc, wsum[(size_t)m * W + w]);
    }
    G1A r = g1a_identity();
    if (!g1x_is_identity(acc)) {
        // x = X / ZZ = X ZZZ / (ZZ ZZZ), y = Y / ZZZ = Y ZZ / (ZZ ZZZ): one inversion
        const Fp inv = fp_inv(fp_mul(acc.zz, acc.zzz));
        r.x = fp_mul(fp_mul(acc.x, acc.zzz), inv);
        r.y = fp_mul(fp_mul(acc.y, acc.zz), inv);
    }
    out[m] = r;
}

}  // namespace

MsmShape msm_shape(uint32_t ne, uint32_t nmsm) {
    MsmShape sh;
    sh.ne = ne;
    sh.nmsm = nmsm;
    // about four half-entries per bucket and window: 2^(c-1) ~ per / 4
    const uint32_t per = nmsm ? (2 * ne + nmsm - 1) / nmsm : 0;
    uint32_t c = 4;
    while (c < 13 && (1u << (c + 2)) <= per) c++;
    sh.c = c;
    sh.W = (129 + c - 1) / c;
    sh.bk = 1u << (c - 1);
    sh.K = nmsm * sh.W * sh.bk;
    return sh;
}

size_t msm_scratch_bytes(const MsmShape &sh) {
    const size_t nh = 2ull * sh.ne;
    size_t b = 0;
    b += 16 * nh;                        // half scalars
    b += 4ull * (3 * (size_t)sh.K + 1);  // cnt, cur, off
    b += 4 * nh * sh.W;                  // sorted references
    b += sizeof(G1X) * (size_t)sh.K;     // buckets
    b += sizeof(G1X) * (size_t)sh.nmsm * sh.W;
    return b + 6 * 256;
}

hipError_t launch_msm(const MsmShape &sh, void *scratch, const G1A *pts, const uint32_t *pidx, const Fr *sc,
                      const uint8_t *mid, G1A *out, hipStream_t s) {
    if (sh.ne == 0 || sh.nmsm == 0) return hipErrorInvalidValue;
    const uint32_t nh = 2 * sh.ne;
    uint8_t *p = static_cast<uint8_t *>(scratch);
    auto take = [&p](size_t bytes) {
        uint8_t *r = p;
        p += (bytes + 255) & ~size_t(255);
        return r;
    };
    uint32_t *hs = reinterpret_cast<uint32_t *>(take(16ull * nh));
    uint32_t *cnt = reinterpret_cast<uint32_t *>(take(4ull * sh.K));
    uint32_t *cur = reinterpret_cast<uint32_t *>(take(4ull * sh.K));
    uint32_t *off = reinterpret_cast<uint32_t *>(take(4ull * (sh.K + 1)));
    uint32_t *list = reinterpret_cast<uint32_t *>(take(4ull * nh * sh.W));
    G1X *buckets = reinterpret_cast<G1X *>(take(sizeof(G1X) * (size_t)sh.K));
    G1X *wsum = reinterpret_cast<G1X *>(take(sizeof(G1X) * (size_t)sh.nmsm * sh.W));
    hipLaunchKernelGGL(k_msm_split, dim3(blocks_for(sh.ne, 256)), dim3(256), 0, s, hs, sc, sh.ne);
    hipLaunchKernelGGL(k_msm_zero, dim3(blocks_for(sh.K, 256)), dim3(256), 0, s, cnt, sh.K);
    hipLaunchKernelGGL(k_msm_count, dim3(blocks_for(nh, 256)), dim3(256), 0, s, cnt, hs, mid, nh, sh.c, sh.W);
    hipLaunchKernelGGL(k_msm_scan, dim3(1), dim3(1024), 0, s, off, cur, cnt, sh.K);
    hipLaunchKernelGGL(k_msm_scatter, dim3(blocks_for(nh, 256)), dim3(256), 0, s, list, cur, hs, mid, pidx, nh,
                       sh.c, sh.W);
    hipLaunchKernelGGL(k_msm_buckets, dim3(blocks_for(sh.K, 64)), dim3(64), 0, s, buckets, off, list, pts,
                       sh.K);
    hipLaunchKernelGGL(k_msm_windows, dim3(sh.nmsm * sh.W), dim3(kWinThreads), 0, s, wsum, buckets, sh.bk);
    hipLaunchKernelGGL(k_msm_combine, dim3(blocks_for(sh.nmsm, 64)), dim3(64), 0, s, out, wsum, sh.nmsm, sh.W,
                       sh.c);
    return hipGetLastError();
}

}  // namespace vkzg_amd
