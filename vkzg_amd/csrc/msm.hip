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
//   scan     bucket offsets (one block; three passes above 16k buckets)
//   scatter  point references sorted by key               (atomics)
//   buckets  one lane per bucket sums its points          (XYZZ, checked);
//            buckets of more than 24 points by a team of 64 lanes (large
//            launches: 16)
//   windows  blocks per (msm, window): sum_b (b+1) B_b by running sums over
//            segments of at most 4 (large launches: 16) buckets, then trees
//            in LDS
//   combine  four lanes per MSM: Horner over windows (c doublings each),
//            each doubling and addition shared by the quad, and the
//            conversion to affine.
// The point order inside a bucket depends on atomics, but the sum is the same
// group element, so results are deterministic after normalisation.
#include "msm.h"

#include "glv.h"
#include "proof_common.h"

namespace vkzg_amd {

namespace {

constexpr uint32_t kWinThreads = 128;
constexpr uint32_t kHotLen = 24; // longer buckets are summed by a team of lanes (k_msm_hot)
// Launch regimes.  Small sums (verification) are latency-bound: 4 buckets per
// window thread, and a hot bucket (often hundreds of points: a narrow top
// window) summed by 64 lanes.  From kBigK buckets on (commitments of many
// blobs) the launches are throughput-bound: 16 buckets per window thread
// spread its fixed costs (the offset multiple and the tree), and hot buckets
// (thousands, of some 30 points) are summed by teams of 16 lanes.
constexpr uint32_t kBigK = 1u << 18;
constexpr uint32_t kWinSegSmall = 4, kWinSegBig = 16;
constexpr uint32_t kHotBlocks = 2048;
constexpr uint32_t kScanOneBlock = 16384;  // larger scans use k_msm_scan_tiles
constexpr uint32_t kScanTile = 4096;       // buckets per block of k_msm_scan_tiles
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

// Three-pass form for large K.  Pass 1: each block scans its tile of
// kScanTile counts (4 per thread) into tile-local exclusive offsets and
// writes the tile total.
__global__ void __launch_bounds__(1024) k_msm_scan_tiles(uint32_t *off, uint32_t *tot, const uint32_t *cnt,
                                                         uint32_t K) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t lo = blockIdx.x * kScanTile + 4 * t;
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = lo + i < K ? cnt[lo + i] : 0;
        sum += v[i];
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t x = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (lo + i < K) off[lo + i] = run;
        run += v[i];
    }
    if (t == 1023) tot[blockIdx.x] = part[1023];
}

// Pass 2 (one block): exclusive scan of the ntiles <= 1024 tile totals in
// place; off[K] = the grand total.
__global__ void __launch_bounds__(1024) k_msm_scan_totals(uint32_t *off, uint32_t *tot, uint32_t ntiles,
                                                          uint32_t K) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t mine = t < ntiles ? tot[t] : 0;
    part[t] = mine;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t x = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    if (t < ntiles) tot[t] = part[t] - mine;
    if (t == 1023) off[K] = part[1023];
}

// Pass 3: add each tile's offset; cur = off.
__global__ void k_msm_scan_add(uint32_t *off, uint32_t *cur, const uint32_t *tot, uint32_t K) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const uint32_t o = off[k] + tot[k / kScanTile];
    off[k] = o;
    cur[k] = o;
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

// The point a sorted reference names: phi(P) when bit 30 is set, negated when bit 31 is.
VKZG_DEV G1A msm_point(const G1A *pts, uint32_t ref) {
    G1A p = pts[ref & kPtMask];
    if (ref & (1u << 30)) p.x = fp_mul(p.x, fp_const(cfp::BETA_M));
    return g1a_neg_if(p, (ref >> 31) != 0);
}

// A bucket summed with the checked addition (a doubling or cancellation
// inside it, flagged by the call-free loop of k_msm_buckets).
VKZG_DEV_COLD void bucket_checked(G1X *acc, const G1A *pts, const uint32_t *list, uint32_t lo, uint32_t hi) {
    G1X a = g1x_identity();
    for (uint32_t q = lo; q < hi; q++) a = g1x_madd(a, msm_point(pts, list[q]));
    *acc = a;
}

// One lane per bucket.  A bucket with more than kHotLen points (the few
// digits of a narrow top window collect most points of every scalar) goes to
// the hot list instead, for k_msm_hot.
__global__ void __launch_bounds__(64) k_msm_buckets(G1X *buckets, const uint32_t *off, const uint32_t *list, const G1A *pts,
                              uint32_t K, uint32_t *hot, uint32_t *hot_n) {
    const uint32_t key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= K) return;
    const uint32_t lo = off[key], hi = off[key + 1];
    if (hi - lo > kHotLen) {
        hot[atomicAdd(hot_n, 1u)] = key;
        return;
    }
    G1X acc = g1x_identity();
    bool exc = false;
#pragma unroll 1
    for (uint32_t q = lo; q < hi; q++) acc = g1x_madd_fast(acc, msm_point(pts, list[q]), exc);
    if (exc) bucket_checked(&acc, pts, list, lo, hi);
    buckets[key] = acc;
}

// Hot buckets, TEAM lanes each (64 / TEAM per block; blocks stride over the
// list): the team's lanes sum interleaved shares of the bucket's points, then
// a tree in LDS.  The loop bound is uniform over the block (barriers).
template <uint32_t TEAM>
__global__ void __launch_bounds__(64) k_msm_hot(G1X *buckets, const uint32_t *off, const uint32_t *list,
                                                const G1A *pts, const uint32_t *hot, const uint32_t *hot_n) {
    constexpr uint32_t kTeams = 64 / TEAM;
    __shared__ G1X sh[64];
    const uint32_t t = threadIdx.x, l = t % TEAM;
    const uint32_t n = *hot_n;
#pragma unroll 1
    for (uint32_t q0 = blockIdx.x * kTeams; q0 < n; q0 += gridDim.x * kTeams) {
        const uint32_t q = q0 + t / TEAM;
        uint32_t key = 0;
        G1X acc = g1x_identity();
        if (q < n) {
            key = hot[q];
            const uint32_t lo = off[key], hi = off[key + 1];
#pragma unroll 1
            for (uint32_t i = lo + l; i < hi; i += TEAM) acc = g1x_madd(acc, msm_point(pts, list[i]));
        }
        sh[t] = acc;
        __syncthreads();
        for (uint32_t stride = TEAM / 2; stride > 0; stride >>= 1) {
            if (l < stride) sh[t] = g1x_add(sh[t], sh[t + stride]);
            __syncthreads();
        }
        if (q < n && l == 0) buckets[key] = sh[t];
        __syncthreads();
    }
}

// [k] p for 0 < k < 2^16 (double-and-add from the top set bit, checked operations)
VKZG_DEV G1X g1x_mul_small(const G1X &p, uint32_t k) {
    G1X acc = p;
#pragma unroll 1
    for (int b = 30 - __builtin_clz(k); b >= 0; b--) {
        acc = g1x_dbl(acc);
        if ((k >> b) & 1u) acc = g1x_add(acc, p);
    }
    return acc;
}

// Block (msm, window, part): sum over the part's buckets b of (b + 1) B_b.
// Thread t owns buckets [lo, lo + seg): running sums from the top give
// R = sum B_b and T = sum (b - lo + 1) B_b, and its share is T + lo R; a
// tree in LDS adds the shares.  A window's parts (enough of them for at most
// kWinSegSmall or kWinSegBig buckets per thread) are added by k_msm_parts.
__global__ void __launch_bounds__(kWinThreads) k_msm_windows(G1X *wpart, const G1X *buckets, uint32_t bk,
                                                             uint32_t parts) {
    __shared__ G1X sh[kWinThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t win = blockIdx.x / parts, part = blockIdx.x % parts;
    const G1X *B = buckets + (size_t)win * bk;
    const uint32_t per_part = bk / parts;
    const uint32_t seg = per_part > kWinThreads ? per_part / kWinThreads : 1;
    const uint32_t lo = part * per_part + t * seg;
    G1X share = g1x_identity();
    if (t * seg < per_part) {
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
    if (t == 0) wpart[blockIdx.x] = sh[0];
}

// wsum[win] = sum of the window's parts (a tree over at most 64 parts).
__global__ void __launch_bounds__(64) k_msm_parts(G1X *wsum, const G1X *wpart, uint32_t parts) {
    __shared__ G1X sh[64];
    const uint32_t t = threadIdx.x;
    sh[t] = t < parts ? wpart[(size_t)blockIdx.x * parts + t] : g1x_identity();
    __syncthreads();
    for (uint32_t stride = 32; stride > 0; stride >>= 1) {
        if (t < stride) sh[t] = g1x_add(sh[t], sh[t + stride]);
        __syncthreads();
    }
    if (t == 0) wsum[blockIdx.x] = sh[0];
}

// Horner over the windows with the checked addition (the rare path of k_msm_combine).
VKZG_DEV_COLD void combine_checked(G1X *acc, const G1X *ws, uint32_t W, uint32_t c) {
    G1X a = ws[W - 1];
    for (uint32_t w = W - 1; w-- > 0;) {
        for (uint32_t i = 0; i < c; i++) a = g1x_dbl(a);
        a = g1x_add(a, ws[w]);
    }
    *acc = a;
}

// out[m] = sum_w 2^(c w) wsum[m W + w], affine ((0, 0) for the identity),
// by the quad m = tid / 4.  The doubling chain is this launch's critical
// path, so the loop uses the call-free addition; an equal or opposite pair
// (flagged) redoes it checked.  Quads exit together.
__global__ void __launch_bounds__(64) k_msm_combine(G1A *out, const G1X *wsum, uint32_t nmsm, uint32_t W,
                                                    uint32_t c) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t m = tid >> 2, q = tid & 3u;
    if (m >= nmsm) return;
    const G1X *ws = wsum + (size_t)m * W;
    G1X acc = ws[W - 1];
    bool exc = false;
#pragma unroll 1
    for (uint32_t w = W - 1; w-- > 0;) {
#pragma unroll 1
        for (uint32_t i = 0; i < c; i++) acc = g1x_dbl_quad(acc, q);
        acc = g1x_add_quad(acc, ws[w], q, exc);
    }
    if (exc) combine_checked(&acc, ws, W, c);
    G1A r = g1a_identity();
    if (!g1x_is_identity(acc)) {
        // x = X / ZZ = X ZZZ / (ZZ ZZZ), y = Y / ZZZ = Y ZZ / (ZZ ZZZ): one inversion
        const Fp inv = fp_inv(fp_mul(acc.zz, acc.zzz));
        r.x = fp_mul(fp_mul(acc.x, acc.zzz), inv);
        r.y = fp_mul(fp_mul(acc.y, acc.zz), inv);
    }
    if (q == 0) out[m] = r;
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
    const bool big = sh.K >= kBigK;
    const uint32_t seg = big ? kWinSegBig : kWinSegSmall;
    sh.parts = sh.bk > kWinThreads * seg ? sh.bk / (kWinThreads * seg) : 1;
    return sh;
}

size_t msm_scratch_bytes(const MsmShape &sh) {
    const size_t nh = 2ull * sh.ne;
    size_t b = 0;
    b += 16 * nh;                        // half scalars
    b += 4ull * (4 * (size_t)sh.K + 3);  // cnt (+ hot count), cur, off, hot list
    b += 4ull * (sh.K / kScanTile + 1);  // scan tile totals
    b += 4 * nh * sh.W;                  // sorted references
    b += sizeof(G1X) * (size_t)sh.K;     // buckets
    b += sizeof(G1X) * (size_t)sh.nmsm * sh.W * (sh.parts + 1);
    return b + 16 * 256;  // each of the 10 regions is rounded up to 256 bytes
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
    uint32_t *cnt = reinterpret_cast<uint32_t *>(take(4ull * (sh.K + 1)));  // cnt[K]: hot-bucket count
    uint32_t *hot = reinterpret_cast<uint32_t *>(take(4ull * sh.K));
    uint32_t *cur = reinterpret_cast<uint32_t *>(take(4ull * sh.K));
    uint32_t *off = reinterpret_cast<uint32_t *>(take(4ull * (sh.K + 1)));
    const uint32_t ntiles = (sh.K + kScanTile - 1) / kScanTile;
    uint32_t *tot = reinterpret_cast<uint32_t *>(take(4ull * (sh.K / kScanTile + 1)));
    uint32_t *list = reinterpret_cast<uint32_t *>(take(4ull * nh * sh.W));
    G1X *buckets = reinterpret_cast<G1X *>(take(sizeof(G1X) * (size_t)sh.K));
    G1X *wsum = reinterpret_cast<G1X *>(take(sizeof(G1X) * (size_t)sh.nmsm * sh.W));
    G1X *wpart = reinterpret_cast<G1X *>(take(sizeof(G1X) * (size_t)sh.nmsm * sh.W * sh.parts));
    if ((size_t)(p - static_cast<uint8_t *>(scratch)) > msm_scratch_bytes(sh)) return hipErrorInvalidValue;
    if (sh.K > kScanOneBlock && ntiles > 1024) return hipErrorInvalidValue;  // beyond 4M buckets
    hipLaunchKernelGGL(k_msm_split, dim3(blocks_for(sh.ne, 256)), dim3(256), 0, s, hs, sc, sh.ne);
    hipLaunchKernelGGL(k_msm_zero, dim3(blocks_for(sh.K + 1, 256)), dim3(256), 0, s, cnt, sh.K + 1);
    hipLaunchKernelGGL(k_msm_count, dim3(blocks_for(nh, 256)), dim3(256), 0, s, cnt, hs, mid, nh, sh.c, sh.W);
    if (sh.K <= kScanOneBlock) {
        hipLaunchKernelGGL(k_msm_scan, dim3(1), dim3(1024), 0, s, off, cur, cnt, sh.K);
    } else {
        hipLaunchKernelGGL(k_msm_scan_tiles, dim3(ntiles), dim3(1024), 0, s, off, tot, cnt, sh.K);
        hipLaunchKernelGGL(k_msm_scan_totals, dim3(1), dim3(1024), 0, s, off, tot, ntiles, sh.K);
        hipLaunchKernelGGL(k_msm_scan_add, dim3(blocks_for(sh.K, 256)), dim3(256), 0, s, off, cur, tot, sh.K);
    }
    hipLaunchKernelGGL(k_msm_scatter, dim3(blocks_for(nh, 256)), dim3(256), 0, s, list, cur, hs, mid, pidx, nh,
                       sh.c, sh.W);
    hipLaunchKernelGGL(k_msm_buckets, dim3(blocks_for(sh.K, 64)), dim3(64), 0, s, buckets, off, list, pts,
                       sh.K, hot, cnt + sh.K);
    const dim3 hot_grid(sh.K < kHotBlocks ? sh.K : kHotBlocks);
    if (sh.K >= kBigK)
        hipLaunchKernelGGL(k_msm_hot<16>, hot_grid, dim3(64), 0, s, buckets, off, list, pts, hot, cnt + sh.K);
    else
        hipLaunchKernelGGL(k_msm_hot<64>, hot_grid, dim3(64), 0, s, buckets, off, list, pts, hot, cnt + sh.K);
    hipLaunchKernelGGL(k_msm_windows, dim3(sh.nmsm * sh.W * sh.parts), dim3(kWinThreads), 0, s, wpart, buckets,
                       sh.bk, sh.parts);
    hipLaunchKernelGGL(k_msm_parts, dim3(sh.nmsm * sh.W), dim3(64), 0, s, wsum, wpart, sh.parts);
    hipLaunchKernelGGL(k_msm_combine, dim3(blocks_for(sh.nmsm * 4ull, 64)), dim3(64), 0, s, out, wsum, sh.nmsm, sh.W,
                       sh.c);
    return hipGetLastError();
}

}  // namespace vkzg_amd
