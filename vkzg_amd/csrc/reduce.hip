// FK20 phase B reduction, recombination and proof encoding.
#include "proof_common.h"

namespace vkzg_amd {

namespace {
// One output's reduction recomputed with the checked point operations.
__device__ __attribute__((noinline)) void reduce_slow(G1X *out, const G1X *B) {
    G1X run = g1x_identity(), sum = g1x_identity();
    for (int k = (int)kBuckets - 1; k >= 0; k--) {
        run = g1x_add(run, B[k]);
        sum = g1x_add(sum, run);
    }
    *out = sum;
}

// One XYZZ point per thread parked in LDS, limb-major (word k of thread t at
// k * 256 + t), so the running sum waits there while the bucket sum is
// updated: three live points would not fit 256 VGPRs without spilling.
struct G1XStash {
    uint32_t *w;  // this thread's word 0
    VKZG_DEV void put(const G1X &p) const {
#pragma unroll
        for (int i = 0; i < kFpLimbs; i++) {
            w[(0 * kFpLimbs + i) * 256] = p.x.v[i];
            w[(1 * kFpLimbs + i) * 256] = p.y.v[i];
            w[(2 * kFpLimbs + i) * 256] = p.zz.v[i];
            w[(3 * kFpLimbs + i) * 256] = p.zzz.v[i];
        }
    }
    VKZG_DEV G1X get() const {
        G1X p;
#pragma unroll
        for (int i = 0; i < kFpLimbs; i++) {
            p.x.v[i] = w[(0 * kFpLimbs + i) * 256];
            p.y.v[i] = w[(1 * kFpLimbs + i) * 256];
            p.zz.v[i] = w[(2 * kFpLimbs + i) * 256];
            p.zzz.v[i] = w[(3 * kFpLimbs + i) * 256];
        }
        return p;
    }
};

// The lane index, recomputed where it is called: the asm is opaque, so what
// derives from it is not hoisted out of a loop and held (or spilled) across
// the loop's point chains.
VKZG_DEV uint32_t lane_here() {
    uint32_t l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

// A uniform value made opaque at this point (same reason).
VKZG_DEV uint32_t uniform_here(uint32_t v) {
    asm volatile("" : "+s"(v));
    return v;
}

// One point operation of a bucket range: on one lane (Q = 1), or shared by
// the four lanes of a quad that all hold the operands (Q = 4: its products
// in dependent rounds of up to four, g1x_add_quad / g1x_dbl_quad).
template <uint32_t Q>
VKZG_DEV G1X range_add(const G1X &a, const G1X &b, bool &exc) {
    if constexpr (Q == 4) return g1x_add_quad(a, b, lane_here() & 3u, exc);
    else return g1x_add_fast(a, b, exc);
}
template <uint32_t Q>
VKZG_DEV G1X range_dbl(const G1X &a) {
    if constexpr (Q == 4) return g1x_dbl_quad(a, lane_here() & 3u);
    else return g1x_dbl_core(a);
}

// cvals[blob][a] = sum_{k=1..kBuckets} k * buckets[blob][a][k-1]
// (field.glsl.inc:478-517).  LANES lanes share an output: 4, 8, 16 or 32
// bucket ranges of one lane each, or (LANES = 64) 16 ranges of a quad each.
// Range l owns buckets [per l, per l + per), per = kBuckets / ranges:
// rs_l = sum_i (i+1) B[per l + i] by running sums, s_l = sum_i B[per l + i];
// the output is sum_l (rs_l + per l * s_l).  More ranges shorten each
// range's dependent chain and add total work (the scan over the ranges);
// quads shorten it again (an addition waits on four rounds of products
// instead of fourteen products) at four times the lanes, for batches that
// leave SIMDs idle.  The running sum rs_l stays in LDS between its updates
// (G1XStash).
//
// Persistent: every wave takes items of 64 / LANES consecutive outputs from
// the counter q[0] (cleared by k_glv_phi) until the queue is drained and
// leaves after one failed take.  Flagged outputs (a doubling or cancellation
// met by the fast additions) are recomputed by k_reduce_fix.
template <int LANES>
__global__ void __launch_bounds__(256, 2)
    k_bucket_reduce(G1X *cvals, uint8_t *flags, const G1X *buckets, uint32_t nout, uint32_t *q,
                    uint32_t force_checked) {
    constexpr uint32_t Q = LANES == 64 ? 4u : 1u;  // lanes per bucket range
    constexpr uint32_t ranges = LANES / Q;
    constexpr uint32_t per = kBuckets / ranges;
    constexpr uint32_t teams = 64 / LANES;
    __shared__ uint32_t stash_words[4 * kFpLimbs * 256];
    const uint32_t wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
    const uint32_t nitems = (nout + teams - 1) / teams;
    // Lane-dependent values are rederived per item from lane_here(), so none
    // is held across the chains.
#pragma unroll 1
    for (;;) {
        const uint32_t lane = lane_here();
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(q, 1u);
        item = __shfl(item, 0);
        if (item >= nitems) break;
        const uint32_t l = (lane / Q) % ranges;
        const uint32_t o = item * teams + lane / LANES;
        const G1XStash stash{stash_words + wave_base + lane};
        G1X s = g1x_identity();
        stash.put(g1x_identity());
        bool exc = uniform_here(force_checked) != 0;
        if (o < nout) {
            const G1X *B = buckets + (size_t)o * kBuckets + l * per;
            // The item lists leave some buckets empty (the identity), the
            // same ones for every blob.  An empty bucket right after the
            // range's first non-empty one leaves s unchanged while the
            // running sum still equals s: that update is a doubling, taken
            // here rather than flagged.  `single` records running sum == s.
            bool single = false;
#pragma unroll 1
            for (int i = per - 1; i >= 0; i--) {
                const G1X b = B[i];
                const bool empty = fp_is_zero(b.zz);
                s = range_add<Q>(s, b, exc);
                G1X rs = stash.get();
                if (empty && single) {
                    rs = range_dbl<Q>(rs);
                    single = false;
                } else {
                    single = fp_is_zero(rs.zz) && !fp_is_zero(s.zz);
                    rs = range_add<Q>(rs, s, exc);
                }
                stash.put(rs);
            }
        }
        // sum_l per l s_l = per * sum_{m >= 1} S_m, with the suffix sums
        // S_m = sum_{l >= m} s_l of the team (a log-step scan over shuffles),
        // then log2(per) doublings: no range-dependent double-and-add chain.
        // Every condition below is uniform in a quad.
        const uint32_t l2 = (lane_here() / Q) % ranges;
        G1X x = s;
#pragma unroll 1
        for (uint32_t d = 1; d < ranges; d <<= 1) {
            const G1X y = g1x_shfl_down(x, d * Q);
            if (l2 + d < ranges) x = range_add<Q>(x, y, exc);
        }
        if (l2 == 0) x = g1x_identity();
#pragma unroll 1
        for (uint32_t k = 1; k < per; k <<= 1) x = range_dbl<Q>(x);
        G1X v = range_add<Q>(stash.get(), x, exc);
#pragma unroll 1
        for (uint32_t m = 1; m < ranges; m <<= 1) v = range_add<Q>(v, g1x_shfl_xor(v, m * Q), exc);
        const uint64_t flagged = __ballot(exc);
        const uint32_t lane2 = lane_here();
        const uint32_t o2 = item * teams + lane2 / LANES;
        const uint64_t team_mask = LANES == 64 ? ~0ull : ((1ull << LANES) - 1);
        const bool redo = ((flagged >> (lane2 / LANES * LANES)) & team_mask) != 0;
        if (o2 < nout && lane2 % LANES == 0) {
            cvals[o2] = v;
            flags[o2] = redo ? 1 : 0;
        }
    }
}

// Recompute the flagged outputs of k_bucket_reduce with checked additions.
__global__ void __launch_bounds__(256) k_reduce_fix(G1X *cvals, const uint8_t *flags, const G1X *buckets, uint32_t nout) {
    const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout || !flags[o]) return;
    reduce_slow(&cvals[o], buckets + (size_t)o * kBuckets);
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
// compressed and written at brev7(k) (k_compress_proofs.comp).  The cyclic
// half comes from its two sub-convolutions: C+[a] = Yc[a] + Yn[a] and
// C+[a+32] = Yc[a] - Yn[a] for a < 32 (cvals[..][a] = Yc[a],
// cvals[..][32+a] = Yn[a], cvals[..][64+a] = C-[a]).
//
// Q = 4 (while the quads fit one wave per SIMD): the four lanes of a quad
// share the three additions (g1x_add_quad, four rounds of products each);
// a flagged doubling or cancellation redoes them checked, uniformly in the
// quad.  The inversion is the same on every lane of the quad.
template <uint32_t Q>
__global__ void __launch_bounds__(64 * Q) k_finalize(uint8_t *proofs, const G1X *cvals, uint32_t nblobs,
                                                     uint32_t force_checked) {
    __builtin_amdgcn_s_setprio(3);  // dependent inversion chain: issue first
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = tid / Q, qd = tid % Q;
    if (g >= nblobs * 64) return;
    const uint32_t blob = g / 64, a = g % 64;
    const G1X yc = cvals[blob * 128 + (a & 31u)];
    const G1X yn = cvals[blob * 128 + 32 + (a & 31u)];
    const G1X cm = cvals[blob * 128 + a + 64];
    G1X s0, s1;
    if constexpr (Q == 4) {
        bool exc = force_checked != 0;
        const G1X cp = g1x_add_quad(yc, a < 32 ? yn : g1x_neg(yn), qd, exc);
        s0 = g1x_add_quad(cp, cm, qd, exc);
        s1 = g1x_add_quad(cp, g1x_neg(cm), qd, exc);
        if (exc) {
            const G1X cc = g1x_add(yc, a < 32 ? yn : g1x_neg(yn));
            s0 = g1x_add(cc, cm);
            s1 = g1x_add(cc, g1x_neg(cm));
        }
    } else {
        const G1X cp = g1x_add(yc, a < 32 ? yn : g1x_neg(yn));
        s0 = g1x_add(cp, cm);
        s1 = g1x_add(cp, g1x_neg(cm));
    }
    if (qd != 0) return;  // the inversion and the stores: one lane of the quad
    G1A p0, p1;
    g1x_pair_to_affine(p0, p1, s0, s1);
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

}  // namespace

hipError_t launch_bucket_reduce(G1X *cvals, uint8_t *flags, const G1X *buckets, uint32_t nblobs,
                                uint32_t lanes, uint32_t grid, uint32_t *queue, uint32_t force_checked,
                                hipStream_t s) {
    const uint32_t nout = nblobs * 128;
    const uint32_t teams = 64 / (lanes == 4 || lanes == 16 || lanes == 32 || lanes == 64 ? lanes : 8);
    const uint32_t waves = (nout + teams - 1) / teams;
    if (grid > (waves + 3) / 4) grid = (waves + 3) / 4;
    switch (lanes) {
        case 4:
            hipLaunchKernelGGL(k_bucket_reduce<4>, dim3(grid), dim3(256), 0, s, cvals, flags, buckets, nout,
                               queue, force_checked);
            break;
        case 16:
            hipLaunchKernelGGL(k_bucket_reduce<16>, dim3(grid), dim3(256), 0, s, cvals, flags, buckets, nout,
                               queue, force_checked);
            break;
        case 32:
            hipLaunchKernelGGL(k_bucket_reduce<32>, dim3(grid), dim3(256), 0, s, cvals, flags, buckets, nout,
                               queue, force_checked);
            break;
        case 64:
            hipLaunchKernelGGL(k_bucket_reduce<64>, dim3(grid), dim3(256), 0, s, cvals, flags, buckets, nout,
                               queue, force_checked);
            break;
        default:
            hipLaunchKernelGGL(k_bucket_reduce<8>, dim3(grid), dim3(256), 0, s, cvals, flags, buckets, nout,
                               queue, force_checked);
            break;
    }
    hipLaunchKernelGGL(k_reduce_fix, dim3(blocks_for(nout, 256)), dim3(256), 0, s, cvals, flags, buckets,
                       nout);
    return hipGetLastError();
}

hipError_t launch_finalize(uint8_t *proofs, const G1X *cvals, uint32_t nblobs, bool quad,
                           uint32_t force_checked, hipStream_t s) {
    if (quad)
        hipLaunchKernelGGL(k_finalize<4>, dim3(blocks_for(nblobs * 256ull, 256)), dim3(256), 0, s, proofs,
                           cvals, nblobs, force_checked);
    else
        hipLaunchKernelGGL(k_finalize<1>, dim3(blocks_for(nblobs * 64ull, 64)), dim3(64), 0, s, proofs,
                           cvals, nblobs, force_checked);
    return hipGetLastError();
}

}  // namespace vkzg_amd
