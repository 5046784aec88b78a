// Cell-recovery kernels (the reference's recoverBatch, src/vkzg.cpp:1086-1179).
//
// The reference computes P = coset_iNTT(coset_NTT(iNTT(E * NTT(Z))) / coset_NTT(Z))
// and evaluates NTT(P).  Z(X) = z(X^64) with z of degree <= 64 in X^64, so both
// of Z's 8192-point evaluations take only 128 distinct values:
//   Z(omega^k)     = z(omega_128^(k mod 128))            (zero on missing cells)
//   Z(7 omega^k)   = z(7^64 omega_128^(k mod 128))
// which removes two of the six transforms and the per-element Fermat
// inversion of k_ew_div.comp.  The remaining four transforms T1 (inverse),
// T2 (forward), T3 (inverse) and T4 (forward) are four-step NTT_8192s,
// 8192 = 64 x 128, over x[n1 + 64 n2]:
//   pass A  y[n1][k2] = omega^(s n1 k2) sum_n2 x[n1 + 64 n2] omega_128^(s n2 k2)
//   pass B  X[k2 + 128 k1] = sum_n1 y[n1][k2] omega_64^(s n1 k1)
// Pass B of one transform and pass A of the next meet in registers: the
// pass-A column n1 = c of the next transform holds X[c + 64 n2], i.e. the
// pass-B columns k2 = c (n2 even) and k2 = c + 64 (n2 odd).  So one wave that
// owns the column pair (c, c + 64) runs both, and the four transforms take
// five kernels instead of eight passes:
//   k_rec_first          T1 pass A: cells * Z(omega^k), reduced mod r
//   k_rec_mid<0>         T1 pass B, * 7^k / 8192, T2 pass A
//   k_rec_mid<1>         T2 pass B, * 1 / Z(7 omega^k), T3 pass A
//   k_rec_mid<2>         T3 pass B, * 7^-k / 8192, T4 pass A
//   k_rec_last           T4 pass B -> canonical cell bytes
// The vanishing values each wave needs (8 of Z on the domain in the first
// kernel, 8 of 1/Z on the coset in k_rec_mid<1>) are computed by that wave.
// The default form, k_rec3_* ("three kernels" below), splits the other way
// round and needs only three kernels; the five-kernel forms remain as tuning
// options (radix 4 and 8).
//
// Work items are waves: item (blob, q) owns the pass-A columns n1 in
// [4q, 4q + 4) and the pass-B columns k2 in [4q, 4q + 4) and
// [64 + 4q, 64 + 4q + 4); 16 items per blob.  Every lane holds 8 values and
// runs radix-8 (and radix-4) butterflies in registers; values change lanes
// through a wave-private 16 KiB LDS region, with no barrier beyond the block
// (of 1, 2 or 4 such waves).
//
// Between kernels y[n1][k2] is stored per blob as S[q][n1][h][j] with
// k2 = 4q + j + 64h (q < 16, n1 < 64, h < 2, j < 4), so the 512 values a wave
// reads next are one contiguous 16 KiB block.  Cell wire position w holds the
// evaluation at natural index brev13(w); with i = n1 + 64 n2 that is
// w = 128 brev6(n1) + brev7(n2), so a pass-A column is two whole cells, and
// a pass-B column k2 is the one cell brev7(k2).
#include "field_frl.h"
#include "kernels.h"
#include "ntt.h"

namespace vkzg_amd {

namespace {

constexpr int kItemsPerBlob = 16;

// (u, t) <- (u + t w, u - t w), and the same with w = 1.
VKZG_DEV void bfly(Fr &u, Fr &t, const Fr &w) {
    const Fr tw = fr_mul(t, w);
    t = fr_sub(u, tw);
    u = fr_add(u, tw);
}
VKZG_DEV void bfly1(Fr &u, Fr &t) {
    const Fr s = fr_add(u, t);
    t = fr_sub(u, t);
    u = s;
}

// out[e] = sum_b v[b] w8^(b e) in place (natural order in and out): radix-2
// decimation in time over the bit-reversed inputs; w1, w2, w3 are w8^1..3 of
// the transform's direction.  5 products.
VKZG_DEV void dft8(Fr (&v)[8], const Fr &w1, const Fr &w2, const Fr &w3) {
    Fr x0 = v[0], x1 = v[4], x2 = v[2], x3 = v[6], x4 = v[1], x5 = v[5], x6 = v[3], x7 = v[7];
    bfly1(x0, x1);
    bfly1(x2, x3);
    bfly1(x4, x5);
    bfly1(x6, x7);
    bfly1(x0, x2);
    bfly(x1, x3, w2);
    bfly1(x4, x6);
    bfly(x5, x7, w2);
    bfly1(x0, x4);
    bfly(x1, x5, w1);
    bfly(x2, x6, w2);
    bfly(x3, x7, w3);
    v[0] = x0, v[1] = x1, v[2] = x2, v[3] = x3, v[4] = x4, v[5] = x5, v[6] = x6, v[7] = x7;
}

// out[g] = sum_a v[a] w4^(a g), w4 of the transform's direction.  1 product.
VKZG_DEV void dft4(Fr &v0, Fr &v1, Fr &v2, Fr &v3, const Fr &w4) {
    Fr x0 = v0, x1 = v2, x2 = v1, x3 = v3;
    bfly1(x0, x1);
    bfly1(x2, x3);
    bfly1(x0, x2);
    bfly(x1, x3, w4);
    v0 = x0, v1 = x1, v2 = x2, v3 = x3;
}

// A wave's 512-value LDS exchange region, as pairs of 16-byte halves; the
// position of logical index idx swaps bits 3-5 with bits 6-8 xor'ed in, which
// spreads the strided patterns below over the banks.
struct Exch {
    uint4 *base;
    VKZG_DEV static uint32_t pos(uint32_t idx) { return idx ^ (((idx >> 6) & 7u) << 3); }
    VKZG_DEV void put(uint32_t idx, const Fr &v) const {
        uint4 *p = base + 2 * pos(idx);
        p[0] = make_uint4(v.v[0], v.v[1], v.v[2], v.v[3]);
        p[1] = make_uint4(v.v[4], v.v[5], v.v[6], v.v[7]);
    }
    VKZG_DEV Fr get(uint32_t idx) const {
        const uint4 *p = base + 2 * pos(idx);
        const uint4 a = p[0], b = p[1];
        Fr r;
        r.v[0] = a.x, r.v[1] = a.y, r.v[2] = a.z, r.v[3] = a.w;
        r.v[4] = b.x, r.v[5] = b.y, r.v[6] = b.z, r.v[7] = b.w;
        return r;
    }
};

// Writes and reads of an exchange region by the lanes of one wave: the
// region is the wave's own, so a block-wide barrier (blocks of 1-4 waves)
// only has to order each wave's LDS accesses.
VKZG_DEV void exch_sync() { __syncthreads(); }

VKZG_DEV Fr ld_fr(const Fr *p) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    const uint4 a = q[0], b = q[1];
    Fr r;
    r.v[0] = a.x, r.v[1] = a.y, r.v[2] = a.z, r.v[3] = a.w;
    r.v[4] = b.x, r.v[5] = b.y, r.v[6] = b.z, r.v[7] = b.w;
    return r;
}
VKZG_DEV void st_fr(Fr *p, const Fr &v) {
    uint4 *q = reinterpret_cast<uint4 *>(p);
    q[0] = make_uint4(v.v[0], v.v[1], v.v[2], v.v[3]);
    q[1] = make_uint4(v.v[4], v.v[5], v.v[6], v.v[7]);
}

VKZG_DEV Fr fr_shfl(const Fr &v, int src) {
    Fr r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = (uint32_t)__shfl((int)v.v[i], src);
    return r;
}
VKZG_DEV Fr fr_shfl_xor(const Fr &v, int m) {
    Fr r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = (uint32_t)__shfl_xor((int)v.v[i], m);
    return r;
}
// The transform's direction: w = omega_8192^(+-1) and its table of powers.
struct Dir {
    const Fr *tab;  // tab[k] = w^k
    Fr w8_1, w8_2, w8_3;
    VKZG_DEV explicit Dir(const Fr *t) : tab(t) {
        w8_1 = tab[1024];
        w8_2 = tab[2048];
        w8_3 = tab[3072];
    }
    VKZG_DEV Fr pow(uint32_t k) const { return tab[k & 8191u]; }
};

// Pass B on the wave's 8 columns cc (k2 = 4q + (cc & 3) + 64 (cc >> 2)).
// In: lane L = 8a + cc holds v[b] = y[a + 8b][cc].  Out: lane L = 8e + cc
// holds v[f] = X[e + 8f] of its column (the pass-B output k1 = e + 8f).
VKZG_DEV void pass_b(Fr (&v)[8], const Dir &d, const Exch &x, uint32_t lane) {
    const uint32_t cc = lane & 7u, a = lane >> 3;
    dft8(v, d.w8_1, d.w8_2, d.w8_3);
#pragma unroll
    for (int e = 1; e < 8; e++) v[e] = fr_mul(v[e], d.pow(128u * a * (uint32_t)e));
    exch_sync();
#pragma unroll
    for (int e = 0; e < 8; e++) x.put(cc * 64u + (uint32_t)e * 8u + a, v[e]);
    exch_sync();
    const uint32_t e = lane >> 3;
#pragma unroll
    for (int aa = 0; aa < 8; aa++) v[aa] = x.get(cc * 64u + e * 8u + (uint32_t)aa);
    dft8(v, d.w8_1, d.w8_2, d.w8_3);
}

// Pass A on the wave's 4 columns c (n1 = 4q + c), with its four-step
// twiddle.  In: lane L = 16c + a holds v[b] = x[a + 16b] of column c.  Out:
// v[r] = y[n1][k2] with k2 = pass_a_k2(lane, r).
VKZG_DEV uint32_t pass_a_k2(uint32_t lane, int r) {
    const uint32_t m = lane & 15u;
    return (m >> 1) + 8u * (2u * (m & 1u) + (uint32_t)(r >> 2)) + 32u * (uint32_t)(r & 3);
}
VKZG_DEV void pass_a(Fr (&v)[8], const Dir &d, const Exch &x, uint32_t lane, uint32_t n1) {
    const uint32_t c = lane >> 4, a = lane & 15u;
    // DFT_8 over b, twiddle w128^(a e)
    dft8(v, d.w8_1, d.w8_2, d.w8_3);
#pragma unroll
    for (int e = 1; e < 8; e++) v[e] = fr_mul(v[e], d.pow(64u * a * (uint32_t)e));
    exch_sync();
#pragma unroll
    for (int e = 0; e < 8; e++) x.put(c * 128u + a * 8u + (uint32_t)e, v[e]);
    exch_sync();
    // DFT_16 over a = a0 + 4 a1 for e = m >> 1: first DFT_4 over a1 for this
    // lane's two a0 = 2 (m & 1) + j, twiddle w16^(a0 g)
    const uint32_t m = lane & 15u, e = m >> 1, odd = m & 1u;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const uint32_t a0 = 2u * odd + (uint32_t)j;
        Fr g0 = x.get(c * 128u + (a0 + 0u) * 8u + e);
        Fr g1 = x.get(c * 128u + (a0 + 4u) * 8u + e);
        Fr g2 = x.get(c * 128u + (a0 + 8u) * 8u + e);
        Fr g3 = x.get(c * 128u + (a0 + 12u) * 8u + e);
        dft4(g0, g1, g2, g3, d.w8_2);
        v[j * 4 + 0] = g0;
        v[j * 4 + 1] = fr_mul(g1, d.pow(512u * a0));
        v[j * 4 + 2] = fr_mul(g2, d.pow(1024u * a0));
        v[j * 4 + 3] = fr_mul(g3, d.pow(1536u * a0));
    }
    // v[4 j + g] = W[e][a0 = 2 odd + j][g]; the DFT_4 over a0 needs, for this
    // lane's g = 2 odd + t, the W of all four a0: a second exchange (half of
    // it with the pair lane m ^ 1).
    exch_sync();
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int g = 0; g < 4; g++) x.put(c * 128u + e * 16u + (2u * odd + (uint32_t)j) * 4u + (uint32_t)g, v[j * 4 + g]);
    exch_sync();
    Fr out[8];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const uint32_t g = 2u * odd + (uint32_t)t;
        Fr c0 = x.get(c * 128u + e * 16u + 0u * 4u + g);
        Fr c1 = x.get(c * 128u + e * 16u + 1u * 4u + g);
        Fr c2 = x.get(c * 128u + e * 16u + 2u * 4u + g);
        Fr c3 = x.get(c * 128u + e * 16u + 3u * 4u + g);
        dft4(c0, c1, c2, c3, d.w8_2);
        out[t * 4 + 0] = c0;
        out[t * 4 + 1] = c1;
        out[t * 4 + 2] = c2;
        out[t * 4 + 3] = c3;
    }
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = fr_mul(out[r], d.pow(n1 * pass_a_k2(lane, r)));
}

// Index of y[n1][k2] in a blob's stored intermediate S[q][n1][h][j].
VKZG_DEV uint32_t s_index(uint32_t n1, uint32_t k2) {
    return ((k2 & 63u) >> 2) * 512u + n1 * 8u + (k2 & 3u) + 4u * (k2 >> 6);
}

// Position of the k-th (from 0) set bit of w, for k < popcount(w): a binary
// search on the popcounts of the low halves.
VKZG_DEV uint32_t nth_set_bit(uint64_t w, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t half = 32; half >= 1; half >>= 1) {
        const uint32_t c = __popcll(w & ((1ull << half) - 1ull));
        if (k >= c) {
            k -= c;
            w >>= half;
            pos += half;
        }
    }
    return pos;
}

// Product over the cells c with (present[c] != 0) == want of (x - w128^brev7(c)),
// for 64 / PARTS values per wave.  Two ballots give the wanted cells as a
// 128-bit mask, the same in every lane; part p = lane % PARTS of a value
// takes the wanted cells of rank p, p + PARTS, p + 2 PARTS, ... in two
// interleaved chains, so only wanted cells cost a product, and the PARTS
// lanes of a value combine by xor shuffles, each ending with the whole
// product.  Called by whole waves (the ballots).
template <int PARTS>
VKZG_DEV Fr vanishing_product(Fr x, const uint8_t *present, bool want, const Fr *roots, uint32_t lane) {
    const uint64_t m0 = __ballot((present[lane] != 0) == want);
    const uint64_t m1 = __ballot((present[64u + lane] != 0) == want);
    const uint32_t n0 = __popcll(m0), n = n0 + __popcll(m1);
    const auto factor = [&](uint32_t t) {
        const uint32_t cell = t < n0 ? nth_set_bit(m0, t) : 64u + nth_set_bit(m1, t - n0);
        return fr_sub(x, roots[(64u * brev_bits(cell, 7)) & 8191u]);
    };
    Fr acc0 = fr_one(), acc1 = fr_one();
#pragma unroll 1
    for (uint32_t t = lane & (PARTS - 1u); t < n; t += 2 * PARTS) {
        acc0 = fr_mul(acc0, factor(t));
        if (t + PARTS < n) acc1 = fr_mul(acc1, factor(t + PARTS));
    }
    Fr prod = fr_mul(acc0, acc1);
#pragma unroll
    for (int m = 1; m < PARTS; m <<= 1) prod = fr_mul(prod, fr_shfl_xor(prod, m));
    return prod;
}

// T1 pass A.  IN_MODE 0: `in` is the blobs' cell bytes (reduced mod r) and
// the transform is inverse, on E(omega^i) Z(omega^i) with Z from the missing
// cells.  IN_MODE 1: `in` is Fr coefficients in natural order (the blob
// polynomial) and the transform is forward, with no factor: the first half
// of the evaluation of a polynomial.  Writes S.
template <int IN_MODE>
__global__ void __launch_bounds__(256, 2)
    k_rec_first(Fr *out, const void *in, const uint8_t *present, const Fr *roots, const Fr *roots_inv,
                uint32_t nitems) {
    extern __shared__ uint4 lds[];  // 16 KiB per wave
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;  // whole waves: the block barriers below see only live waves' peers exit
    const uint32_t blob = item / kItemsPerBlob, q = item % kItemsPerBlob;
    const Exch x{lds + 1024 * wave};
    const uint32_t c = lane >> 4, a = lane & 15u, n1 = 4u * q + c;
    Fr v[8];
    if (IN_MODE == 0) {
        // Z(omega^i) for i & 127 = n1 + 64 (a & 1): value v = c + 4 (a & 1)
        // of the 8 this wave computes (k = 4q + (v & 3) + 64 (v >> 2)).
        const uint32_t vv = lane >> 3;
        const uint32_t k = 4u * q + (vv & 3u) + 64u * (vv >> 2);
        const Fr zd = vanishing_product<8>(roots[(64u * k) & 8191u], present + 128ull * blob, false, roots, lane);
        // times R^2: one product takes a canonical element to E Z in Montgomery form
        const Fr zr = fr_mul(fr_shfl(zd, (int)((c + 4u * (a & 1u)) * 8u)), fr_const(cfr::R2));
        // x[a + 16b] sits at wire 128 brev6(n1) + 8 brev4(a) + brev3(b):
        // 256 contiguous bytes per lane
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(in) +
                             ((size_t)blob * 8192 + 128u * brev_bits(n1, 6) + 8u * brev_bits(a, 4)) * 8;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint32_t w = brev_bits((uint32_t)b, 3);
            const uint4 lo = *reinterpret_cast<const uint4 *>(cw + 8 * w);
            const uint4 hi = *reinterpret_cast<const uint4 *>(cw + 8 * w + 4);
            const uint32_t words[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            Fr e = fr_from_be_words(words);
            fr_reduce_once(e.v, e.v);
            fr_reduce_once(e.v, e.v);
            v[b] = fr_mul(e, zr);
        }
        pass_a(v, Dir(roots_inv), x, lane, n1);
    } else {
        const Fr *src = reinterpret_cast<const Fr *>(in) + (size_t)blob * 8192;
#pragma unroll
        for (int b = 0; b < 8; b++) v[b] = ld_fr(src + n1 + 64u * (a + 16u * (uint32_t)b));
        pass_a(v, Dir(roots), x, lane, n1);
    }
    Fr *dst = out + (size_t)blob * 8192;
#pragma unroll
    for (int r = 0; r < 8; r++) st_fr(dst + s_index(n1, pass_a_k2(lane, r)), v[r]);
}

// Pass B of T1 / T2 / T3, its pointwise factor, and pass A of the next
// transform.  STAGE 0: inverse, * 7^k / 8192, forward.  STAGE 1: forward,
// * 1 / Z(7 omega^k), inverse.  STAGE 2: inverse, * 7^-k / 8192, forward.
template <int STAGE>
__global__ void __launch_bounds__(256, 2)
    k_rec_mid(Fr *out, const Fr *in, const uint8_t *present, const Fr *roots, const Fr *roots_inv,
              const Fr *shift, uint32_t nitems) {
    extern __shared__ uint4 lds[];  // 16 KiB per wave
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;
    const uint32_t blob = item / kItemsPerBlob, q = item % kItemsPerBlob;
    const Exch x{lds + 1024 * wave};
    const Fr *src = in + (size_t)blob * 8192 + q * 512u;
    Fr v[8];
#pragma unroll
    for (int b = 0; b < 8; b++) v[b] = ld_fr(src + 64u * (uint32_t)b + lane);
    const bool b_inverse = STAGE != 1;
    pass_b(v, Dir(b_inverse ? roots_inv : roots), x, lane);
    // lane L = 8e + cc holds X[k2 + 128 (e + 8f)], k2 = 4q + (cc & 3) + 64 (cc >> 2)
    const uint32_t cc = lane & 7u, e = lane >> 3;
    const uint32_t k2 = 4u * q + (cc & 3u) + 64u * (cc >> 2);
    if (STAGE == 1) {
        // 1 / Z(7 omega^k) = K^-1 prod over present cells (y - w128^brev7(c)),
        // y = 7^64 w128^k2, K = 7^8192 - 1 (the product over all 128 cells)
        const uint32_t vv = lane >> 3;
        const uint32_t kv = 4u * q + (vv & 3u) + 64u * (vv >> 2);
        const Fr y = fr_mul(roots[(64u * kv) & 8191u], fr_const(cfr::SEVEN_POW64_M));
        const Fr p = vanishing_product<8>(y, present + 128ull * blob, true, roots, lane);
        const Fr izc = fr_mul(fr_shfl(p, (int)(cc * 8u)), fr_const(cfr::INV_SEVEN_POW8192_MINUS_1_M));
#pragma unroll
        for (int f = 0; f < 8; f++) v[f] = fr_mul(v[f], izc);
    } else {
#pragma unroll
        for (int f = 0; f < 8; f++) v[f] = fr_mul(v[f], shift[k2 + 128u * (e + 8u * (uint32_t)f)]);
    }
    // next pass A column c = cc & 3 takes x[n2] = X'[k1 = n2 >> 1] of the
    // pass-B column with h = n2 & 1 = cc >> 2
    exch_sync();
#pragma unroll
    for (int f = 0; f < 8; f++) x.put((cc & 3u) * 128u + 2u * (e + 8u * (uint32_t)f) + (cc >> 2), v[f]);
    exch_sync();
    const uint32_t c = lane >> 4, a = lane & 15u, n1 = 4u * q + c;
#pragma unroll
    for (int b = 0; b < 8; b++) v[b] = x.get(c * 128u + a + 16u * (uint32_t)b);
    pass_a(v, Dir(b_inverse ? roots : roots_inv), x, lane, n1);
    Fr *dst = out + (size_t)blob * 8192;
#pragma unroll
    for (int r = 0; r < 8; r++) st_fr(dst + s_index(n1, pass_a_k2(lane, r)), v[r]);
}

// T4 pass B: canonical big-endian bytes.  Lane L = 8e + cc writes the
// elements k1 = e + 8f of cell brev7(k2) at positions 8 brev3(e) + brev3(f):
// 256 contiguous bytes.
__global__ void __launch_bounds__(256, 2)
    k_rec_last(uint8_t *cells, const Fr *in, const Fr *roots, uint32_t nitems) {
    extern __shared__ uint4 lds[];  // 16 KiB per wave
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;
    const uint32_t blob = item / kItemsPerBlob, q = item % kItemsPerBlob;
    const Exch x{lds + 1024 * wave};
    const Fr *src = in + (size_t)blob * 8192 + q * 512u;
    Fr v[8];
#pragma unroll
    for (int b = 0; b < 8; b++) v[b] = ld_fr(src + 64u * (uint32_t)b + lane);
    pass_b(v, Dir(roots), x, lane);
    const uint32_t cc = lane & 7u, e = lane >> 3;
    const uint32_t k2 = 4u * q + (cc & 3u) + 64u * (cc >> 2);
    uint32_t *dst = reinterpret_cast<uint32_t *>(cells) +
                    ((size_t)blob * 8192 + 64u * brev_bits(k2, 7) + 8u * brev_bits(e, 3)) * 8;
    Fr one = fr_zero();
    one.v[0] = 1;
#pragma unroll
    for (int f = 0; f < 8; f++) {
        uint32_t w[8];
        fr_to_be_words(w, fr_mul(v[f], one));  // Montgomery -> canonical
        uint4 *p = reinterpret_cast<uint4 *>(dst + 8u * brev_bits((uint32_t)f, 3));
        p[0] = make_uint4(w[0], w[1], w[2], w[3]);
        p[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

// ------------------------------------------------------------------ radix 4
//
// The same five kernels with 4 values per lane: item (blob, q) for q < 32
// owns the pass-A columns n1 in {2q, 2q + 1} and the pass-B columns k2 in
// {2q, 2q + 1, 64 + 2q, 65 + 2q}, so a batch has twice as many (shorter)
// waves: more of them share each SIMD, which is what a v_mad_u64_u32 stream
// needs to issue at full rate.  The transforms run as Stockham stages
// (natural order in and out): thread j of a column of length N reads
// d[j + r N / R] (r < R), multiplies by w_(Ns R)^((j mod Ns) r), takes the
// DFT_R and writes d'[(j / Ns) Ns R + j mod Ns + r Ns], for radices
// 4, 4, 4 (N = 64) and 4, 4, 4, 2 (N = 128), through a 256-value LDS region
// per wave.  Between kernels y[n1][k2] is stored per blob as
// S4[q][n1][h][j] with k2 = 2q + j + 64h (q < 32, h, j < 2).
constexpr int kItemsPerBlob4 = 32;

// ------------------------------------------------------------ radix 4, FrL
//
// The radix-4 kernels compute in FrL (field_frl.h: 9 x 29-bit limbs,
// Montgomery radix 2^261, lazy sums).  Their intermediates in HBM (A, B) are
// packed back into 8 words (values below 2r < 2^256, normalised) in the same
// S layout; their twiddle tables are FrL tables of 48-byte entries
// (launch_frl_tables).  Value bounds, in multiples of r, are noted where they
// matter: a product's first operand stays below 36r (64r / 2 for the 1/Z
// product) and every limb below 3 * 2^30 (field_frl.h).

VKZG_DEV uint32_t s4_index(uint32_t n1, uint32_t k2) {
    return ((k2 & 63u) >> 1) * 256u + n1 * 4u + (k2 & 1u) + 2u * (k2 >> 6);
}

// A 48-byte FrL table entry (limbs 0-3, 4-7, 8 and padding).
VKZG_DEV FrL ld_frl_tab(const uint4 *tab, uint32_t k) {
    const uint4 *q = tab + 3u * k;
    const uint4 a = q[0], b = q[1];
    const uint32_t c = q[2].x;
    FrL r;
    r.v[0] = a.x, r.v[1] = a.y, r.v[2] = a.z, r.v[3] = a.w;
    r.v[4] = b.x, r.v[5] = b.y, r.v[6] = b.z, r.v[7] = b.w;
    r.v[8] = c;
    return r;
}
// Packed intermediates: 8 words per value.
VKZG_DEV FrL ld_frl(const Fr *p) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    const uint4 a = q[0], b = q[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return frl_unpack(w);
}
VKZG_DEV void st_frl(Fr *p, const FrL &v) {
    uint32_t w[8];
    frl_pack(w, v);
    uint4 *q = reinterpret_cast<uint4 *>(p);
    q[0] = make_uint4(w[0], w[1], w[2], w[3]);
    q[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// A wave's 256-value FrL exchange region (9 KiB): limbs 0-3 and 4-7 as two
// planes of 16-byte slots (bank-spread as Exch) and limb 8 as a plane of
// words.
struct Exch4 {
    uint4 *base;
    VKZG_DEV static uint32_t pos(uint32_t idx) { return idx ^ ((idx >> 5) & 7u); }
    VKZG_DEV void put(uint32_t idx, const FrL &v) const {
        const uint32_t p = pos(idx);
        base[p] = make_uint4(v.v[0], v.v[1], v.v[2], v.v[3]);
        base[256 + p] = make_uint4(v.v[4], v.v[5], v.v[6], v.v[7]);
        reinterpret_cast<uint32_t *>(base + 512)[p] = v.v[8];
    }
    VKZG_DEV FrL get(uint32_t idx) const {
        const uint32_t p = pos(idx);
        const uint4 a = base[p], b = base[256 + p];
        FrL r;
        r.v[0] = a.x, r.v[1] = a.y, r.v[2] = a.z, r.v[3] = a.w;
        r.v[4] = b.x, r.v[5] = b.y, r.v[6] = b.z, r.v[7] = b.w;
        r.v[8] = reinterpret_cast<const uint32_t *>(base + 512)[p];
        return r;
    }
};
constexpr uint32_t kExch4Uint4 = 576;  // 9 KiB per wave

// The transform's direction: tab[k] = w^k (FrL table), w4 = w^2048.
struct Dir4 {
    const uint4 *tab;
    FrL w4;
    VKZG_DEV explicit Dir4(const uint4 *t) : tab(t) { w4 = ld_frl_tab(tab, 2048); }
    VKZG_DEV FrL pow(uint32_t k) const { return ld_frl_tab(tab, k & 8191u); }
};

// out[e] = sum_b v[b] w4^(b e) (natural order in and out).  In: v0
// normalised (< V r), v1..v3 normalised below 2r.  Out: limbs below
// 3 * 2^30, values below (V + 10) r.  One product.
VKZG_DEV void dft4(FrL &v0, FrL &v1, FrL &v2, FrL &v3, const FrL &w4) {
    const FrL a = frl_add(v0, v2), b = frl_sub4(v0, v2);  // < V+2, < V+4
    const FrL c = frl_add(v1, v3), d = frl_sub4(v1, v3);  // < 4 (limbs < 2^30), < 6
    const FrL tw = frl_mul(d, w4);                        // < 2
    v0 = frl_add(a, c);                                   // < V+6
    v2 = frl_sub8(a, c);                                  // < V+10
    v1 = frl_add(b, tw);                                  // < V+6
    v3 = frl_sub4(b, tw);                                 // < V+8
}

// One radix-4 Stockham stage of a column at LDS offset `col` (length N,
// N / 4 threads, thread j): v <- d[j + r N/4] twiddled and transformed; the
// caller stores v at (j / Ns) 4 Ns + j mod Ns + r Ns.  v[0] is carried
// untwiddled, so it is normalised first; the twiddle products take v[1..3]
// back below 2r.
template <int N, int NS>
VKZG_DEV void stage4_load(FrL (&v)[4], const Dir4 &d, const Exch4 &x, uint32_t col, uint32_t j) {
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = x.get(col + j + (uint32_t)r * (N / 4));
    const uint32_t k = j % NS;
    v[0] = frl_norm(v[0]);
#pragma unroll
    for (int r = 1; r < 4; r++) v[r] = frl_mul(v[r], d.pow((8192u / (NS * 4)) * k * (uint32_t)r));
    dft4(v[0], v[1], v[2], v[3], d.w4);
}
template <int NS>
VKZG_DEV void stage4_store(const FrL (&v)[4], const Exch4 &x, uint32_t col, uint32_t j) {
    const uint32_t base = (j / NS) * NS * 4u + j % NS;
#pragma unroll
    for (int r = 0; r < 4; r++) x.put(col + base + (uint32_t)r * NS, v[r]);
}

// Pass B (DFT_64) of the wave's 4 columns cc (k2 = 2q + (cc & 1) + 64 (cc >> 1)):
// lane L = 4j + cc enters with v[r] = y[j + 16r][cc] (normalised, below 2r)
// and leaves with v[r] = X[j + 16r] of its column (below 32r).
VKZG_DEV void pass_b4(FrL (&v)[4], const Dir4 &d, const Exch4 &x, uint32_t lane) {
    const uint32_t cc = lane & 3u, j = lane >> 2, col = cc * 64u;
    dft4(v[0], v[1], v[2], v[3], d.w4);  // < 12
    exch_sync();
    stage4_store<1>(v, x, col, j);
    exch_sync();
    stage4_load<64, 4>(v, d, x, col, j);  // < 22
    exch_sync();
    stage4_store<4>(v, x, col, j);
    exch_sync();
    stage4_load<64, 16>(v, d, x, col, j);  // < 32; NS = 16: output index j + 16 r
}

// Pass A (DFT_128) of the wave's 2 columns c (n1 = 2q + c), with its
// four-step twiddle: lane L = 32c + j enters with v[r] = x[j + 32r]
// (normalised, below 2r) and leaves with v[t] = y[n1][k2] for
// k2 = pass_a4_k2(j, t), normalised and below 2r.
VKZG_DEV uint32_t pass_a4_k2(uint32_t j, int t) {
    // t = 0, 1: j' = j, outputs j', j' + 64; t = 2, 3: j' = j + 32
    return j + 32u * (uint32_t)(t >> 1) + 64u * (uint32_t)(t & 1);
}
// The same DFT_128 with output v[t] (index pass_a4_k2(j, t)) multiplied by
// fac(t) (normalised, below 2r) instead of the four-step twiddle.
template <class F>
VKZG_DEV void pass_a4f(FrL (&v)[4], const Dir4 &d, const Exch4 &x, uint32_t lane, const F &fac) {
    const uint32_t c = lane >> 5, j = lane & 31u, col = c * 128u;
    dft4(v[0], v[1], v[2], v[3], d.w4);  // < 12
    exch_sync();
    stage4_store<1>(v, x, col, j);
    exch_sync();
    stage4_load<128, 4>(v, d, x, col, j);  // < 22
    exch_sync();
    stage4_store<4>(v, x, col, j);
    exch_sync();
    stage4_load<128, 16>(v, d, x, col, j);  // < 32
    exch_sync();
    stage4_store<16>(v, x, col, j);
    exch_sync();
    // radix-2 stage (Ns = 64): j' = j and j + 32, d[j'] and d[j' + 64],
    // twiddle w128^j' = w^(64 j')
    const FrL u0 = frl_norm(x.get(col + j)), u2 = frl_norm(x.get(col + j + 32u));
    const FrL t1 = frl_mul(x.get(col + j + 64u), d.pow(64u * j));           // < 2
    const FrL t3 = frl_mul(x.get(col + j + 96u), d.pow(64u * (j + 32u)));  // < 2
    v[0] = frl_mul(frl_add(u0, t1), fac(0));   // first operands < 34
    v[1] = frl_mul(frl_sub4(u0, t1), fac(1));  // < 36
    v[2] = frl_mul(frl_add(u2, t3), fac(2));
    v[3] = frl_mul(frl_sub4(u2, t3), fac(3));
}
VKZG_DEV void pass_a4(FrL (&v)[4], const Dir4 &d, const Exch4 &x, uint32_t lane, uint32_t n1) {
    const uint32_t j = lane & 31u;
    pass_a4f(v, d, x, lane, [&](int t) { return d.pow(n1 * pass_a4_k2(j, t)); });
}

// Both vanishing values' work split, in FrL: the product over the cells c
// with (present[c] != 0) == want of (x - w128^brev7(c)), for 4 values per
// wave (16 lanes each; see vanishing_product).  x normalised below 2r.
VKZG_DEV FrL vanishing_product4(const FrL &x, const uint8_t *present, bool want, const uint4 *roots,
                                uint32_t lane) {
    constexpr uint32_t kParts = 16;
    const uint64_t m0 = __ballot((present[lane] != 0) == want);
    const uint64_t m1 = __ballot((present[64u + lane] != 0) == want);
    const uint32_t n0 = __popcll(m0), n = n0 + __popcll(m1);
    const auto factor = [&](uint32_t t) {  // below 6r, limbs below 2^29 + 2^30
        const uint32_t cell = t < n0 ? nth_set_bit(m0, t) : 64u + nth_set_bit(m1, t - n0);
        return frl_sub4(x, ld_frl_tab(roots, (64u * brev_bits(cell, 7)) & 8191u));
    };
    FrL acc0 = frl_const(cfrl::ONE_M), acc1 = frl_const(cfrl::ONE_M);
#pragma unroll 1
    for (uint32_t t = lane & (kParts - 1u); t < n; t += 2 * kParts) {
        acc0 = frl_mul(factor(t), acc0);
        if (t + kParts < n) acc1 = frl_mul(factor(t + kParts), acc1);
    }
    FrL prod = frl_mul(acc0, acc1);
#pragma unroll
    for (int m = 1; m < (int)kParts; m <<= 1) prod = frl_mul(prod, frl_shfl_xor(prod, m));
    return prod;
}

template <int IN_MODE>
__global__ void __launch_bounds__(256, 2)
    k_rec4_first(Fr *out, const void *in, const uint8_t *present, const uint4 *roots, const uint4 *roots_inv,
                 uint32_t nitems) {
    extern __shared__ uint4 lds[];  // 9 KiB per wave
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;
    const uint32_t blob = item / kItemsPerBlob4, q = item % kItemsPerBlob4;
    const Exch4 x{lds + kExch4Uint4 * wave};
    const uint32_t c = lane >> 5, j = lane & 31u, n1 = 2u * q + c;
    FrL v[4];
    if (IN_MODE == 0) {
        // Z(omega^i) for i & 127 = n1 + 64 (j & 1): value vv = c + 2 (j & 1)
        // of the 4 this wave computes (k = 2q + (vv & 1) + 64 (vv >> 1)).
        // x[j + 32r] sits at wire 128 brev6(n1) + 4 brev5(j) + brev2(r): the
        // loads go out first, the vanishing value is computed meanwhile
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(in) +
                             ((size_t)blob * 8192 + 128u * brev_bits(n1, 6) + 4u * brev_bits(j, 5)) * 8;
        uint4 raw[4][2];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t w = brev_bits((uint32_t)r, 2);
            raw[r][0] = *reinterpret_cast<const uint4 *>(cw + 8 * w);
            raw[r][1] = *reinterpret_cast<const uint4 *>(cw + 8 * w + 4);
        }
        const uint32_t vv = lane >> 4;
        const uint32_t k = 2u * q + (vv & 1u) + 64u * (vv >> 1);
        const FrL zd = vanishing_product4(ld_frl_tab(roots, (64u * k) & 8191u), present + 128ull * blob, false,
                                          roots, lane);
        // times 2^522: one product takes a plain cell element to E Z in Montgomery form
        const FrL zr = frl_mul(frl_shfl(zd, (int)((c + 2u * (j & 1u)) * 16u)), frl_const(cfrl::R2));
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint4 lo = raw[r][0], hi = raw[r][1];
            const uint32_t words[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            const Fr e = fr_from_be_words(words);  // any 256-bit value: below 3r, no reduction needed
            v[r] = frl_mul(frl_unpack(e.v), zr);
        }
        pass_a4(v, Dir4(roots_inv), x, lane, n1);
    } else {
        // Fr coefficients in the 8 x 32-bit Montgomery form (x 2^256)
        const Fr *src = reinterpret_cast<const Fr *>(in) + (size_t)blob * 8192;
        const FrL conv = frl_const(cfrl::FROM_R256);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const Fr e = ld_fr(src + n1 + 64u * (j + 32u * (uint32_t)r));
            v[r] = frl_mul(frl_unpack(e.v), conv);
        }
        pass_a4(v, Dir4(roots), x, lane, n1);
    }
    Fr *dst = out + (size_t)blob * 8192;
#pragma unroll
    for (int t = 0; t < 4; t++) st_frl(dst + s4_index(n1, pass_a4_k2(j, t)), v[t]);
}

template <int STAGE>
__global__ void __launch_bounds__(256, 2)
    k_rec4_mid(Fr *out, const Fr *in, const uint8_t *present, const uint4 *roots, const uint4 *roots_inv,
               const uint4 *shift, uint32_t nitems) {
    extern __shared__ uint4 lds[];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;
    const uint32_t blob = item / kItemsPerBlob4, q = item % kItemsPerBlob4;
    const Exch4 x{lds + kExch4Uint4 * wave};
    const Fr *src = in + (size_t)blob * 8192 + q * 256u;
    FrL v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = ld_frl(src + 64u * (uint32_t)r + lane);
    const bool b_inverse = STAGE != 1;
    // lane L = 4j + cc ends pass B with X[k2 + 128 (j + 16 r)], k2 = 2q + (cc & 1) + 64 (cc >> 1)
    const uint32_t cc = lane & 3u, jb = lane >> 2;
    const uint32_t k2 = 2u * q + (cc & 1u) + 64u * (cc >> 1);
    FrL izc;
    if (STAGE == 1) {  // before pass B: overlaps the loads above
        const uint32_t vv = lane >> 4;
        const uint32_t kv = 2u * q + (vv & 1u) + 64u * (vv >> 1);
        const FrL y = frl_mul(ld_frl_tab(roots, (64u * kv) & 8191u), frl_const(cfrl::SEVEN_POW64_M));
        const FrL p = vanishing_product4(y, present + 128ull * blob, true, roots, lane);
        izc = frl_mul(frl_shfl(p, (int)(cc * 16u)), frl_const(cfrl::INV_SEVEN_POW8192_MINUS_1_M));
    }
    pass_b4(v, Dir4(b_inverse ? roots_inv : roots), x, lane);  // < 32
    if (STAGE == 1) {
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = frl_mul(v[r], izc);  // 32 x 2 < 70
    } else {
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = frl_mul(v[r], ld_frl_tab(shift, k2 + 128u * (jb + 16u * (uint32_t)r)));
    }
    // next pass A column c = cc & 1 takes x[n2] = X'[k1 = n2 >> 1] of the
    // pass-B column with h = n2 & 1 = cc >> 1
    exch_sync();
#pragma unroll
    for (int r = 0; r < 4; r++) x.put((cc & 1u) * 128u + 2u * (jb + 16u * (uint32_t)r) + (cc >> 1), v[r]);
    exch_sync();
    const uint32_t c = lane >> 5, j = lane & 31u, n1 = 2u * q + c;
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = x.get(c * 128u + j + 32u * (uint32_t)r);
    pass_a4(v, Dir4(b_inverse ? roots : roots_inv), x, lane, n1);
    Fr *dst = out + (size_t)blob * 8192;
#pragma unroll
    for (int t = 0; t < 4; t++) st_frl(dst + s4_index(n1, pass_a4_k2(j, t)), v[t]);
}

__global__ void __launch_bounds__(256, 2)
    k_rec4_last(uint8_t *cells, const Fr *in, const uint4 *roots, uint32_t nitems) {
    extern __shared__ uint4 lds[];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;
    const uint32_t blob = item / kItemsPerBlob4, q = item % kItemsPerBlob4;
    const Exch4 x{lds + kExch4Uint4 * wave};
    const Fr *src = in + (size_t)blob * 8192 + q * 256u;
    FrL v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = ld_frl(src + 64u * (uint32_t)r + lane);
    pass_b4(v, Dir4(roots), x, lane);  // < 32
    // X[k1 = j + 16r] of cell brev7(k2) at element 4 brev4(j) + brev2(r)
    const uint32_t cc = lane & 3u, j = lane >> 2;
    const uint32_t k2 = 2u * q + (cc & 1u) + 64u * (cc >> 1);
    uint32_t *dst = reinterpret_cast<uint32_t *>(cells) +
                    ((size_t)blob * 8192 + 64u * brev_bits(k2, 7) + 4u * brev_bits(j, 4)) * 8;
    const FrL one = frl_const(cfrl::ONE_PLAIN);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        // Montgomery -> plain: at most r (r only for 0), then canonical
        Fr e;
        frl_pack(e.v, frl_mul(v[r], one));
        fr_reduce_once(e.v, e.v);
        uint32_t w[8];
        fr_to_be_words(w, e);
        uint4 *p = reinterpret_cast<uint4 *>(dst + 8u * brev_bits((uint32_t)r, 2));
        p[0] = make_uint4(w[0], w[1], w[2], w[3]);
        p[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

// ------------------------------------------------------------ three kernels
//
// A cell is the coset of mu_64 at the natural indices k = k2 + 128 k1
// (k1 < 64: cell brev7(k2), element brev6(k1)), and Z is constant on it.
// With the four-step split taken cell first (k = k2 + 128 k1, coefficient
// c = c1 + 64 c2), an 8192-point transform is a per-cell DFT_64 over k1, the
// twiddle w^(k2 c1) and a per-column DFT_128 over k2 (forward transforms run
// the same steps in reverse).  T2's per-cell DFT_64, the division by the
// per-cell constant Z(7 w^k) and T3's per-cell inverse DFT_64, with their
// twiddles, then cancel to a factor 64, and the recovery is
//   k_rec3_in    per cell: E * Z(w^k), inverse DFT_64, * w^-(k2 c1)
//   k_rec3_cols  per column c1: inverse DFT_128, * 7^c / 8192, DFT_128,
//                * 64 / Z(7 w^k), inverse DFT_128, * 7^-c / 8192, DFT_128,
//                * w^(k2 c1)
//   k_rec3_out   per cell: DFT_64, canonical cell bytes
// -- the same linear map as the five kernels (the same bytes on any input)
// with two twiddle layers and two rounds of DFT_64 fewer.  Items are waves,
// 32 per blob in each kernel: 4 cells (k2 = 4q + cc) or 2 columns
// (c1 = 2q + c).  M[c1][k2] sits per blob in 256-byte tiles of 2 columns x
// 4 cells (m3_index), so cell-wise and column-wise waves both move whole
// tiles; k_rec3_cols updates its 2 columns (one contiguous 256-value block)
// in place.  k_rec3_in also leaves 64 / Z(7 w^k) for the blob's 128 classes
// k2 at the start of the blob's B.
VKZG_DEV uint32_t m3_index(uint32_t c1, uint32_t k2) {
    return ((c1 >> 1) * 32u + (k2 >> 2)) * 8u + (c1 & 1u) * 4u + (k2 & 3u);
}

__global__ void __launch_bounds__(256, 2)
    k_rec3_in(Fr *M, Fr *izc, const uint8_t *cells, const uint8_t *present, const uint4 *roots,
              const uint4 *roots_inv, uint32_t nitems) {
    extern __shared__ uint4 lds[];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;
    const uint32_t blob = item / kItemsPerBlob4, q = item % kItemsPerBlob4;
    const Exch4 x{lds + kExch4Uint4 * wave};
    const uint32_t cc = lane & 3u, j = lane >> 2, k2 = 4u * q + cc;
    // E[k2 + 128 k1] for k1 = j + 16 r: element brev6(k1) = 4 brev4(j) + brev2(r)
    // of cell brev7(k2); the loads go out first
    const uint32_t *cw = reinterpret_cast<const uint32_t *>(cells) +
                         ((size_t)blob * 8192 + 64u * brev_bits(k2, 7) + 4u * brev_bits(j, 4)) * 8;
    uint4 raw[4][2];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t w = brev_bits((uint32_t)r, 2);
        raw[r][0] = *reinterpret_cast<const uint4 *>(cw + 8 * w);
        raw[r][1] = *reinterpret_cast<const uint4 *>(cw + 8 * w + 4);
    }
    // Z(w^k) = z(w128^k2) and 64 / Z(7 w^k) for the wave's 4 classes
    // (16 lanes each: value vv = lane >> 4, class 4q + vv)
    const uint8_t *pres = present + 128ull * blob;
    const uint32_t kv = 4u * q + (lane >> 4);
    const FrL xd = ld_frl_tab(roots, (64u * kv) & 8191u);
    const FrL zd = vanishing_product4(xd, pres, false, roots, lane);
    const FrL pc = vanishing_product4(frl_mul(xd, frl_const(cfrl::SEVEN_POW64_M)), pres, true, roots, lane);
    if ((lane & 15u) == 0)
        st_frl(izc + (size_t)blob * 8192 + kv, frl_mul(pc, frl_const(cfrl::SIXTY_FOUR_INV_SEVEN_POW8192_MINUS_1_M)));
    // times 2^522: one product takes a plain cell element to E Z in Montgomery form
    const FrL zr = frl_mul(frl_shfl(zd, (int)(16u * cc)), frl_const(cfrl::R2));
    FrL v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint4 lo = raw[r][0], hi = raw[r][1];
        const uint32_t words[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        const Fr e = fr_from_be_words(words);  // any 256-bit value: below 3r, no reduction needed
        v[r] = frl_mul(frl_unpack(e.v), zr);
    }
    pass_b4(v, Dir4(roots_inv), x, lane);  // inverse DFT_64: v[r] at c1 = j + 16 r, < 32
    Fr *dst = M + (size_t)blob * 8192;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t c1 = j + 16u * (uint32_t)r;
        st_frl(dst + m3_index(c1, k2), frl_mul(v[r], ld_frl_tab(roots_inv, (k2 * c1) & 8191u)));
    }
}

__global__ void __launch_bounds__(256, 2)
    k_rec3_cols(Fr *M, const Fr *izc, const uint4 *roots, const uint4 *roots_inv, const uint4 *shift_fwd,
                const uint4 *shift_inv, uint32_t nitems) {
    extern __shared__ uint4 lds[];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;
    const uint32_t blob = item / kItemsPerBlob4, q = item % kItemsPerBlob4;
    const Exch4 x{lds + kExch4Uint4 * wave};
    const uint32_t c = lane >> 5, j = lane & 31u, c1 = 2u * q + c;
    Fr *blk = M + (size_t)blob * 8192;
    FrL v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = ld_frl(blk + m3_index(c1, j + 32u * (uint32_t)r));
    // 64 / Z(7 w^k) at this lane's outputs k2 = pass_a4_k2(j, t)
    FrL iz[4];
#pragma unroll
    for (int t = 0; t < 4; t++) iz[t] = ld_frl(izc + (size_t)blob * 8192 + pass_a4_k2(j, t));
    const Dir4 inv(roots_inv), fwd(roots);
    // pass_a4f leaves v[t] at index pass_a4_k2(j, t) = j, j + 64, j + 32, j + 96;
    // the next one takes v[r] at j + 32 r
    const auto reorder = [&]() {
        const FrL s = v[1];
        v[1] = v[2];
        v[2] = s;
    };
    // T1: inverse DFT_128 over k2 -> c2, * 7^c / 8192
    pass_a4f(v, inv, x, lane, [&](int t) { return ld_frl_tab(shift_fwd, c1 + 64u * pass_a4_k2(j, t)); });
    reorder();
    // T2: DFT_128 over c2 -> k2, * 64 / Z(7 w^k)
    pass_a4f(v, fwd, x, lane, [&](int t) { return iz[t]; });
    reorder();
    // T3: inverse DFT_128 over k2 -> c2, * 7^-c / 8192
    pass_a4f(v, inv, x, lane, [&](int t) { return ld_frl_tab(shift_inv, c1 + 64u * pass_a4_k2(j, t)); });
    reorder();
    // T4: DFT_128 over c2 -> k2, * w^(k2 c1)
    pass_a4f(v, fwd, x, lane, [&](int t) { return fwd.pow(c1 * pass_a4_k2(j, t)); });
#pragma unroll
    for (int t = 0; t < 4; t++) st_frl(blk + m3_index(c1, pass_a4_k2(j, t)), v[t]);
}

__global__ void __launch_bounds__(256, 2)
    k_rec3_out(uint8_t *cells, const Fr *M, const uint4 *roots, uint32_t nitems) {
    extern __shared__ uint4 lds[];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t item = blockIdx.x * (blockDim.x >> 6) + wave;
    if (item >= nitems) return;
    const uint32_t blob = item / kItemsPerBlob4, q = item % kItemsPerBlob4;
    const Exch4 x{lds + kExch4Uint4 * wave};
    const uint32_t cc = lane & 3u, j = lane >> 2, k2 = 4u * q + cc;
    const Fr *src = M + (size_t)blob * 8192;
    FrL v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = ld_frl(src + m3_index(j + 16u * (uint32_t)r, k2));
    pass_b4(v, Dir4(roots), x, lane);  // DFT_64 over c1: v[r] at k1 = j + 16 r, < 32
    uint32_t *dst = reinterpret_cast<uint32_t *>(cells) +
                    ((size_t)blob * 8192 + 64u * brev_bits(k2, 7) + 4u * brev_bits(j, 4)) * 8;
    const FrL one = frl_const(cfrl::ONE_PLAIN);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        Fr e;
        frl_pack(e.v, frl_mul(v[r], one));  // Montgomery -> plain: at most r (r only for 0)
        fr_reduce_once(e.v, e.v);
        uint32_t w[8];
        fr_to_be_words(w, e);
        uint4 *p = reinterpret_cast<uint4 *>(dst + 8u * brev_bits((uint32_t)r, 2));
        p[0] = make_uint4(w[0], w[1], w[2], w[3]);
        p[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
}

// FrL table from an Fr (8 x 32-bit Montgomery) one: out[k] = in[k] 2^5, as
// 48-byte entries.
__global__ void k_frl_table(uint4 *out, const Fr *in, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Fr t = fr_mul(in[k], fr_const(cfrl::TO_R261));  // canonical x 2^261
    const FrL l = frl_unpack(t.v);
    out[3 * k] = make_uint4(l.v[0], l.v[1], l.v[2], l.v[3]);
    out[3 * k + 1] = make_uint4(l.v[4], l.v[5], l.v[6], l.v[7]);
    out[3 * k + 2] = make_uint4(l.v[8], 0u, 0u, 0u);
}

uint32_t rec_grid(uint32_t nitems, uint32_t waves) { return (nitems + waves - 1) / waves; }
constexpr uint32_t kExchBytes = 1024 * sizeof(uint4);  // LDS per wave

}  // namespace

hipError_t launch_frl_table(uint4 *out, const Fr *in, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_frl_table, dim3((n + 255) / 256), dim3(256), 0, s, out, in, n);
    return hipGetLastError();
}

hipError_t launch_rec_first(Fr *out, const void *in, int in_mode, const uint8_t *present, const RecTables &t,
                            uint32_t nblobs, uint32_t radix, uint32_t waves, hipStream_t s) {
    const bool r4 = radix != 8;  // radix 2 (the three kernels) evaluates cells like radix 4
    const uint32_t nitems = nblobs * (r4 ? kItemsPerBlob4 : kItemsPerBlob);
    const dim3 grid(rec_grid(nitems, waves)), block(64 * waves);
    const uint32_t lds = (r4 ? kExch4Uint4 * sizeof(uint4) : kExchBytes) * waves;
    if (r4 && in_mode == 0)
        hipLaunchKernelGGL(k_rec4_first<0>, grid, block, lds, s, out, in, present, t.roots4, t.roots_inv4, nitems);
    else if (r4)
        hipLaunchKernelGGL(k_rec4_first<1>, grid, block, lds, s, out, in, present, t.roots4, t.roots_inv4, nitems);
    else if (in_mode == 0)
        hipLaunchKernelGGL(k_rec_first<0>, grid, block, lds, s, out, in, present, t.roots, t.roots_inv, nitems);
    else
        hipLaunchKernelGGL(k_rec_first<1>, grid, block, lds, s, out, in, present, t.roots, t.roots_inv, nitems);
    return hipGetLastError();
}

hipError_t launch_rec_mid(Fr *out, const Fr *in, int stage, const uint8_t *present, const RecTables &t,
                          uint32_t nblobs, uint32_t radix, uint32_t waves, hipStream_t s) {
    const bool r4 = radix != 8;  // radix 2 (the three kernels) evaluates cells like radix 4
    const uint32_t nitems = nblobs * (r4 ? kItemsPerBlob4 : kItemsPerBlob);
    const dim3 grid(rec_grid(nitems, waves)), block(64 * waves);
    const uint32_t lds = (r4 ? kExch4Uint4 * sizeof(uint4) : kExchBytes) * waves;
    if (r4) {
        if (stage == 0)
            hipLaunchKernelGGL(k_rec4_mid<0>, grid, block, lds, s, out, in, present, t.roots4, t.roots_inv4,
                               t.shift_fwd4, nitems);
        else if (stage == 1)
            hipLaunchKernelGGL(k_rec4_mid<1>, grid, block, lds, s, out, in, present, t.roots4, t.roots_inv4,
                               (const uint4 *)nullptr, nitems);
        else
            hipLaunchKernelGGL(k_rec4_mid<2>, grid, block, lds, s, out, in, present, t.roots4, t.roots_inv4,
                               t.shift_inv4, nitems);
    } else {
        if (stage == 0)
            hipLaunchKernelGGL(k_rec_mid<0>, grid, block, lds, s, out, in, present, t.roots, t.roots_inv,
                               t.shift_fwd, nitems);
        else if (stage == 1)
            hipLaunchKernelGGL(k_rec_mid<1>, grid, block, lds, s, out, in, present, t.roots, t.roots_inv,
                               (const Fr *)nullptr, nitems);
        else
            hipLaunchKernelGGL(k_rec_mid<2>, grid, block, lds, s, out, in, present, t.roots, t.roots_inv,
                               t.shift_inv, nitems);
    }
    return hipGetLastError();
}

hipError_t launch_rec3(uint8_t *cells_out, Fr *M, Fr *izc, const uint8_t *cells, const uint8_t *present,
                       const RecTables &t, uint32_t nblobs, uint32_t waves, hipStream_t s, int step) {
    const uint32_t nitems = nblobs * kItemsPerBlob4;
    const dim3 grid(rec_grid(nitems, waves)), block(64 * waves);
    const uint32_t lds = kExch4Uint4 * sizeof(uint4) * waves;
    if (step == 0)
        hipLaunchKernelGGL(k_rec3_in, grid, block, lds, s, M, izc, cells, present, t.roots4, t.roots_inv4, nitems);
    else if (step == 1)
        hipLaunchKernelGGL(k_rec3_cols, grid, block, lds, s, M, izc, t.roots4, t.roots_inv4, t.shift_fwd4,
                           t.shift_inv4, nitems);
    else
        hipLaunchKernelGGL(k_rec3_out, grid, block, lds, s, cells_out, M, t.roots4, nitems);
    return hipGetLastError();
}

hipError_t launch_rec_last(uint8_t *cells, const Fr *in, const RecTables &t, uint32_t nblobs, uint32_t radix,
                           uint32_t waves, hipStream_t s) {
    const bool r4 = radix != 8;  // radix 2 (the three kernels) evaluates cells like radix 4
    const uint32_t nitems = nblobs * (r4 ? kItemsPerBlob4 : kItemsPerBlob);
    const dim3 grid(rec_grid(nitems, waves)), block(64 * waves);
    const uint32_t lds = (r4 ? kExch4Uint4 * sizeof(uint4) : kExchBytes) * waves;
    if (r4) hipLaunchKernelGGL(k_rec4_last, grid, block, lds, s, cells, in, t.roots4, nitems);
    else hipLaunchKernelGGL(k_rec_last, grid, block, lds, s, cells, in, t.roots, nitems);
    return hipGetLastError();
}

}  // namespace vkzg_amd
