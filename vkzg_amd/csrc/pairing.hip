// Optimal ate pairing on BLS12-381 for the batch cell-proof check
// e(P1, Q1) e(P2, Q2) == 1 (consensus specs, verify_cell_kzg_proof_batch;
// the reference has no verifier, include/vkzg.h:14-18).
//
// Both G2 points of the check are setup constants, so their Miller-loop lines
// are computed once, in affine coordinates, when the G2 setup is loaded
// (k_g2_lines).  A call then needs only:
//   k_pair_lines  one block per loop step: both lines evaluated at P1, P2
//                 and multiplied together (68 steps);
//   k_pair_final  one block: the Miller accumulation f <- f^2 L_s and the
//                 final exponentiation, with every Fp12 product spread over
//                 144 lanes (tower.h).
//
// Miller loop over |z| = 0xd201000000010000 from the top bit down: a
// doubling line per bit, an addition line per set bit below the top (63 + 5
// = 68 lines per point).  The line through T with slope lam (slope on the
// twist), evaluated at P = (xP, yP) in E(Fp) and multiplied by w^3:
//   l = (lam xT - yT) - lam xP w^2 + yP w^3.
// The factor w^3 lies in Fp4 and the dropped vertical lines in Fp6; the final
// exponentiation (p^12 - 1)/r kills both.  As in the oracle, the Miller value
// is not inverted for z < 0, which changes no pairing product's equality
// with 1.
//
// Final exponentiation: f^(p^6 - 1) = conj(f) / f, then ^(p^2 + 1) with the
// p^2-Frobenius, then the hard part through
//   3 (p^4 - p^2 + 1)/r = (z - 1)^2 (z + p)(z^2 + p^2 - 1) + 3
// (Hayashida, Hayasaka and Teruya, "Efficient final exponentiation via
// cyclotomic structure for pairings over families of elliptic curves", 2020).
// The result is therefore e(P, Q)^3 in the oracle's convention; a product is
// 1 exactly when its cube is (3 does not divide r).
#include "pairing.h"

#include "proof_common.h"

namespace vkzg_amd {

namespace {

constexpr uint32_t kPairThreads = kFp12Threads;

VKZG_DEV Fp frob2_const(uint32_t m) {
    switch (m) {
        case 0: return fp_const(cfp::FROB2_0);
        case 1: return fp_const(cfp::FROB2_1);
        case 2: return fp_const(cfp::FROB2_2);
        case 3: return fp_const(cfp::FROB2_3);
        case 4: return fp_const(cfp::FROB2_4);
        default: return fp_const(cfp::FROB2_5);
    }
}
VKZG_DEV Fp2 frob1_const(uint32_t m) {
    switch (m) {
        case 0: return {fp_const(cfp::FROB1_0_C0), fp_const(cfp::FROB1_0_C1)};
        case 1: return {fp_const(cfp::FROB1_1_C0), fp_const(cfp::FROB1_1_C1)};
        case 2: return {fp_const(cfp::FROB1_2_C0), fp_const(cfp::FROB1_2_C1)};
        case 3: return {fp_const(cfp::FROB1_3_C0), fp_const(cfp::FROB1_3_C1)};
        case 4: return {fp_const(cfp::FROB1_4_C0), fp_const(cfp::FROB1_4_C1)};
        default: return {fp_const(cfp::FROB1_5_C0), fp_const(cfp::FROB1_5_C1)};
    }
}

// 96-byte compressed G2 (x.c1 then x.c0, big-endian; flags in the top bits
// of the first byte as for G1) -> affine on E'.  inf[i] = 1 for the point at
// infinity; err |= 1 for a malformed encoding or an x with no point.
__global__ void __launch_bounds__(64) k_g2_decompress(G2A *out, uint32_t *inf, const uint8_t *in, uint32_t n, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(in + 96ull * i);
    uint32_t w[24];
#pragma unroll
    for (int k = 0; k < 24; k++) w[k] = __builtin_bswap32(src[k]);
    const uint32_t b0 = w[0] >> 24;
    bool bad = (b0 & 0x80u) == 0;
    G2A r;
    r.x = fp2_zero();
    r.y = fp2_zero();
    if (b0 & 0x40u) {
        uint32_t rest = w[0] & 0x3fffffffu;
#pragma unroll
        for (int k = 1; k < 24; k++) rest |= w[k];
        if (rest) bad = true;
        out[i] = r;
        inf[i] = 1;
        if (bad) atomicOr(err, 1u);
        return;
    }
    uint32_t x1w[12], x0w[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
        x1w[k] = w[11 - k];
        x0w[k] = w[23 - k];
    }
    x1w[11] &= 0x1fffffffu;
    const Fp x1 = fp_from_words(x1w), x0 = fp_from_words(x0w);
    if (!fp_limbs_lt_p(x1) || !fp_limbs_lt_p(x0)) bad = true;
    r.x = {fp_from_canonical(x0), fp_from_canonical(x1)};
    const Fp four = fp_const(cfp::FOUR_M);
    const Fp2 rhs = fp2_add(fp2_mul(fp2_sqr(r.x), r.x), Fp2{four, four});  // x^3 + 4(1 + u)
    Fp2 y;
    if (!fp2_sqrt(y, rhs)) bad = true;
    const bool want_large = (b0 & 0x20u) != 0;
    if (fp2_is_lex_largest(y) != want_large) y = fp2_neg(y);
    r.y = y;
    out[i] = r;
    inf[i] = 0;
    if (bad) atomicOr(err, 1u);
}

// Affine Miller-loop lines of nq points: lines[q][s] = (lam xT - yT, -lam)
// for the 68 steps s in loop order.  One lane per point; err |= 2 on a zero
// denominator (impossible for a point of order r).
__global__ void __launch_bounds__(64) k_g2_lines(Fp2 *lines, const G2A *qs, uint32_t nq, uint32_t *err) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const G2A Q = qs[q];
    G2A T = Q;
    Fp2 *out = lines + (size_t)q * 2 * kPairSteps;
    uint32_t s = 0;
    bool bad = false;
#pragma unroll 1
    for (int i = 62; i >= 0; i--) {
        {
            const Fp2 x2 = fp2_sqr(T.x);
            const Fp2 num = fp2_add(fp2_add(x2, x2), x2);
            const Fp2 den = fp2_add(T.y, T.y);
            if (fp2_is_zero(den)) bad = true;
            const Fp2 lam = fp2_mul(num, fp2_inv(den));
            out[2 * s] = fp2_sub(fp2_mul(lam, T.x), T.y);
            out[2 * s + 1] = fp2_neg(lam);
            s++;
            const Fp2 x3 = fp2_sub(fp2_sub(fp2_sqr(lam), T.x), T.x);
            const Fp2 y3 = fp2_sub(fp2_mul(lam, fp2_sub(T.x, x3)), T.y);
            T.x = x3;
            T.y = y3;
        }
        if ((kZAbs >> i) & 1u) {
            const Fp2 num = fp2_sub(Q.y, T.y);
            const Fp2 den = fp2_sub(Q.x, T.x);
            if (fp2_is_zero(den)) bad = true;
            const Fp2 lam = fp2_mul(num, fp2_inv(den));
            out[2 * s] = fp2_sub(fp2_mul(lam, T.x), T.y);
            out[2 * s + 1] = fp2_neg(lam);
            s++;
            const Fp2 x3 = fp2_sub(fp2_sub(fp2_sqr(lam), T.x), Q.x);
            const Fp2 y3 = fp2_sub(fp2_mul(lam, fp2_sub(T.x, x3)), T.y);
            T.x = x3;
            T.y = y3;
        }
    }
    if (bad) atomicOr(err, 2u);
}

// Block s: L[s] = l1_s(P1) * l2_s(P2), P2 negated when neg2.  A point at
// infinity contributes the line 1.
__global__ void __launch_bounds__(kPairThreads) k_pair_lines(Fp *L, const Fp2 *lines1, const Fp2 *lines2,
                                                             const G1A *p1, const G1A *p2, uint32_t neg2) {
    __shared__ Fp l1[12], l2[12], res[12];
    const uint32_t t = threadIdx.x, s = blockIdx.x;
    if (t < 12) {
        l1[t] = fp_zero();
        l2[t] = fp_zero();
    }
    __syncthreads();
    if (t < 2) {
        const Fp2 *ln = (t == 0 ? lines1 : lines2) + 2 * s;
        G1A P = t == 0 ? *p1 : *p2;
        Fp *l = t == 0 ? l1 : l2;
        if (g1a_is_identity(P)) {
            l[0] = fp_one();
        } else {
            if (t == 1 && neg2) P.y = fp_neg(P.y);
            const Fp2 c0 = ln[0], c2 = ln[1];
            l[0] = c0.c0;
            l[1] = c0.c1;
            l[4] = fp_mul(c2.c0, P.x);
            l[5] = fp_mul(c2.c1, P.x);
            l[6] = P.y;
        }
    }
    __syncthreads();
    fp12_mul_lds(res, l1, l2);
    if (t < 12) L[(size_t)s * 12 + t] = res[t];
}

__device__ void fp12_copy(Fp *dst, const Fp *src) {
    if (threadIdx.x < 12) dst[threadIdx.x] = src[threadIdx.x];
    __syncthreads();
}
// Fp12 products leave coefficients below 12p (tower.h); the helpers that
// negate or add them reduce to [0, 2p) first.
__device__ void fp12_conj_inplace(Fp *a) {
    const uint32_t t = threadIdx.x;
    if (t < 12) {
        const Fp v = fp_reduce16(a[t]);
        a[t] = ((t >> 1) & 1u) ? fp_neg(v) : v;
    }
    __syncthreads();
}
// dst = src^|z| (dst != src)
__device__ void fp12_pow_zabs(Fp *dst, const Fp *src) {
    fp12_copy(dst, src);
#pragma unroll 1
    for (int i = 62; i >= 0; i--) {
        fp12_mul_lds(dst, dst, dst);
        if ((kZAbs >> i) & 1u) fp12_mul_lds(dst, dst, src);
    }
}
__device__ void fp12_frob1(Fp *dst, const Fp *src) {
    const uint32_t t = threadIdx.x;
    if (t < 6) {
        const Fp2 a = {fp_reduce16(src[2 * t]), fp_reduce16(src[2 * t + 1])};
        const Fp2 r = fp2_mul(fp2_conj(a), frob1_const(t));
        dst[2 * t] = r.c0;
        dst[2 * t + 1] = r.c1;
    }
    __syncthreads();
}
__device__ void fp12_frob2(Fp *dst, const Fp *src) {
    const uint32_t t = threadIdx.x;
    if (t < 12) dst[t] = fp_mul(src[t], frob2_const(t >> 1));
    __syncthreads();
}

// Miller accumulation over the 68 step products L, final exponentiation,
// status[0] = 1 when the result is 1 (else 0); out12 (may be null) receives
// the result (12 Fp, Montgomery).
__global__ void __launch_bounds__(kPairThreads) k_pair_final(uint32_t *status, Fp *out12, const Fp *L) {
    __shared__ Fp f[12], g[12], h[12], A[12], B[12], C[12], D[12];
    __shared__ Fp Ls[kPairSteps * 12];  // the step products, staged in LDS (45 KiB)
    __shared__ uint32_t ok;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < kPairSteps * 12; i += blockDim.x) Ls[i] = L[i];
    fp12_set_one(f, t);
    __syncthreads();
    uint32_t s = 0;
#pragma unroll 1
    for (int i = 62; i >= 0; i--) {
        fp12_mul_lds(f, f, f);
        fp12_mul_lds(f, f, Ls + 12 * s++);
        if ((kZAbs >> i) & 1u) fp12_mul_lds(f, f, Ls + 12 * s++);
    }
    // ---- easy part: f^((p^6 - 1)(p^2 + 1))
    fp12_copy(g, f);
    fp12_conj_inplace(g);   // conj(f)
    fp12_mul_lds(h, f, g);  // f conj(f) = n0 + n1 w^2 + n2 w^4 (in Fp6)
    fp12_reduce(h);
    if (t == 0) {
        const Fp2 n0 = {h[0], h[1]}, n1 = {h[4], h[5]}, n2 = {h[8], h[9]};
        // inverse in Fp2[v]/(v^3 - xi), v = w^2
        const Fp2 c0 = fp2_sub(fp2_sqr(n0), fp2_mul_xi(fp2_mul(n1, n2)));
        const Fp2 c1 = fp2_sub(fp2_mul_xi(fp2_sqr(n2)), fp2_mul(n0, n1));
        const Fp2 c2 = fp2_sub(fp2_sqr(n1), fp2_mul(n0, n2));
        const Fp2 den = fp2_add(fp2_mul(n0, c0), fp2_mul_xi(fp2_add(fp2_mul(n2, c1), fp2_mul(n1, c2))));
        const Fp2 di = fp2_inv(den);
        const Fp2 i0 = fp2_mul(c0, di), i1 = fp2_mul(c1, di), i2 = fp2_mul(c2, di);
        for (int k = 0; k < 12; k++) h[k] = fp_zero();
        h[0] = i0.c0;
        h[1] = i0.c1;
        h[4] = i1.c0;
        h[5] = i1.c1;
        h[8] = i2.c0;
        h[9] = i2.c1;
    }
    __syncthreads();
    fp12_mul_lds(h, g, h);  // 1/f = conj(f) / (f conj(f))
    fp12_mul_lds(g, g, h);  // f^(p^6 - 1)
    fp12_frob2(h, g);
    fp12_mul_lds(f, h, g);  // F = f^((p^6 - 1)(p^2 + 1)), in the cyclotomic subgroup
    // ---- hard part: F^(3 (p^4 - p^2 + 1)/r); conj is the inverse here
    fp12_pow_zabs(A, f);
    fp12_mul_lds(A, A, f);
    fp12_conj_inplace(A);  // A = F^(z - 1)
    fp12_pow_zabs(B, A);
    fp12_mul_lds(B, B, A);
    fp12_conj_inplace(B);  // B = A^(z - 1)
    fp12_pow_zabs(C, B);
    fp12_conj_inplace(C);
    fp12_frob1(g, B);
    fp12_mul_lds(C, C, g);  // C = B^(z + p)
    fp12_pow_zabs(D, C);
    fp12_pow_zabs(A, D);    // C^(z^2)
    fp12_frob2(g, C);
    fp12_mul_lds(A, A, g);
    fp12_copy(g, C);
    fp12_conj_inplace(g);
    fp12_mul_lds(A, A, g);  // C^(z^2 + p^2 - 1)
    fp12_mul_lds(g, f, f);
    fp12_mul_lds(g, g, f);  // F^3
    fp12_mul_lds(A, A, g);
    fp12_reduce(A);
    if (t == 0) ok = 1;
    __syncthreads();
    if (t < 12) {
        const bool eq = fp_eq(A[t], t == 0 ? fp_one() : fp_zero());
        if (!eq) ok = 0;
        if (out12) out12[t] = A[t];
    }
    __syncthreads();
    if (t == 0) status[0] = ok;
}

__global__ void k_fp_canonical(Fp *a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = fp_to_canonical(a[i]);
}

}  // namespace

hipError_t launch_fp_canonical(Fp *a, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fp_canonical, dim3(blocks_for(n, 64)), dim3(64), 0, s, a, n);
    return hipGetLastError();
}

hipError_t launch_g2_decompress(G2A *out, uint32_t *inf, const uint8_t *in, uint32_t n, uint32_t *err,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_g2_decompress, dim3(blocks_for(n, 64)), dim3(64), 0, s, out, inf, in, n, err);
    return hipGetLastError();
}

hipError_t launch_g2_lines(Fp2 *lines, const G2A *qs, uint32_t nq, uint32_t *err, hipStream_t s) {
    hipLaunchKernelGGL(k_g2_lines, dim3(blocks_for(nq, 64)), dim3(64), 0, s, lines, qs, nq, err);
    return hipGetLastError();
}

hipError_t launch_pairing_check(uint32_t *status, Fp *out12, Fp *L, const Fp2 *lines1, const Fp2 *lines2,
                                const G1A *p1, const G1A *p2, bool neg2, hipStream_t s) {
    hipLaunchKernelGGL(k_pair_lines, dim3(kPairSteps), dim3(kPairThreads), 0, s, L, lines1, lines2, p1, p2,
                       neg2 ? 1u : 0u);
    hipLaunchKernelGGL(k_pair_final, dim3(1), dim3(kPairThreads), 0, s, status, out12, L);
    return hipGetLastError();
}

}  // namespace vkzg_amd
