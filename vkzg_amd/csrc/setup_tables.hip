// Prover creation: batched normalisation and the phase-A multiples table.
#include "proof_common.h"

namespace vkzg_amd {

namespace {

// Batched Jacobian -> affine (Montgomery's trick, k_normalize.comp:29-56).
// Thread t converts the points t, t+T, t+2T, ... (T = thread count) so that
// neighbouring lanes touch neighbouring points.
__global__ void __launch_bounds__(256) k_jac_to_affine(G1A *out, const G1J *in, Fp *prefix, uint64_t count, uint32_t chunk,
                                uint64_t nthreads) {
    __builtin_amdgcn_s_setprio(3);  // dependent inversion chain: issue first
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

// out[w][b] = 2^bit(w) * xext[b] for the 8192 bases.
__global__ void k_table_chains(G1J *out, const G1A *xext, WindowPlan plan) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= 8192) return;
    G1J p = g1j_from_affine(xext[b]);
#pragma unroll 1
    for (uint32_t w = 0; w < plan.W; w++) {
        out[(size_t)w * 8192 + b] = p;
#pragma unroll 1
        for (uint32_t k = 0; k < plan.c + plan.wide(w); k++) p = g1j_dbl(p);
    }
}

// The running sum d0*P, (d0+1)*P, ... with checked additions (rare fallback).
__device__ __attribute__((noinline)) void table_run_slow(G1J *scratch, Fp *prefix, Fp *acc_out,
                                                         G1A P, uint32_t d0, uint32_t S,
                                                         uint32_t thread_count, uint32_t lt) {
    G1J q = g1j_identity();
    for (int bit = 31 - __builtin_clz(d0); bit >= 0; bit--) {
        q = g1j_dbl(q);
        if ((d0 >> bit) & 1u) q = g1j_madd(q, P);
    }
    Fp acc = fp_one();
    for (uint32_t m = 0; m < S; m++) {
        const size_t slot = (size_t)m * thread_count + lt;
        scratch[slot] = q;
        prefix[slot] = acc;
        if (!fp_is_zero(q.z)) acc = fp_mul(acc, q.z);
        q = g1j_madd(q, P);
    }
    *acc_out = acc;
}

// Thread g = (b * slots + v) * nseg + seg writes the S entries of base b
// from entry v H + seg S: the multiples d in [d0, d0 + S) of P = pw[w][b] for
// the window w that owns slot v (d0 = seg S + 1, plus H in a wide window's
// second slot), a running sum normalised with one inversion per thread.
__global__ void __launch_bounds__(64) k_table_multiples(G1AT *table, const G1A *pw, WindowPlan plan,
                                  uint64_t thread_begin, uint32_t thread_count, G1J *scratch,
                                  Fp *prefix) {
    const uint32_t lt = blockIdx.x * blockDim.x + threadIdx.x;
    if (lt >= thread_count) return;
    const uint64_t g = thread_begin + lt;
    // A base's entries() split into slots of H entries: one per c-bit window,
    // two per wide window; a slot into nseg runs of S consecutive multiples.
    const uint32_t H = plan.H();
    const uint32_t S = H < 64u ? H : 64u;
    const uint32_t nseg = H / S;
    const uint32_t slots = plan.W + plan.k, narrow = plan.W - plan.k;
    const uint32_t seg = (uint32_t)(g % nseg);
    const uint64_t bv = g / nseg;  // b * slots + v
    const uint32_t v = (uint32_t)(bv % slots);
    const uint32_t b = (uint32_t)(bv / slots);
    const uint32_t w = v < narrow ? v : narrow + (v - narrow) / 2;
    const uint32_t half = v < narrow ? 0u : (v - narrow) % 2;
    const G1A P = pw[(size_t)w * 8192 + b];
    const uint32_t d0 = (half * nseg + seg) * S + 1;

    // d0 * P by double-and-add (no addition there meets equal points), then
    // the running sum; the one doubling case, 1*P + P, is taken explicitly.
    bool exc = false;
    G1J q = g1j_identity();
#pragma unroll 1
    for (int bit = 31 - __builtin_clz(d0); bit >= 0; bit--) {
        q = g1j_dbl(q);
        if ((d0 >> bit) & 1u) q = g1j_madd_fast(q, P, exc);
    }
    Fp acc = fp_one();
#pragma unroll 1
    for (uint32_t m = 0; m < S; m++) {
        const size_t slot = (size_t)m * thread_count + lt;
        scratch[slot] = q;
        prefix[slot] = acc;
        if (!fp_is_zero(q.z)) acc = fp_mul(acc, q.z);
        if (d0 + m == 1) q = g1j_dbl(q);
        else q = g1j_madd_fast(q, P, exc);
    }
    if (exc) table_run_slow(scratch, prefix, &acc, P, d0, S, thread_count, lt);
    Fp inv = fp_inv(acc);
    G1AT *dst = table + bv * H + seg * S;  // entry(w) + d0 - 1 of base b
#pragma unroll 1
    for (int m = (int)S - 1; m >= 0; m--) {
        const size_t slot = (size_t)m * thread_count + lt;
        const G1J p = scratch[slot];
        if (fp_is_zero(p.z)) {
            dst[m].p = g1a_identity();
            continue;
        }
        const Fp zi = fp_mul(inv, prefix[slot]);
        inv = fp_mul(inv, p.z);
        const Fp zi2 = fp_sqr(zi);
        G1A r;
        r.x = fp_mul(p.x, zi2);
        r.y = fp_mul(p.y, fp_mul(zi2, zi));
        dst[m].p = r;
    }
}

}  // namespace

hipError_t launch_jac_to_affine(G1A *out, const G1J *in, Fp *prefix, uint64_t count, uint32_t chunk,
                                hipStream_t s) {
    const uint64_t nthreads = (count + chunk - 1) / chunk;
    // four-wave workgroups: spread over a CU's SIMDs (one-wave workgroups
    // can stack on some SIMDs while others idle)
    hipLaunchKernelGGL(k_jac_to_affine, dim3(blocks_for(nthreads, 256)), dim3(256), 0, s, out, in,
                       prefix, count, chunk, nthreads);
    return hipGetLastError();
}

hipError_t launch_table_chains(G1J *out, const G1A *xext, WindowPlan plan, hipStream_t s) {
    hipLaunchKernelGGL(k_table_chains, dim3(8192 / 64), dim3(64), 0, s, out, xext, plan);
    return hipGetLastError();
}

namespace {
// Reference tables-file layout of the position table: entry (b = j*64 + i, d)
// at words ((b * 32 + d) * 24), x then y, each 12 little-endian words of the
// reference's Montgomery form x 2^384 mod p; the identity is (0, 0).
// pw[d * 8192 + b] = 2^(8d) xext[b] (k_table_chains with c = 8, W = 32).
__global__ void k_v1_positions(uint32_t *out, const G1A *pw) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 8192u * 32u) return;
    const uint32_t b = g / 32, d = g % 32;
    const G1A pt = pw[(size_t)d * 8192 + b];
    const Fp ref_r = fp_const(cfp::REF_R);
    uint32_t w[12];
    uint32_t *dst = out + (size_t)g * 24;
    fp_to_words(w, fp_canonical(fp_mul(pt.x, ref_r)));
#pragma unroll
    for (int k = 0; k < 12; k++) dst[k] = w[k];
    fp_to_words(w, fp_canonical(fp_mul(pt.y, ref_r)));
#pragma unroll
    for (int k = 0; k < 12; k++) dst[12 + k] = w[k];
}
}  // namespace

hipError_t launch_v1_positions(uint32_t *out, const G1A *pw, hipStream_t s) {
    hipLaunchKernelGGL(k_v1_positions, dim3(8192u * 32u / 64u), dim3(64), 0, s, out, pw);
    return hipGetLastError();
}

hipError_t launch_table_multiples(G1AT *table, const G1A *pw, WindowPlan plan,
                                  uint64_t thread_begin, uint32_t thread_count, G1J *scratch,
                                  Fp *prefix, hipStream_t s) {
    hipLaunchKernelGGL(k_table_multiples, dim3(blocks_for(thread_count, 64)), dim3(64), 0, s, table,
                       pw, plan, thread_begin, thread_count, scratch, prefix);
    return hipGetLastError();
}

}  // namespace vkzg_amd
