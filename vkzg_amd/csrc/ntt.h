// Radix-2 number-theoretic transforms over Fr, staged in LDS.
//
// LDS holds elements limb-major ("SoA": limb l of element idx at
// lds[l * pitch + idx]) so that consecutive lanes touch consecutive words and
// avoid bank conflicts.  Transforms are decimation-in-time: input in
// bit-reversed order, output in natural order.  `roots` is the table of
// omega_8192^k (or its inverse) in Montgomery form; a transform of size 2^LOG
// uses every 2^(13-LOG)-th entry.
#pragma once

#include "field.h"

namespace vkzg_amd {

VKZG_DEV Fr lds_load_fr(const uint32_t *lds, int pitch, int idx) {
    Fr r;
#pragma unroll
    for (int l = 0; l < 8; l++) r.v[l] = lds[l * pitch + idx];
    return r;
}

VKZG_DEV void lds_store_fr(uint32_t *lds, int pitch, int idx, const Fr &v) {
#pragma unroll
    for (int l = 0; l < 8; l++) lds[l * pitch + idx] = v.v[l];
}

VKZG_DEV uint32_t brev_bits(uint32_t x, int bits) { return __builtin_bitreverse32(x) >> (32 - bits); }

// `ncols` independent transforms of size 2^LOG at lds offsets col * 2^LOG.
// Called by all `nthreads` threads of the block; ends with a barrier.
template <int LOG>
__device__ void lds_ntt_dit(uint32_t *lds, int pitch, int ncols, const Fr *__restrict__ roots,
                            int tid, int nthreads) {
    constexpr int n = 1 << LOG;
    const int nbf = ncols * (n / 2);
#pragma unroll 1
    for (int s = 1; s <= LOG; s++) {
        const int half = 1 << (s - 1);
#pragma unroll 1
        for (int b = tid; b < nbf; b += nthreads) {
            const int col = b >> (LOG - 1);
            const int r = b & (n / 2 - 1);
            const int j = r & (half - 1);
            const int i0 = col * n + ((r >> (s - 1)) << s) + j;
            const int i1 = i0 + half;
            const Fr u = lds_load_fr(lds, pitch, i0);
            const Fr v = fr_mul(lds_load_fr(lds, pitch, i1), roots[j << (13 - s)]);
            lds_store_fr(lds, pitch, i0, fr_add(u, v));
            lds_store_fr(lds, pitch, i1, fr_sub(u, v));
        }
        __syncthreads();
    }
}

}  // namespace vkzg_amd
