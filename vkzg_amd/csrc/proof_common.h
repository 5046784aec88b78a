// Helpers shared by the proof kernels.
#pragma once

#include "curve.h"
#include "kernels.h"
#include "ntt.h"

namespace vkzg_amd {

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

VKZG_DEV G1X g1x_shfl_xor(const G1X &p, int m) {
    G1X o;
#pragma unroll
    for (int k = 0; k < kFpLimbs; k++) {
        o.x.v[k] = __shfl_xor(p.x.v[k], m);
        o.y.v[k] = __shfl_xor(p.y.v[k], m);
        o.zz.v[k] = __shfl_xor(p.zz.v[k], m);
        o.zzz.v[k] = __shfl_xor(p.zzz.v[k], m);
    }
    return o;
}

}  // namespace vkzg_amd
