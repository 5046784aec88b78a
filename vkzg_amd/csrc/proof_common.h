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

// The point lane (this lane + d) holds (lanes past the wave's end read their own).
VKZG_DEV G1X g1x_shfl_down(const G1X &p, uint32_t d) {
    G1X o;
#pragma unroll
    for (int k = 0; k < kFpLimbs; k++) {
        o.x.v[k] = __shfl_down(p.x.v[k], d);
        o.y.v[k] = __shfl_down(p.y.v[k], d);
        o.zz.v[k] = __shfl_down(p.zz.v[k], d);
        o.zzz.v[k] = __shfl_down(p.zzz.v[k], d);
    }
    return o;
}

}  // namespace vkzg_amd
