// Variable-base multi-scalar multiplication on G1 (msm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "curve.h"

namespace vkzg_amd {

struct MsmShape {
    uint32_t ne = 0;    // entries (point index, canonical scalar, MSM id)
    uint32_t nmsm = 0;  // independent sums
    uint32_t c = 0;     // window bits
    uint32_t W = 0;     // windows of a 128-bit half scalar, ceil(129 / c)
    uint32_t bk = 0;    // buckets per window, 2^(c-1)
    uint32_t K = 0;     // buckets in all, nmsm * W * bk
    uint32_t parts = 1; // window-sum blocks per window (bk / (128 * buckets per thread))
};

MsmShape msm_shape(uint32_t ne, uint32_t nmsm);
size_t msm_scratch_bytes(const MsmShape &sh);

// out[m] = sum over entries e with mid[e] == m of [sc[e]] pts[pidx[e]], affine
// ((0, 0) for the identity).  sc: canonical scalars below r; pidx < 2^30;
// scratch: msm_scratch_bytes(sh) bytes of device memory.
hipError_t launch_msm(const MsmShape &sh, void *scratch, const G1A *pts, const uint32_t *pidx, const Fr *sc,
                      const uint8_t *mid, G1A *out, hipStream_t s);

}  // namespace vkzg_amd
