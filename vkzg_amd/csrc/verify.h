// Verification and commitment kernels (verify.hip) and their launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "curve.h"

namespace vkzg_amd {

// Error bits of a verification call (the err word).
constexpr uint32_t kVfyBadPoint = 1;    // malformed commitment or proof encoding
constexpr uint32_t kVfyNotInGroup = 2;  // a commitment or proof outside G1
constexpr uint32_t kVfyBadCell = 4;     // a cell element >= r

hipError_t launch_vfy_points(G1A *pts, const uint8_t *bytes, uint32_t n, const G1A *srs, uint32_t nsrs,
                             uint32_t *err, hipStream_t s);
hipError_t launch_vfy_subgroup(const G1A *pts, uint32_t n, uint32_t *err, hipStream_t s);
hipError_t launch_vfy_cells(Fr *ev, const uint8_t *cells, uint32_t n, uint32_t *err, hipStream_t s);
hipError_t launch_vfy_weights(Fr *rp, const Fr *r_canon, uint32_t n, hipStream_t s);
hipError_t launch_vfy_interp(Fr *part, Fr *agg, const uint32_t *chunks, uint32_t nchunks, const uint32_t *order,
                             const Fr *ev, const Fr *rp, const Fr *roots_inv, hipStream_t s);
hipError_t launch_vfy_entries(Fr *sc, uint32_t *pidx, uint8_t *mid, const Fr *rp, const uint32_t *cgrp_off,
                              const uint32_t *cgrp, const uint32_t *cell_idx, const Fr *roots, const Fr *agg,
                              uint32_t n, uint32_t u, hipStream_t s);
hipError_t launch_commit_scalars(Fr *sc, uint32_t *pidx, uint8_t *mid, uint32_t *err, const uint8_t *blobs,
                                 uint32_t nb, const Fr *roots_inv, hipStream_t s);
hipError_t launch_compress_points(uint8_t *out, const G1A *pts, uint32_t n, hipStream_t s);

}  // namespace vkzg_amd
