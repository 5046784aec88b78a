// Pairing kernels (pairing.hip) and their launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "tower.h"

namespace vkzg_amd {

constexpr uint32_t kPairSteps = 68;  // Miller-loop lines per G2 point: 63 doublings + 5 additions

// n 96-byte compressed G2 points -> affine; inf[i] = 1 at infinity; err |= 1 when malformed.
hipError_t launch_g2_decompress(G2A *out, uint32_t *inf, const uint8_t *in, uint32_t n, uint32_t *err,
                                hipStream_t s);
// Miller-loop lines of nq points: lines[(q * kPairSteps + s) * 2 + {0, 1}].
hipError_t launch_g2_lines(Fp2 *lines, const G2A *qs, uint32_t nq, uint32_t *err, hipStream_t s);
// status[0] = [e(P1, Q1) e(+-P2, Q2) == 1] for the G2 points whose lines are
// given (P2 negated when neg2); out12 (may be null) gets the product of the
// pairings, cubed; L: kPairSteps * 12 Fp of device scratch.
hipError_t launch_pairing_check(uint32_t *status, Fp *out12, Fp *L, const Fp2 *lines1, const Fp2 *lines2,
                                const G1A *p1, const G1A *p2, bool neg2, hipStream_t s);

// Montgomery -> canonical integers (28-bit limbs), in place.
hipError_t launch_fp_canonical(Fp *a, uint32_t n, hipStream_t s);

}  // namespace vkzg_amd
