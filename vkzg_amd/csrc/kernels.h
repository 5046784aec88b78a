// Host-side launch wrappers for the gfx950 kernels, and the device layouts
// they share.  Every wrapper enqueues on `stream` and returns the launch
// status; none synchronises.
//
// Layouts (all arrays of structs, little-endian 32-bit limbs, Montgomery form
// unless stated):
//   Fr roots[8192]            omega_8192^k;  roots_inv: omega_8192^-k
//   Fr shift_fwd[8192]        7^k / 8192;    shift_inv: 7^-k / 8192
//   G1A srs[4096]             the G1 monomial setup, affine
//   G1A xext[128][64]         x_ext_fft[j][i] (FK20 setup in the Fourier domain)
//   G1AT table[128][64][E]    for the window plan (c, W, k) (WindowPlan):
//                             table[j][i][entry(w) + d - 1] =
//                             d * 2^bit(w) * xext[j][i], d <= 2^(width(w)-1),
//                             one 128-byte line per entry
//   Fr scal[B][128][64]       phase-A scalars, CANONICAL (not Montgomery), as
//                             sign (bit 255) and magnitude (< 2^254)
//   G1X u[B][128][G]          phase-A partial sums, one per window group g < G
//   G1J lad[B][16][128]       2^(8d) * f[j], d < kLadPos, for the folds f of u:
//                             Fc (j < 32), Fn (32 <= j < 64), F- (j >= 64);
//                             normalised into G1A lad_aff with the same
//                             layout; G1A lad_phi: phi(lad_aff), where
//                             phi(x, y) = (beta x, y) = lambda * P (GLV);
//                             small batches: G1X (XYZZ) points and their
//                             images in the same two buffers' memory
//   G1X buckets[B][128][128]  phase-B buckets, [output a][bucket k]
//   G1X cvals[B][128]         Yc (a < 32), Yn (32 <= a < 64), C- (a >= 64)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vkzg_amd {

// Phase-B scalars are split as k = k1 + lambda k2 with the G1 endomorphism
// phi = [lambda], balanced and signed (|k1|, |k2| < 2^126.5), and the
// magnitudes written in signed kDigitBits-bit digits: kLadPos of them per half
// (16 bytes, the top one below 2^6.5, so no carry position), so each ladder
// needs 2^(kDigitBits d) u for d < kLadPos only.
// Digit magnitudes 1..kBuckets are the buckets of every phase-B output.
// 7-bit digits (64 buckets) halve the bucket reduction but cost phase B more
// than that on the MI355X (19.6-19.8k against 19.9-20.0k blobs/s at 256
// blobs, on one box).
constexpr uint32_t kDigitBits = 8;
constexpr uint32_t kLadPos = (127 + kDigitBits - 1) / kDigitBits;
static_assert(kDigitBits == 8 && kLadPos == 16, "the top byte of a magnitude below 2^126.5 never carries");
constexpr uint32_t kBuckets = 1u << (kDigitBits - 1);
// The ladder may be split into 1, 2 or 4 segments; segment m of M covers
// positions [lad_start(M, m), lad_start(M, m + 1)) (chains of 7 and 3 steps
// of kDigitBits doublings for M = 2 and 4), and phase A then also
// computes 2^(kDigitBits start) u directly.
constexpr uint32_t kMaxLadSplits = 4;
__host__ __device__ constexpr uint32_t lad_start(uint32_t M, uint32_t m) { return (kLadPos * m + M - 1) / M; }

struct Fr;
struct Fp;

// Phase-A window plan: W signed-digit windows over a scalar's magnitude
// (below 2^254, with the sign in bit 255: 255 bits always suffice, and the top
// window never carries out), the top k of them c + 1 bits wide and the
// others c bits wide.  Window w starts at bit bit(w); its multiples 1..2^(width-1)
// sit at entries [entry(w), entry(w) + 2^(width-1)) of a base's entries().
struct WindowPlan {
    uint32_t c, W, k;
    __host__ __device__ uint32_t H() const { return 1u << (c - 1); }
    __host__ __device__ uint32_t wide_below(uint32_t w) const { return w > W - k ? w - (W - k) : 0u; }
    __host__ __device__ uint32_t wide(uint32_t w) const { return w >= W - k ? 1u : 0u; }
    __host__ __device__ uint32_t bit(uint32_t w) const { return w * c + wide_below(w); }
    __host__ __device__ uint32_t entry(uint32_t w) const { return (w + wide_below(w)) * H(); }
    __host__ __device__ uint32_t entries() const { return (W + k) * H(); }
};
// The fewest windows of c bits, k of them widened by one bit, that cover 255
// bits: (14, 0) is 19 windows, (14, 3) 18 (15 of 14 bits and 3 of 15).
inline WindowPlan window_plan(uint32_t c, uint32_t k) {
    const uint32_t W = (255 - k + c - 1) / c;
    return WindowPlan{c, W, k < W ? k : W};
}
struct G1A;
struct G1AT;
struct G1J;
struct G1X;

// ------------------------------------------------------------------ setup
hipError_t launch_roots(Fr *roots, Fr *roots_inv, Fr *shift_fwd, Fr *shift_inv, hipStream_t s);
hipError_t launch_decompress(G1A *out, const uint8_t *in, uint32_t n, uint32_t *err, hipStream_t s);
hipError_t launch_xext_fft(G1J *out, const G1A *srs, const Fr *roots, hipStream_t s);
hipError_t launch_jac_to_affine(G1A *out, const G1J *in, Fp *prefix, uint64_t count, uint32_t chunk,
                                hipStream_t s);
hipError_t launch_table_chains(G1J *out, const G1A *xext, WindowPlan plan, hipStream_t s);
// The reference's v1 position table (8192 x 32 entries of 24 words) from pw = 2^(8d) xext.
hipError_t launch_v1_positions(uint32_t *out, const G1A *pw, hipStream_t s);
hipError_t launch_table_multiples(G1AT *table, const G1A *pw, WindowPlan plan,
                                  uint64_t thread_begin, uint32_t thread_count, G1J *scratch,
                                  Fp *prefix, hipStream_t s);
hipError_t launch_phase_b_scalars(Fr *out, const Fr *roots, hipStream_t s);

// ----------------------------------------------------------------- proofs
// Blob b is read at blobs + b * blob_stride.  Also clears queue[0], the tile
// counter of the phase-A launch that follows, and, when coeffs is non-null,
// writes the blob polynomial's monomial coefficients there (Fr[B][8192],
// Montgomery, zero above 4096).
// With `splits` > 1 it writes `splits` scalar sets per blob,
// scal[blob][m][j][i] = scal[blob][0][j][i] * 2^(8 start_m) mod r (each as
// sign and magnitude).
hipError_t launch_fk20_scalars(Fr *scal, const uint8_t *blobs, uint64_t blob_stride, uint32_t *err,
                               const Fr *roots, const Fr *roots_inv, uint32_t *queue, Fr *coeffs,
                               uint32_t splits, uint32_t nblobs, hipStream_t s);
// Persistent: `grid` blocks drain (msm, window group) tiles from queue[0].
// force_checked != 0 (a test switch) sends every tile down the checked
// recompute, as if the fast additions had met a doubling or cancellation; the
// same holds for launch_phase_b and launch_bucket_reduce.
hipError_t launch_phase_a(G1X *part, const Fr *scal, const G1AT *table, WindowPlan plan,
                          uint32_t nmsm, uint32_t groups, uint32_t grid, uint32_t *queue,
                          uint32_t force_checked, hipStream_t s);
// usum[msm] = sum over g < groups of part[msm * groups + g] (groups <= 24,
// nmsm a multiple of 32).
// quad: four lanes share each addition (force_checked: the checked recompute).
hipError_t launch_sum_groups(G1X *usum, const G1X *part, uint32_t groups, uint32_t nmsm, bool quad,
                             uint32_t force_checked, hipStream_t s);
// u: one sum per MSM, [nblobs][splits][128].
hipError_t launch_ladder(G1J *lad, const G1X *u, uint32_t splits, uint32_t nblobs, uint32_t lanes,
                         hipStream_t s);
// Also clears the work counters of the phase-B and bucket-reduction launches
// that follow (queue_b[0], queue_r[0]).
// The ladder points in XYZZ form and their endomorphism images (small
// batches, whose phase B adds XYZZ points on quads); clears the same counters.
hipError_t launch_ladder_xyzz(G1X *pts, G1X *phi, const G1J *lad, uint64_t count, uint32_t *queue_b,
                              uint32_t *queue_r, hipStream_t s);
hipError_t launch_glv_phi(G1A *phi, const G1A *pts, uint64_t count, uint32_t *queue_b, uint32_t *queue_r,
                          hipStream_t s);
// Persistent (`grid` blocks drain queue[0]), then the recompute of flagged
// buckets; flags[B][128][128] is scratch.  quad: four lanes share a bucket;
// xyzz (with quad): lad_pts / lad_phi are G1X (launch_ladder_xyzz), else G1A.
hipError_t launch_phase_b(G1X *buckets, uint8_t *flags, const void *lad_pts, const void *lad_phi,
                          const uint32_t *items, const uint32_t *offsets, uint32_t nblobs, uint32_t grid,
                          bool quad, bool xyzz, uint32_t *queue, uint32_t force_checked, hipStream_t s);
// `lanes` (4, 8, 16 or 32) lanes per output; persistent (`grid` blocks drain
// queue[0]), then the recompute of flagged outputs; flags[B][128] is scratch.
hipError_t launch_bucket_reduce(G1X *cvals, uint8_t *flags, const G1X *buckets, uint32_t nblobs,
                                uint32_t lanes, uint32_t grid, uint32_t *queue, uint32_t force_checked,
                                hipStream_t s);
// quad: four lanes share each output pair's additions (force_checked: their
// checked recompute).
hipError_t launch_finalize(uint8_t *proofs, const G1X *cvals, uint32_t nblobs, bool quad,
                           uint32_t force_checked, hipStream_t s);

// --------------------------------------------------------------- recovery
// The four transforms of a recovery as five kernels (recover.hip), each
// over one-wave work items: 16 per blob with 8 values per lane (radix 8) or
// 32 per blob with 4 (radix 4); `waves` (1, 2 or 4) items per block.  A and
// B are per-blob Fr[8192] intermediates in the layout of the chosen radix.
// Radix 8 computes in the 8 x 32-bit Fr form and reads the Fr tables;
// radix 4 computes in FrL (9 x 29-bit limbs, Montgomery radix 2^261) and
// reads FrL tables of 48-byte entries (launch_frl_table).
struct RecTables {
    const Fr *roots, *roots_inv, *shift_fwd, *shift_inv;          // w^k, w^-k, 7^k / 8192, 7^-k / 8192
    const uint4 *roots4, *roots_inv4, *shift_fwd4, *shift_inv4;  // the same, FrL
};
// out[k] = in[k] as a 48-byte FrL entry (in: the 8 x 32-bit Montgomery form).
hipError_t launch_frl_table(uint4 *out, const Fr *in, uint32_t n, hipStream_t s);
// in_mode 0: `in` is the cells, the first inverse pass with the vanishing
// factor; in_mode 1: `in` is Fr coefficients in natural order (Montgomery),
// the first forward pass of their evaluation (followed by launch_rec_last).
hipError_t launch_rec_first(Fr *out, const void *in, int in_mode, const uint8_t *present, const RecTables &t,
                            uint32_t nblobs, uint32_t radix, uint32_t waves, hipStream_t s);
// stage 0: * 7^k / 8192; stage 1: * 1/Z on the coset (from present);
// stage 2: * 7^-k / 8192.
hipError_t launch_rec_mid(Fr *out, const Fr *in, int stage, const uint8_t *present, const RecTables &t,
                          uint32_t nblobs, uint32_t radix, uint32_t waves, hipStream_t s);
// Canonical big-endian cell bytes.
// The three-kernel recovery (cells -> M, M in place, M -> cells; step 0, 1,
// 2), radix 4 and `waves` per block: M and izc are per-blob Fr[8192] buffers
// (izc: its first 128 values).
hipError_t launch_rec3(uint8_t *cells_out, Fr *M, Fr *izc, const uint8_t *cells, const uint8_t *present,
                       const RecTables &t, uint32_t nblobs, uint32_t waves, hipStream_t s, int step);
hipError_t launch_rec_last(uint8_t *cells, const Fr *in, const RecTables &t, uint32_t nblobs, uint32_t radix,
                           uint32_t waves, hipStream_t s);

// ------------------------------------------------------------ test hooks
hipError_t launch_debug_fp_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, uint32_t n,
                               hipStream_t s);
hipError_t launch_debug_fp_inv(uint32_t *out, const uint32_t *a, uint32_t n, hipStream_t s);
// `blocks` one-wave blocks, each lane n dependent Fp products (rate measurement).
hipError_t launch_debug_fp_mul_chain(Fp *out, uint32_t blocks, uint32_t n, hipStream_t s);
// `blocks` one-wave blocks, each lane n x 128 v_mad_u64_u32 in 8 independent chains.
hipError_t launch_debug_mad_chain(uint64_t *out, uint32_t blocks, uint32_t n, hipStream_t s);
hipError_t launch_debug_fr_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, uint32_t n,
                               hipStream_t s);

}  // namespace vkzg_amd
