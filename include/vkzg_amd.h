/*
 * Extensions of the vkzg C ABI for the MI355X build.  None of these has a
 * counterpart in the reference; they are additive (vkzg.h and vkzg_options
 * are unchanged) and exist for callers that:
 *   - need a setup other than the mainnet one (seeded test setups),
 *   - keep blobs and cells in device memory (benchmarks, GPU pipelines),
 *   - inspect the resident tables or test the device field arithmetic.
 */
#ifndef VKZG_AMD_H
#define VKZG_AMD_H

#include "vkzg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Like vkzg_prover_new, for a caller-supplied G1 monomial setup: `len` must be
 * 4096 * 48 bytes of compressed points ([tau^i]G1, i < 4096).  A malformed
 * point -> VKZG_ERR_SETUP. */
vkzg_result vkzg_amd_prover_new_with_setup(vkzg_prover **out, const vkzg_options *opts,
                                           const uint8_t *g1_monomial, size_t len);

/* vkzg_compute_proofs_batch on device pointers (hipMalloc'd on the prover's
 * device).  Work is enqueued on `hip_stream` (a hipStream_t; NULL = the
 * prover's own stream) and the call returns once it has completed.  Blob
 * validation reports VKZG_ERR_INVALID_BLOB per chunk as in the host entry. */
vkzg_result vkzg_amd_compute_proofs_device(vkzg_prover *p, void *d_proofs, const void *d_blobs,
                                           size_t num_blobs, void *hip_stream);

/* vkzg_recover_cells_batch on device pointers.  `h_cell_present` is the host
 * copy of the presence mask (num_blobs x 128 bytes) used for argument
 * validation; `d_cell_present` is the same mask in device memory. */
vkzg_result vkzg_amd_recover_cells_device(vkzg_prover *p, void *d_cells_out, const void *d_cells,
                                          const void *d_cell_present,
                                          const uint8_t *h_cell_present, size_t num_blobs,
                                          void *hip_stream);

typedef struct {
    uint32_t window_bits;      /* phase-A table window c */
    uint32_t windows;          /* W = ceil(256 / c) */
    uint64_t table_bytes;      /* resident phase-A table */
    uint64_t working_bytes;    /* per-prover working set */
    uint32_t max_batch_size;   /* chunk size actually used */
    double setup_ms;           /* wall time of prover creation */
    uint32_t last_splits;      /* sub-streams the last call's batches were spread over */
} vkzg_amd_info;

/* Fills `out` (never fails for a non-NULL prover). */
vkzg_result vkzg_amd_prover_info(const vkzg_prover *p, vkzg_amd_info *out);

/* Per-stage timing: when enabled, the prover records HIP events between its
 * kernel launches on the stream they run on; vkzg_amd_last_stage_times then
 * returns the elapsed milliseconds of each stage of the last call (summed over
 * its chunks) and the stage names (static strings).  Returns the stage count. */
vkzg_result vkzg_amd_set_stage_timing(vkzg_prover *p, int enable);
int vkzg_amd_last_stage_times(vkzg_prover *p, float *ms, const char **names, int max);

/* Test hooks (p may be NULL: current device): n products of canonical little-endian limb vectors (12 x u32
 * for Fp, 8 x u32 for Fr; Fr inputs may be any 256-bit value and are reduced
 * mod r first), computed by the device Montgomery code. */
vkzg_result vkzg_amd_debug_fp_mul(vkzg_prover *p, uint32_t *out, const uint32_t *a,
                                  const uint32_t *b, size_t n);
vkzg_result vkzg_amd_debug_fr_mul(vkzg_prover *p, uint32_t *out, const uint32_t *a,
                                  const uint32_t *b, size_t n);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VKZG_AMD_H */
