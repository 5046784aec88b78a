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
 * validation; `d_cell_present` is the same mask in device memory.  The three
 * device pointers must be 16-byte aligned (hipMalloc's are); others return
 * VKZG_ERR_BADARGS. */
vkzg_result vkzg_amd_recover_cells_device(vkzg_prover *p, void *d_cells_out, const void *d_cells,
                                          const void *d_cell_present,
                                          const uint8_t *h_cell_present, size_t num_blobs,
                                          void *hip_stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory during a call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): vkzg_amd_recover_cells_device with the
 * same arguments and results, whose scratch (2 x 8192 Fr per blob) is a
 * fixed 8 MiB file-scope device array instead of the prover's
 * working set.  The call allocates nothing (no hipMalloc, hipFree or graph).
 * At most 16 blobs per call: more returns VKZG_ERR_BADARGS.  Every GPU holds
 * one copy of the array until the process ends, shared by all calls on that
 * GPU: calls that overlap on one GPU race and must be serialised by the
 * caller. */
vkzg_result vkzg_amd_recover_cells_device_noalloc(vkzg_prover *p, void *d_cells_out, const void *d_cells,
                                                  const void *d_cell_present,
                                                  const uint8_t *h_cell_present, size_t num_blobs,
                                                  void *hip_stream);

/* ---------------------------------------------------- cells and proofs
 * The consensus-spec pair of functions around the reference's two entry
 * points (the reference leaves both to the caller, include/vkzg.h:14-18):
 *
 * compute_cells_and_kzg_proofs for a batch: the 128 cells (2x Reed-Solomon
 * extension, n x 262,144 bytes, canonical big-endian, cell order) and the 128
 * proofs (n x 6,144 bytes, the bytes vkzg_compute_proofs_batch returns) of
 * each blob.  Either output may be NULL, not both.  Errors as
 * vkzg_compute_proofs_batch (INVALID_BLOB for a chunk with a non-canonical
 * element, nothing written for that chunk). */
vkzg_result vkzg_amd_compute_cells_and_proofs(vkzg_prover *p, uint8_t *cells, uint8_t *proofs,
                                              const uint8_t *blobs, size_t num_blobs);

/* recover_cells_and_kzg_proofs for a batch: vkzg_recover_cells_batch's cells
 * plus the 128 proofs of each recovered blob, computed on the device from the
 * recovered cells (no host round trip).  Same buffer contract and errors as
 * vkzg_recover_cells_batch; proofs_out is n x 6,144 bytes. */
vkzg_result vkzg_amd_recover_cells_and_proofs(vkzg_prover *p, uint8_t *cells_out, uint8_t *proofs_out,
                                              const uint8_t *cells, const uint8_t *cell_present,
                                              size_t num_blobs);

/* -------------------------------------------- verification and commitments
 * The consensus-spec functions on either side of the reference's two entry
 * points (the reference has no verifier and no commitment function,
 * include/vkzg.h:14-18), on the device. */

/* Loads the G2 side of the setup for verification: the 65 G2 monomial points
 * [tau^i]G2, i = 0..64, 96-byte compressed each (the trusted setup's G2
 * list), matching the prover's G1 setup.  The Miller-loop lines of
 * [tau^64]G2 and [tau^0]G2 are precomputed on the device and kept resident.
 * BADARGS for a wrong length; SETUP for a malformed point or a point at
 * infinity at index 0 or 64. */
#define VKZG_AMD_G2_SETUP_BYTES (65 * 96)
vkzg_result vkzg_amd_load_g2_setup(vkzg_prover *p, const uint8_t *g2_monomial, size_t len);

/* verify_cell_kzg_proof_batch (consensus specs, polynomial-commitments-
 * sampling): *ok = 1 when every cell k (2048 bytes) with index
 * cell_indices[k] is consistent with commitments[k] (48 bytes) and
 * proofs[k] (48 bytes), checked all at once with Fiat-Shamir weights (the
 * specs' challenge: SHA-256 over every input) and one product of two
 * pairings; *ok = 0 when not.  num_cells == 0 gives *ok = 1.
 * BADARGS: NULL arguments, a cell index >= 128, a malformed or non-G1
 * commitment or proof, or a cell element >= r (as the specs' assertions);
 * SETUP: no G2 setup loaded. */
vkzg_result vkzg_amd_verify_cell_kzg_proof_batch(vkzg_prover *p, int *ok, const uint8_t *commitments,
                                                 const uint64_t *cell_indices, const uint8_t *cells,
                                                 const uint8_t *proofs, size_t num_cells);

/* blob_to_kzg_commitment for a batch: commitments (n x 48 bytes) of blobs
 * (n x 131,072 bytes) against the prover's G1 setup.  INVALID_BLOB for a
 * chunk with a non-canonical element (nothing written for it). */
vkzg_result vkzg_amd_blob_to_commitments(vkzg_prover *p, uint8_t *commitments, const uint8_t *blobs,
                                         size_t num_blobs);

/* Writes the reference's tables file, version 1 (the file its loader reads
 * from $VKZG_TABLES_PATH, src/precomputed_tables.cpp:16-58), for this
 * prover's setup: the 2^(8d) position table of x_ext_fft, the root tables,
 * the phase-B item lists of the reference's 8-bit digit recoding and
 * Montgomery 1/4096.  This build never reads such a file; it derives its own
 * tables on the device.  IO when the file cannot be written. */
vkzg_result vkzg_amd_write_v1_tables(vkzg_prover *p, const char *path);

typedef struct {
    uint32_t window_bits;      /* phase-A table window c */
    uint32_t windows;          /* W = ceil(256 / c) */
    uint64_t table_bytes;      /* resident phase-A table */
    uint64_t working_bytes;    /* per-prover working set */
    uint32_t max_batch_size;   /* chunk size actually used */
    double setup_ms;           /* wall time of prover creation */
    uint32_t last_splits;      /* sub-streams the last call's batches were spread over */
    uint32_t graphs_enabled;   /* 1 unless VKZG_AMD_GRAPHS=0 */
    uint64_t graph_captures;   /* pipelines captured into HIP graphs by the synchronous calls */
    uint64_t graph_replays;    /* calls served by replaying a captured graph */
    uint32_t autotuned;        /* 1 after vkzg_amd_autotune */
    double autotune_ms;        /* its wall time */
    uint32_t phase_b_items;    /* bucket items (point additions) one blob's 128 phase-B outputs add up */
    uint64_t resident_table_bytes; /* the main phase-A table plus any smaller ones the autotuner kept */
    uint32_t extra_table_windows;  /* bit c set: a smaller phase-A table of window c is resident */
} vkzg_amd_info;

/* Fills `out` (never fails for a non-NULL prover). */
vkzg_result vkzg_amd_prover_info(const vkzg_prover *p, vkzg_amd_info *out);

/* ------------------------------------------------------------- tuning
 * Launch shape of the proof and recovery pipelines per batch-size class
 * (<= 1, <= 4, <= 16, <= 64 and more blobs).  Every tuning gives
 * bit-identical results.
 * No reference counterpart: the reference fixes one launch shape. */
typedef struct {
    uint32_t phase_a_groups;        /* window groups (work-queue tiles) per phase-A MSM, 1-20 */
    uint32_t phase_a_blocks_per_cu; /* persistent phase-A grid in blocks per CU, 1-2 */
    uint32_t normalize_chunk;       /* points per batched inversion, 1-256; 0 = the smallest chunk
                                       whose lanes fit in one wave per SIMD (the default) */
    uint32_t splits;                /* sub-streams a batch is spread over, 1-4 */
    uint32_t recovery_waves;        /* recovery work items (one wave each) per block: 1, 2 or 4 */
    uint32_t ladder_splits;         /* doubling-ladder segments, 1, 2 or 4; phase A computes the start
                                       of each (more phase-A work, shorter dependent chains) */
    uint32_t reduce_lanes;          /* bucket-reduction lanes per output: 4, 8, 16 or 32 (one lane per
                                       bucket range), or 64 (16 ranges, four lanes sharing each addition) */
    uint32_t recovery_radix;        /* recovery: 2 (three kernels, 4 values per lane), 4 or 8 (five kernels, 4 or 8 values per lane) */
    uint32_t phase_a_window_bits;   /* phase-A table window: 0 = the main table, else a smaller
                                       resident one (vkzg_amd_info.extra_table_windows) */
} vkzg_amd_tuning;

/* Builds (if not yet resident) a phase-A table of window `window_bits`
 * (10-13, below the main table's) beside the main one, so that tunings may
 * name it (phase_a_window_bits); it stays until the prover is freed.  The
 * autotuner builds such candidates itself and keeps those a size class picks.
 * BADARGS out of range, MALLOC when it does not fit in free device memory. */
vkzg_result vkzg_amd_load_window_table(vkzg_prover *p, uint32_t window_bits);

/* The tuning used for a batch of `num_blobs` blobs. */
vkzg_result vkzg_amd_get_tuning(const vkzg_prover *p, size_t num_blobs, vkzg_amd_tuning *out);
/* Sets the tuning of the size class of `num_blobs`; BADARGS when a field is out of range. */
vkzg_result vkzg_amd_set_tuning(vkzg_prover *p, size_t num_blobs, const vkzg_amd_tuning *t);
/* Replays of captured HIP graphs on (enable != 0, the default unless
 * VKZG_AMD_GRAPHS=0) or off (every call launches its kernels directly). */
vkzg_result vkzg_amd_set_graphs(vkzg_prover *p, int enable);

/* Times candidate tunings on synthetic blobs for every size class up to the
 * prover's batch size and keeps the fastest; *ms_out (optional) receives the
 * time it took.  $VKZG_AMD_AUTOTUNE=1 runs it at prover creation.  Do not run
 * it while a server (below) is using the prover. */
vkzg_result vkzg_amd_autotune(vkzg_prover *p, double *ms_out);

/* ------------------------------------------------------------------ server
 * Continuous batching of proof and recovery requests on one prover.  A
 * dispatcher packs queued requests (oldest first, one kind per batch) into
 * device batches of up to `batch_blobs` blobs and runs up to `slots` batches
 * at once, each with its own device working set; requests arriving while
 * batches run join the next free slot.  `hbm_budget_bytes` caps the device
 * memory of the prover's resident phase-A tables plus the slots' working
 * sets (fewer slots, then smaller batches); a budget the tables alone fill
 * is VKZG_ERR_MALLOC.  Request buffers belong to the
 * caller and must stay valid until vkzg_amd_server_wait returns.  A request
 * with an invalid blob fails alone (VKZG_ERR_INVALID_BLOB, its outputs
 * unspecified); more than 64 missing cells is rejected at submission.
 * Thread-safe. */
typedef struct vkzg_amd_server vkzg_amd_server;

typedef struct {
    uint32_t batch_blobs;      /* blobs per device batch; 0 = the prover's max_batch_size */
    uint32_t slots;            /* batches in flight; 0 = 2 */
    uint64_t hbm_budget_bytes; /* device memory for the resident tables and the slots; 0 = no limit */
} vkzg_amd_server_options;

typedef struct {
    uint64_t submit_ns, start_ns, done_ns; /* steady-clock times */
    uint32_t batches;                      /* device batches the request was spread over */
} vkzg_amd_request_info;

typedef struct {
    uint64_t requests, blobs, batches;
    uint32_t slots, batch_blobs;
    uint64_t slot_bytes;
} vkzg_amd_server_counters;

void vkzg_amd_server_options_default(vkzg_amd_server_options *opts);
vkzg_result vkzg_amd_server_new(vkzg_amd_server **out, vkzg_prover *p,
                                const vkzg_amd_server_options *opts);
/* Drains every queued request, then frees the server. */
void vkzg_amd_server_free(vkzg_amd_server *s);
vkzg_result vkzg_amd_server_submit_proofs(vkzg_amd_server *s, uint8_t *proofs, const uint8_t *blobs,
                                          size_t num_blobs, uint64_t *ticket);
vkzg_result vkzg_amd_server_submit_recovery(vkzg_amd_server *s, uint8_t *cells_out,
                                            const uint8_t *cells, const uint8_t *cell_present,
                                            size_t num_blobs, uint64_t *ticket);
/* *done = 1 once the request has completed. */
vkzg_result vkzg_amd_server_poll(vkzg_amd_server *s, uint64_t ticket, int *done);
/* Blocks until the request completes, returns its result and retires the
 * ticket; `info` (may be NULL) receives its timestamps. */
vkzg_result vkzg_amd_server_wait(vkzg_amd_server *s, uint64_t ticket, vkzg_amd_request_info *info);
vkzg_result vkzg_amd_server_stats(vkzg_amd_server *s, vkzg_amd_server_counters *out);

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
/* n inverses mod p (0 -> 0) of canonical Fp vectors by the device's
 * binary-GCD inversion, cross-checked in the kernel against the Fermat power
 * (all-ones output on disagreement). */
vkzg_result vkzg_amd_debug_fp_inv(vkzg_prover *p, uint32_t *out, const uint32_t *a, size_t n);
vkzg_result vkzg_amd_debug_fr_mul(vkzg_prover *p, uint32_t *out, const uint32_t *a,
                                  const uint32_t *b, size_t n);
/* Issue-rate measurement of the device Fp multiplication: every lane of
 * CUs x 4 x waves_per_simd one-wave blocks (1-8 waves per SIMD) runs `iters`
 * dependent products; *muls_per_s is the chip's rate, *one_wave_cycles the
 * cycles per product of a single wave alone (its dependency latency). */
vkzg_result vkzg_amd_debug_fp_mul_rate(vkzg_prover *p, uint32_t waves_per_simd, uint32_t iters,
                                       double *muls_per_s, double *one_wave_cycles);
/* Issue-rate measurement of v_mad_u64_u32 (the 32x32->64 multiply-add every
 * limb product is): every lane of CUs x 4 x waves_per_simd one-wave blocks
 * (1-8 waves per SIMD) runs iters x 128 of them in 8 independent chains;
 * *mads_per_s is the chip's lane rate, *clock_mhz the device's nominal clock. */
vkzg_result vkzg_amd_debug_mad_rate(vkzg_prover *p, uint32_t waves_per_simd, uint32_t iters,
                                    double *mads_per_s, double *clock_mhz);
/* Test switch: with enable != 0, every later proof pipeline on p sends all of
 * phase A, phase B and the bucket reduction through their checked recompute
 * (the path the fast point additions take when they meet a doubling or a
 * cancellation, which random inputs never reach).  Proofs are unchanged; only
 * slower. */
vkzg_result vkzg_amd_debug_force_checked(vkzg_prover *p, int enable);
/* n device pairings e(P_i, Q_i)^3 (P_i 48-byte compressed G1, Q_i 96-byte
 * compressed G2), each as 12 Fp values of 48 bytes, big-endian, in the order
 * (w^0 real, w^0 imaginary, w^1 real, ...) of Fp12 = Fp2[w]/(w^6 - (1 + u)). */
vkzg_result vkzg_amd_debug_pairing(vkzg_prover *p, uint8_t *out, const uint8_t *g1, const uint8_t *g2,
                                   size_t n);
/* sum_i [s_i] P_i by the device MSM (points 48-byte compressed, scalars
 * 32-byte big-endian below r), compressed. */
vkzg_result vkzg_amd_debug_g1_msm(vkzg_prover *p, uint8_t out[48], const uint8_t *points,
                                  const uint8_t *scalars, size_t n);
/* vkzg_amd_verify_cell_kzg_proof_batch that also returns the two G1 inputs
 * of its pairing check, compressed: terms[0..48) = sum r^k proof_k,
 * terms[48..96) = the right-hand G1 point RL (num_cells >= 1). */
vkzg_result vkzg_amd_debug_verify_terms(vkzg_prover *p, int *ok, uint8_t terms[96], const uint8_t *commitments,
                                        const uint64_t *cell_indices, const uint8_t *cells,
                                        const uint8_t *proofs, size_t num_cells);
/* Host-side helpers of verification (no device needed): the Fiat-Shamir
 * challenge r (32 bytes big-endian) of a verification call, and SHA-256 with
 * (portable = 0, when the CPU has the SHA extensions) or without them. */
vkzg_result vkzg_amd_verify_challenge(uint8_t r_out[32], const uint8_t *commitments,
                                      const uint64_t *cell_indices, const uint8_t *cells,
                                      const uint8_t *proofs, size_t num_cells, int portable_sha);
vkzg_result vkzg_amd_sha256(uint8_t out[32], const uint8_t *data, size_t len, int portable);
int vkzg_amd_sha256_hw(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VKZG_AMD_H */
