/*
 * vkzg -- EIP-7594 (PeerDAS) cell KZG proofs and cell recovery, on one AMD
 * Instinct MI355X (gfx950) through HIP.
 *
 * This is the drop-in C ABI: the same eight symbols, enum values, size macros
 * and 4-byte options struct as the reference library's public header, so
 * existing callers (its Java FFM binding, C/C++ programs using dlsym) link
 * against this build unchanged.  Each declaration cites the reference
 * interface it replaces (paths relative to the reference checkout).
 *
 * Every function may be called from several threads on one prover.  Calls on
 * one prover are serialised by the prover; results are independent of what
 * else runs on it.
 *
 * Extensions that have no counterpart in the reference (custom setups,
 * device-pointer entry points, the request server) live in vkzg_amd.h.
 */
#ifndef VKZG_H
#define VKZG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Protocol sizes (reference include/vkzg.h:32-38). */
#define VKZG_FIELD_ELEMENTS_PER_BLOB 4096
#define VKZG_NUM_CELL_PROOFS 128
#define VKZG_BYTES_PER_FIELD_ELEMENT 32
#define VKZG_BYTES_PER_BLOB (VKZG_FIELD_ELEMENTS_PER_BLOB * VKZG_BYTES_PER_FIELD_ELEMENT)
#define VKZG_BYTES_PER_PROOF 48
#define VKZG_FIELD_ELEMENTS_PER_CELL 64
#define VKZG_BYTES_PER_CELL (VKZG_FIELD_ELEMENTS_PER_CELL * VKZG_BYTES_PER_FIELD_ELEMENT)

/* Status codes; the numeric values are ABI (reference include/vkzg.h:42-50). */
typedef enum {
    VKZG_OK = 0,
    VKZG_ERR_BADARGS = 1,     /* null pointer or an argument out of range */
    VKZG_ERR_MALLOC = 2,      /* host or device allocation failed */
    VKZG_ERR_IO = 3,          /* a setup file could not be read */
    VKZG_ERR_SETUP = 4,       /* a setup file or setup point was malformed */
    VKZG_ERR_GPU = 5,         /* no usable GPU, or a HIP call failed */
    VKZG_ERR_INVALID_BLOB = 6 /* a blob field element was not below r */
} vkzg_result;

/* Static description of a status code; never NULL (reference src/vkzg.cpp:170-181). */
const char *vkzg_error_string(vkzg_result r);

/* Exactly 4 bytes: the Java binding allocates 4 (reference Vkzg.java:137-138). */
typedef struct {
    /* Largest number of blobs one device launch covers; larger calls are
     * processed in chunks of this size.  0 selects the default (256 here). */
    uint32_t max_batch_size;
} vkzg_options;

/* Writes the defaults; NULL is ignored (reference src/vkzg.cpp:183-186). */
void vkzg_options_default(vkzg_options *opts);

typedef struct vkzg_prover vkzg_prover;

/* Prover for the Ethereum mainnet trusted setup, built on the current HIP
 * device: the embedded G1 monomial setup is decompressed and the FK20 tables
 * are derived on the GPU and kept resident in HBM.  `opts` may be NULL.
 * VKZG_TABLES_PATH is accepted and ignored: nothing is read from disk
 * (reference src/vkzg.cpp:721-730, src/precomputed_tables.cpp:62-73). */
vkzg_result vkzg_prover_new(vkzg_prover **out, const vkzg_options *opts);

/* Frees a prover and its device memory; NULL is ignored (reference src/vkzg.cpp:732). */
void vkzg_prover_free(vkzg_prover *p);

/* HIP device name, owned by the prover; "" for NULL (reference src/vkzg.cpp:188-190). */
const char *vkzg_prover_device_name(const vkzg_prover *p);

/* Compute-unit count of the device; 0 for NULL (reference src/vkzg.cpp:192-194). */
uint32_t vkzg_prover_gpu_core_count(const vkzg_prover *p);

/*
 * All 128 cell proofs of each of `num_blobs` consecutive blobs
 * (reference src/vkzg.cpp:1183-1199).
 *   blobs:  num_blobs x VKZG_BYTES_PER_BLOB bytes, 4096 big-endian field elements each.
 *   proofs: num_blobs x 128 x 48 bytes, compressed G1, cell order.
 * NULL pointers -> VKZG_ERR_BADARGS (checked before num_blobs); num_blobs == 0
 * -> VKZG_OK.  A non-canonical element -> VKZG_ERR_INVALID_BLOB for its chunk
 * of max_batch_size blobs, whose proofs are not written; earlier chunks stay
 * written.
 */
vkzg_result vkzg_compute_proofs_batch(vkzg_prover *p, uint8_t *proofs, const uint8_t *blobs,
                                      size_t num_blobs);

/*
 * All 128 cells of each of `num_blobs` blobs from at least 64 present cells
 * (reference src/vkzg.cpp:1201-1218).
 *   cells:        num_blobs x 128 x VKZG_BYTES_PER_CELL bytes; a missing cell's
 *                 bytes are ignored, present bytes are reduced mod r.
 *   cell_present: num_blobs x 128 bytes, nonzero = present.
 *   cells_out:    num_blobs x 128 x VKZG_BYTES_PER_CELL bytes.
 * NULL pointers -> VKZG_ERR_BADARGS; num_blobs == 0 -> VKZG_OK; a blob with
 * more than 64 missing cells -> VKZG_ERR_BADARGS before its chunk runs
 * (earlier chunks stay written).
 */
vkzg_result vkzg_recover_cells_batch(vkzg_prover *p, uint8_t *cells_out, const uint8_t *cells,
                                     const uint8_t *cell_present, size_t num_blobs);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VKZG_H */
