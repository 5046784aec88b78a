/*
 * CPU oracle for the vkzg hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of what the reference computes for its two batched
 * entry points (FK20 cell proofs and cell recovery), used by tests/, by the
 * smoke check and by bench.py's CPU-baseline leg as the thing the GPU path is
 * checked against.  Nothing in the shipped library links or calls this code.
 *
 * Arithmetic is exact: Fp/Fr in Montgomery form with 64-bit limbs and
 * unsigned __int128 products.  Byte formats are the ones of the reference's
 * public header (include/vkzg.h:32-38): 32-byte big-endian field elements,
 * 48-byte ZCash-compressed G1 points.
 */
#ifndef VKZG_ORACLE_H
#define VKZG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque FK20 setup: decompressed SRS, x_ext_fft and the phase-A position table. */
typedef struct or_setup or_setup;

/* ---- field element byte helpers (exposed for unit tests) ---- */
/* r (the BLS12-381 group order) and p (the base field prime) as 32/48 big-endian bytes. */
void or_modulus_r(uint8_t out[32]);
void or_modulus_p(uint8_t out[48]);
/* out = a*b mod r for 32-byte big-endian canonical inputs (Montgomery round trip). */
void or_fr_mul_be(uint8_t out[32], const uint8_t a[32], const uint8_t b[32]);
/* out = a*b mod p for 48-byte big-endian canonical inputs. */
void or_fp_mul_be(uint8_t out[48], const uint8_t a[48], const uint8_t b[48]);
/* out = a^-1 mod r (0 -> 0). */
void or_fr_inv_be(uint8_t out[32], const uint8_t a[32]);
/* k-th power of the primitive 8192-th root of unity 7^((r-1)/8192), big-endian. */
void or_root_8192_be(uint8_t out[32], uint32_t k);

/* ---- G1 helpers ---- */
/* Decompress then recompress (returns 0 on success, -1 on invalid encoding). */
int or_g1_roundtrip(uint8_t out[48], const uint8_t in[48]);
/* out = compress(k * G1) for a 32-byte big-endian scalar (any 256-bit value). */
void or_g1_mul_gen(uint8_t out[48], const uint8_t k[32]);
/* out = compress(k * P) for a compressed point P. Returns -1 on a bad encoding. */
int or_g1_mul(uint8_t out[48], const uint8_t p[48], const uint8_t k[32]);
/* out = compress(P + Q). */
int or_g1_add(uint8_t out[48], const uint8_t p[48], const uint8_t q[48]);

/* ---- polynomial / cells ---- */
/* Blob (4096 x 32-byte BE) -> monomial coefficients (4096 x 32-byte BE canonical).
 * Returns 0, or -1 when an element is not canonical (>= r). */
int or_blob_to_poly(uint8_t coeffs[4096 * 32], const uint8_t blob[131072]);
/* Blob -> 128 cells (128 x 64 x 32 bytes), the 2x Reed-Solomon extension. -1 if non-canonical. */
int or_blob_to_cells(uint8_t cells[128 * 2048], const uint8_t blob[131072]);
/* Forward (inverse=0) or inverse (inverse=1, scaled by 1/n) NTT over Fr of size n | 8192,
 * natural order in and out, on 32-byte BE canonical elements (reduced mod r on input). */
void or_fr_ntt(uint8_t *data, uint32_t n, int inverse);

/* ---- recovery ---- */
/* Recover all 128 cells of `num_blobs` blobs.  Same buffer contract as
 * vkzg_recover_cells_batch.  Returns 0, or 1 (BADARGS) when a blob has fewer
 * than 64 present cells. */
int or_recover_cells(uint8_t *cells_out, const uint8_t *cells, const uint8_t *present,
                     size_t num_blobs);

/* ---- FK20 ---- */
/* Build the FK20 setup from 4096 compressed G1 monomial SRS points.  NULL on a bad point. */
or_setup *or_setup_new(const uint8_t *g1_monomial_compressed /* 4096*48 */);
void or_setup_free(or_setup *s);
/* 128 compressed cell proofs per blob.  Returns 0, or 6 (INVALID_BLOB) when a
 * blob element is not canonical (then no proofs of that blob are written). */
int or_compute_proofs(const or_setup *s, uint8_t *proofs, const uint8_t *blobs, size_t num_blobs);
/* The x_ext_fft point for (j, i) as compressed bytes (for table-layout tests). */
void or_setup_xext(const or_setup *s, uint32_t j, uint32_t i, uint8_t out[48]);

/* ---- seeded-secret setups and the closed form ---- */
/* 4096 compressed G1 points [tau^i]G1 for a 32-byte BE tau. */
void or_setup_g1_from_tau(uint8_t *out /* 4096*48 */, const uint8_t tau[32]);
/* Proofs by definition: proof[c] = [q_c(tau)]G1 with q_c = p div (X^64 - w^(64*brev7(c))).
 * Returns 0 or 6 like or_compute_proofs. */
int or_proofs_from_tau(uint8_t *proofs, const uint8_t *blob, const uint8_t tau[32]);

/* ---- reference-style direct H_m form (independent FK20 cross-check) ---- */
/* h[m-1] = sum_k c_{k+64m} s_k for m=1..63, then proofs = brp7(FFT_128(h)).  Slow. */
int or_compute_proofs_direct(const or_setup *s, uint8_t *proofs, const uint8_t *blob);

/* ---- pairing and batch cell-proof verification ---- */
/* G2 points use a 192-byte encoding: x.c0, x.c1, y.c0, y.c1 (48-byte BE each),
 * Fp2 = Fp[u]/(u^2+1); all-zero = infinity.  No G2 data ships with the
 * reference, so G2 uses its own point Q of order r (any nonzero point of the
 * cyclic group G2 gives the same verification equations). */
void or_g2_generator(uint8_t out[192]);
/* A point of E'(Fp2) before cofactor clearing (x = seed+1+i + u for the first
 * square right-hand side). */
int or_g2_raw_point(uint8_t out[192], uint64_t seed);
int or_g2_on_curve(const uint8_t p[192]);
/* out = k * P for a big-endian scalar of k_len bytes (any length). */
int or_g2_mul(uint8_t out[192], const uint8_t p[192], const uint8_t *k_be, uint32_t k_len);
int or_g2_add(uint8_t out[192], const uint8_t p[192], const uint8_t q[192]);
/* [tau^i]Q for i < n. */
void or_g2_powers_from_tau(uint8_t *out /* n*192 */, const uint8_t tau[32], uint32_t n);
/* Constants exposed for tests: G2 cofactor (64 B BE) and (p^6+1)/r (256 B BE). */
void or_g2_cofactor(uint8_t out[64]);
void or_final_exp_hard(uint8_t out[256]);
/* Optimal-ate pairing value (after final exponentiation), 12 Fp elements in
 * tower order d0.c0.c0, d0.c0.c1, d0.c1.c0, ..., d1.c2.c1, 48-byte BE each. */
int or_pairing(uint8_t out[576], const uint8_t p[48], const uint8_t q[192]);
/* 1 when prod_k e(P_k, Q_k) == 1, 0 otherwise, -1 on a bad encoding. */
int or_pairing_check(const uint8_t *g1s /* n*48 */, const uint8_t *g2s /* n*192 */, size_t n);

/* Verification setup: 4096 compressed G1 monomial points and >= 65 G2 powers [tau^i]Q. */
typedef struct or_vsetup or_vsetup;
or_vsetup *or_vsetup_new(const uint8_t *g1_monomial /* 4096*48 */, const uint8_t *g2_powers /* 65*192 */);
void or_vsetup_free(or_vsetup *s);
/* KZG commitment of a blob; 0, or 6 for a non-canonical blob. */
int or_commit(const or_vsetup *s, uint8_t out[48], const uint8_t blob[131072]);
/* Commitments of n blobs (n*48 bytes), computed in parallel; 0 or 6. */
int or_commit_batch(const or_vsetup *s, uint8_t *out, const uint8_t *blobs, size_t n);
/* Batch verification of n (commitment, cell index, cell, proof) items with
 * random weights drawn from `seed`: 1 valid, 0 invalid, -1 malformed input. */
int or_verify_cell_proofs(const or_vsetup *s, const uint8_t *commitments, const uint32_t *cell_indices,
                          const uint8_t *cells, const uint8_t *proofs, size_t n, uint64_t seed);

/* verify_cell_kzg_proof_batch with the weights r^k of a given challenge r;
 * out_a / out_b receive the two G1 inputs of the pairing check (compressed). */
int or_verify_cell_proofs_r(const or_vsetup *s, const uint8_t *commitments, const uint32_t *cell_indices,
                            const uint8_t *cells, const uint8_t *proofs, size_t n, const uint8_t r_be[32],
                            uint8_t out_a[48], uint8_t out_b[48]);
/* ZCash-compressed G2 <-> the 192-byte encoding above. */
int or_g2_compress(uint8_t out[96], const uint8_t in[192]);
int or_g2_decompress(uint8_t out[192], const uint8_t in[96]);
/* sum [s_i] P_i, compressed. */
int or_g1_msm(uint8_t out[48], const uint8_t *points, const uint8_t *scalars, size_t n);

#ifdef __cplusplus
}
#endif
#endif
