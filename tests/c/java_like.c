/*
 * Calls libvkzg.so the way the reference's Java FFM binding does
 * (bindings/java/src/main/java/vkzg/Vkzg.java:53-184): dlopen, dlsym of the
 * eight C symbols, a 4-byte options struct, VKZG_TABLES_PATH set before
 * vkzg_prover_new, then computeCellKzgProofsBatch and recoverCellsBatch.
 * Usage:
 *   java_like LIB --symbols-only
 *       symbol lookup + NULL-argument contract (no GPU needed)
 *   java_like LIB TABLES BLOBS PROOFS N [CELLS PRESENT WANT M]
 *       VKZG_TABLES_PATH=TABLES, prover_new, one N-blob proof call memcmp'd
 *       against PROOFS, and with CELLS given one M-blob recover call
 *       memcmp'd against WANT; prints "proofs ok" / "recover ok".
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef const char *(*error_string_fn)(int);
typedef void (*options_default_fn)(void *);
typedef int (*prover_new_fn)(void **, const void *);
typedef void (*prover_free_fn)(void *);
typedef const char *(*device_name_fn)(const void *);
typedef uint32_t (*core_count_fn)(const void *);
typedef int (*compute_fn)(void *, uint8_t *, const uint8_t *, size_t);
typedef int (*recover_fn)(void *, uint8_t *, const uint8_t *, const uint8_t *, size_t);

static void *need(void *lib, const char *name) {
    void *s = dlsym(lib, name);
    if (!s) {
        printf("FAIL: symbol not found: %s\n", name);
        exit(1);
    }
    return s;
}

static uint8_t *slurp(const char *path, size_t *len) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    *len = (size_t)ftell(f);
    fseek(f, 0, SEEK_SET);
    uint8_t *b = malloc(*len ? *len : 1);
    if (fread(b, 1, *len, f) != *len) {
        fclose(f);
        free(b);
        return NULL;
    }
    fclose(f);
    return b;
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    void *lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!lib) {
        printf("FAIL: dlopen: %s\n", dlerror());
        return 1;
    }
    error_string_fn error_string = (error_string_fn)need(lib, "vkzg_error_string");
    options_default_fn options_default = (options_default_fn)need(lib, "vkzg_options_default");
    prover_new_fn prover_new = (prover_new_fn)need(lib, "vkzg_prover_new");
    prover_free_fn prover_free = (prover_free_fn)need(lib, "vkzg_prover_free");
    device_name_fn device_name = (device_name_fn)need(lib, "vkzg_prover_device_name");
    core_count_fn core_count = (core_count_fn)need(lib, "vkzg_prover_gpu_core_count");
    compute_fn compute = (compute_fn)need(lib, "vkzg_compute_proofs_batch");
    recover_fn recover = (recover_fn)need(lib, "vkzg_recover_cells_batch");
    printf("8 symbols ok\n");

    uint8_t dummy[64] = {0};
    if (compute(NULL, dummy, dummy, 1) != 1 || recover(NULL, dummy, dummy, dummy, 1) != 1 ||
        device_name(NULL)[0] != 0 || core_count(NULL) != 0 || strcmp(error_string(6), "") == 0) {
        printf("FAIL: NULL-argument contract\n");
        return 1;
    }
    if (strcmp(argv[2], "--symbols-only") == 0) return 0;
    if (argc < 6) return 2;

    size_t blob_len = 0, proofs_len = 0;
    uint8_t *blob = slurp(argv[3], &blob_len);
    uint8_t *want = slurp(argv[4], &proofs_len);
    const size_t n = (size_t)atoi(argv[5]);
    if (!blob || !want || blob_len != n * 131072 || proofs_len != n * 6144) {
        printf("FAIL: bad input files\n");
        return 1;
    }
    /* Vkzg.extractPrecomputedTables: the path is set before the prover exists. */
    setenv("VKZG_TABLES_PATH", argv[2], 1);
    uint8_t *opts = calloc(1, 4); /* Java allocates exactly 4 bytes */
    options_default(opts);
    void *prover = NULL;
    int rc = prover_new(&prover, opts);
    if (rc != 0) {
        printf("FAIL: prover_new: %s\n", error_string(rc));
        return 1;
    }
    printf("device %s, %u compute units\n", device_name(prover), core_count(prover));
    uint8_t *proofs = malloc(n * 6144);
    rc = compute(prover, proofs, blob, n);
    if (rc != 0 || memcmp(proofs, want, n * 6144) != 0) {
        printf("FAIL: compute rc=%d (%s) match=%d\n", rc, error_string(rc),
               rc == 0 && memcmp(proofs, want, n * 6144) == 0);
        return 1;
    }
    printf("proofs ok\n");

    if (argc >= 10) {
        size_t cells_len = 0, present_len = 0, want_len = 0;
        uint8_t *cells = slurp(argv[6], &cells_len);
        uint8_t *present = slurp(argv[7], &present_len);
        uint8_t *want_cells = slurp(argv[8], &want_len);
        const size_t m = (size_t)atoi(argv[9]);
        if (!cells || !present || !want_cells || cells_len != m * 262144 || present_len != m * 128 ||
            want_len != m * 262144) {
            printf("FAIL: bad recovery input files\n");
            return 1;
        }
        uint8_t *out = malloc(m * 262144);
        rc = recover(prover, out, cells, present, m);
        if (rc != 0 || memcmp(out, want_cells, m * 262144) != 0) {
            printf("FAIL: recover rc=%d (%s) match=%d\n", rc, error_string(rc),
                   rc == 0 && memcmp(out, want_cells, m * 262144) == 0);
            return 1;
        }
        printf("recover ok\n");
        free(out);
        free(cells);
        free(present);
        free(want_cells);
    }
    prover_free(prover);
    free(opts);
    free(proofs);
    free(blob);
    free(want);
    dlclose(lib);
    return 0;
}
