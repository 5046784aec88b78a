/*
 * Calls libvkzg.so the way the reference's Java FFM binding does
 * (bindings/java/src/main/java/vkzg/Vkzg.java:53-184): dlopen, dlsym of the
 * eight C symbols, a 4-byte options struct, VKZG_TABLES_PATH set before
 * vkzg_prover_new.  Usage:
 *   java_like LIB --symbols-only          symbol lookup + NULL-argument contract
 *   java_like LIB BLOB PROOFS [N]         prover_new, one N-blob call, memcmp
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
    uint8_t *b = malloc(*len);
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

    size_t blob_len = 0, proofs_len = 0;
    uint8_t *blob = slurp(argv[2], &blob_len);
    uint8_t *want = slurp(argv[3], &proofs_len);
    const size_t n = argc > 4 ? (size_t)atoi(argv[4]) : 1;
    if (!blob || !want || blob_len != n * 131072 || proofs_len != n * 6144) {
        printf("FAIL: bad input files\n");
        return 1;
    }
    setenv("VKZG_TABLES_PATH", "/nonexistent/precomputed_tables.bin", 1);
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
    prover_free(prover);
    free(opts);
    free(proofs);
    free(blob);
    free(want);
    dlclose(lib);
    return 0;
}
