/*
 * CPU oracle for the vkzg hot path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * What each part restates (reference paths relative to the reference checkout):
 *   Fp/Fr Montgomery arithmetic ........ src/shaders/field.glsl.inc:69-297 (CIOS,
 *                                          here with 64-bit limbs)
 *   G1 dbl/add/madd .................... src/shaders/field.glsl.inc:322-408
 *   blob -> bit-reversed Lagrange -> iNTT src/shaders/k_blob_to_fr.comp:23-37,
 *                                          src/vkzg.cpp:920-935
 *   circulant columns + NTT_128 ........ src/shaders/k_build_circulant.comp:30-72
 *   phase A digit recoding + buckets ... src/shaders/k_phase_a_sort.comp:31-114,
 *                                          src/shaders/k_phase_a.comp:29-43
 *   bucket reduction sum k*B[k-1] ...... src/shaders/field.glsl.inc:478-517
 *   phase B (split circulant) .......... src/shaders/k_phase_b_split.comp,
 *                                          k_combine_split.comp; restated here as the
 *                                          FK20 definition FFT_128(trunc_64(IFFT_128(u)))
 *                                          that the split form is an algebraic rewrite of
 *                                          (src/layout_defs.h:30-45)
 *   proof compression + brp7 ........... src/shaders/k_compress_proofs.comp:21-46
 *   recovery data flow ................. src/vkzg.cpp:1086-1179 with k_cells_to_fr,
 *                                          k_vanishing_poly, k_ew_mul, k_ew_div,
 *                                          k_pow_seq, k_fr_to_cells
 * The x_ext_fft layout (which SRS point feeds which circulant slot) is not in the
 * reference's sources (its table file is missing); it is derived in or_setup_new
 * from the FK20 identity h[t] = sum_k c[k+64(t+1)] s_k and checked by the tests
 * against the consensus vectors and the closed form.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

typedef struct { uint64_t v[6]; } fp;
typedef struct { uint64_t v[4]; } fr;
typedef struct { fp x, y, z; } g1j; /* Jacobian, z == 0 is the identity */
typedef struct { fp x, y; int inf; } g1a;

/* ------------------------------------------------------------- constants */

static const uint64_t P_MOD[6] = {0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL,
                                  0x6730d2a0f6b0f624ULL, 0x64774b84f38512bfULL,
                                  0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
static const uint64_t R_MOD[4] = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL,
                                  0x3339d80809a1d805ULL, 0x73eda753299d7d48ULL};

static uint64_t P_N0, R_N0;
static fp FP_ONE, FP_R2, FP_FOUR;
static fr FR_ONE, FR_R2;
static uint64_t P_MINUS_2[6], P_SQRT_EXP[6], P_HALF[6];
static uint64_t R_MINUS_2[4];
static fr ROOTS[8192]; /* omega_8192^k, Montgomery */
static fr SHIFT_POW[8192], SHIFT_INV_POW[8192]; /* 7^k, 7^-k */
static g1j G1_GEN;

/* ------------------------------------------------------- generic limb math */

static inline void mont_mul_n(uint64_t *res, const uint64_t *a, const uint64_t *b,
                              const uint64_t *m, uint64_t n0, int N) {
    uint64_t t[8] = {0};
    for (int i = 0; i < N; i++) {
        uint64_t C = 0;
        for (int j = 0; j < N; j++) {
            u128 s = (u128)a[j] * b[i] + t[j] + C;
            t[j] = (uint64_t)s;
            C = (uint64_t)(s >> 64);
        }
        u128 s = (u128)t[N] + C;
        t[N] = (uint64_t)s;
        t[N + 1] = (uint64_t)(s >> 64);
        const uint64_t mm = t[0] * n0;
        s = (u128)mm * m[0] + t[0];
        C = (uint64_t)(s >> 64);
        for (int j = 1; j < N; j++) {
            s = (u128)mm * m[j] + t[j] + C;
            t[j - 1] = (uint64_t)s;
            C = (uint64_t)(s >> 64);
        }
        s = (u128)t[N] + C;
        t[N - 1] = (uint64_t)s;
        t[N] = t[N + 1] + (uint64_t)(s >> 64);
    }
    uint64_t d[8];
    uint64_t borrow = 0;
    for (int j = 0; j < N; j++) {
        u128 s = (u128)t[j] - m[j] - borrow;
        d[j] = (uint64_t)s;
        borrow = (uint64_t)(s >> 64) & 1;
    }
    if (t[N] || !borrow) memcpy(res, d, 8 * N);
    else memcpy(res, t, 8 * N);
}

static inline void mod_add_n(uint64_t *r, const uint64_t *a, const uint64_t *b,
                             const uint64_t *m, int N) {
    uint64_t t[8], d[8], c = 0, borrow = 0;
    for (int j = 0; j < N; j++) {
        u128 s = (u128)a[j] + b[j] + c;
        t[j] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
    for (int j = 0; j < N; j++) {
        u128 s = (u128)t[j] - m[j] - borrow;
        d[j] = (uint64_t)s;
        borrow = (uint64_t)(s >> 64) & 1;
    }
    if (c || !borrow) memcpy(r, d, 8 * N);
    else memcpy(r, t, 8 * N);
}

static inline void mod_sub_n(uint64_t *r, const uint64_t *a, const uint64_t *b,
                             const uint64_t *m, int N) {
    uint64_t t[8], borrow = 0, c = 0;
    for (int j = 0; j < N; j++) {
        u128 s = (u128)a[j] - b[j] - borrow;
        t[j] = (uint64_t)s;
        borrow = (uint64_t)(s >> 64) & 1;
    }
    if (borrow) {
        for (int j = 0; j < N; j++) {
            u128 s = (u128)t[j] + m[j] + c;
            t[j] = (uint64_t)s;
            c = (uint64_t)(s >> 64);
        }
    }
    memcpy(r, t, 8 * N);
}

static inline int is_zero_n(const uint64_t *a, int N) {
    uint64_t acc = 0;
    for (int j = 0; j < N; j++) acc |= a[j];
    return acc == 0;
}

/* -1 / 0 / 1 comparison of plain integers */
static inline int cmp_n(const uint64_t *a, const uint64_t *b, int N) {
    for (int j = N - 1; j >= 0; j--) {
        if (a[j] != b[j]) return a[j] < b[j] ? -1 : 1;
    }
    return 0;
}

static uint64_t neg_inv64(uint64_t m0) {
    uint64_t inv = 1;
    for (int i = 0; i < 7; i++) inv *= 2 - m0 * inv;
    return (uint64_t)0 - inv;
}

/* ------------------------------------------------------------------- Fp */

static inline void fp_mul(fp *r, const fp *a, const fp *b) { mont_mul_n(r->v, a->v, b->v, P_MOD, P_N0, 6); }
static inline void fp_sqr(fp *r, const fp *a) { mont_mul_n(r->v, a->v, a->v, P_MOD, P_N0, 6); }
static inline void fp_add(fp *r, const fp *a, const fp *b) { mod_add_n(r->v, a->v, b->v, P_MOD, 6); }
static inline void fp_sub(fp *r, const fp *a, const fp *b) { mod_sub_n(r->v, a->v, b->v, P_MOD, 6); }
static inline int fp_is_zero(const fp *a) { return is_zero_n(a->v, 6); }
static inline int fp_eq(const fp *a, const fp *b) { return memcmp(a->v, b->v, 48) == 0; }
static inline void fp_neg(fp *r, const fp *a) {
    fp z = {{0}};
    if (fp_is_zero(a)) *r = *a;
    else fp_sub(r, &z, a);
}

static void fp_pow(fp *r, const fp *a, const uint64_t *e, int nl) {
    fp acc = FP_ONE;
    for (int i = nl - 1; i >= 0; i--) {
        for (int b = 63; b >= 0; b--) {
            fp_sqr(&acc, &acc);
            if ((e[i] >> b) & 1) fp_mul(&acc, &acc, a);
        }
    }
    *r = acc;
}

static void fp_inv(fp *r, const fp *a) { fp_pow(r, a, P_MINUS_2, 6); }

static void fp_from_plain(fp *r, const uint64_t *plain) { mont_mul_n(r->v, plain, FP_R2.v, P_MOD, P_N0, 6); }
static void fp_to_plain(uint64_t *plain, const fp *a) {
    const uint64_t one[6] = {1, 0, 0, 0, 0, 0};
    mont_mul_n(plain, a->v, one, P_MOD, P_N0, 6);
}

/* ------------------------------------------------------------------- Fr */

static inline void fr_mul(fr *r, const fr *a, const fr *b) { mont_mul_n(r->v, a->v, b->v, R_MOD, R_N0, 4); }
static inline void fr_add(fr *r, const fr *a, const fr *b) { mod_add_n(r->v, a->v, b->v, R_MOD, 4); }
static inline void fr_sub(fr *r, const fr *a, const fr *b) { mod_sub_n(r->v, a->v, b->v, R_MOD, 4); }

static void fr_pow(fr *r, const fr *a, const uint64_t *e, int nl) {
    fr acc = FR_ONE;
    for (int i = nl - 1; i >= 0; i--) {
        for (int b = 63; b >= 0; b--) {
            fr_mul(&acc, &acc, &acc);
            if ((e[i] >> b) & 1) fr_mul(&acc, &acc, a);
        }
    }
    *r = acc;
}

/* a^(r-2): the reference's per-element division (k_ew_div.comp:27-29). 0 -> 0. */
static void fr_inv(fr *r, const fr *a) { fr_pow(r, a, R_MINUS_2, 4); }

/* Any 256-bit plain value -> Montgomery form of (value mod r).  CIOS with R^2
 * keeps the result below 2r for inputs below 2^256, so one conditional
 * subtraction reduces it fully (k_cells_to_fr.comp:5-6). */
static void fr_from_plain(fr *r, const uint64_t *plain) { mont_mul_n(r->v, plain, FR_R2.v, R_MOD, R_N0, 4); }
static void fr_to_plain(uint64_t *plain, const fr *a) {
    const uint64_t one[4] = {1, 0, 0, 0};
    mont_mul_n(plain, a->v, one, R_MOD, R_N0, 4);
}

static void be_to_limbs(uint64_t *limbs, const uint8_t *be, int N) {
    for (int k = 0; k < N; k++) {
        uint64_t w = 0;
        const uint8_t *src = be + 8 * (N - 1 - k);
        for (int b = 0; b < 8; b++) w = (w << 8) | src[b];
        limbs[k] = w;
    }
}

static void limbs_to_be(uint8_t *be, const uint64_t *limbs, int N) {
    for (int k = 0; k < N; k++) {
        uint64_t w = limbs[k];
        uint8_t *dst = be + 8 * (N - 1 - k);
        for (int b = 7; b >= 0; b--) {
            dst[b] = (uint8_t)w;
            w >>= 8;
        }
    }
}

/* Canonical check then Montgomery (k_blob_to_fr.comp:27-34). */
static int fr_from_be_canonical(fr *r, const uint8_t be[32]) {
    uint64_t plain[4];
    be_to_limbs(plain, be, 4);
    if (cmp_n(plain, R_MOD, 4) >= 0) return -1;
    fr_from_plain(r, plain);
    return 0;
}

static void fr_from_be_reduce(fr *r, const uint8_t be[32]) {
    uint64_t plain[4];
    be_to_limbs(plain, be, 4);
    fr_from_plain(r, plain);
}

static void fr_to_be(uint8_t be[32], const fr *a) {
    uint64_t plain[4];
    fr_to_plain(plain, a);
    limbs_to_be(be, plain, 4);
}

static void fr_from_u64(fr *r, uint64_t x) {
    uint64_t plain[4] = {x, 0, 0, 0};
    fr_from_plain(r, plain);
}

/* ------------------------------------------------------------------- G1 */

static void g1_set_identity(g1j *p) {
    memset(p, 0, sizeof(*p));
    p->y = FP_ONE;
}

static void g1_dbl(g1j *r, const g1j *p) {
    if (fp_is_zero(&p->z)) { *r = *p; return; }
    fp A, B, C, D, E, F, t0, t1;
    fp_sqr(&A, &p->x);
    fp_sqr(&B, &p->y);
    fp_sqr(&C, &B);
    fp_add(&t0, &p->x, &B);
    fp_sqr(&t0, &t0);
    fp_sub(&t0, &t0, &A);
    fp_sub(&t0, &t0, &C);
    fp_add(&D, &t0, &t0);
    fp_add(&E, &A, &A);
    fp_add(&E, &E, &A);
    fp_sqr(&F, &E);
    g1j out;
    fp_add(&t1, &D, &D);
    fp_sub(&out.x, &F, &t1);
    fp_sub(&t1, &D, &out.x);
    fp_mul(&t1, &E, &t1);
    fp_add(&C, &C, &C);
    fp_add(&C, &C, &C);
    fp_add(&C, &C, &C);
    fp_sub(&out.y, &t1, &C);
    fp_mul(&out.z, &p->y, &p->z);
    fp_add(&out.z, &out.z, &out.z);
    *r = out;
}

static void g1_add(g1j *r, const g1j *a, const g1j *b) {
    if (fp_is_zero(&a->z)) { *r = *b; return; }
    if (fp_is_zero(&b->z)) { *r = *a; return; }
    fp Z1Z1, Z2Z2, U1, U2, S1, S2, t;
    fp_sqr(&Z1Z1, &a->z);
    fp_sqr(&Z2Z2, &b->z);
    fp_mul(&U1, &a->x, &Z2Z2);
    fp_mul(&U2, &b->x, &Z1Z1);
    fp_mul(&t, &b->z, &Z2Z2);
    fp_mul(&S1, &a->y, &t);
    fp_mul(&t, &a->z, &Z1Z1);
    fp_mul(&S2, &b->y, &t);
    if (fp_eq(&U1, &U2)) {
        if (fp_eq(&S1, &S2)) { g1_dbl(r, a); return; }
        g1_set_identity(r);
        return;
    }
    fp H, I, J, K, V;
    fp_sub(&H, &U2, &U1);
    fp_add(&I, &H, &H);
    fp_sqr(&I, &I);
    fp_mul(&J, &H, &I);
    fp_sub(&K, &S2, &S1);
    fp_add(&K, &K, &K);
    fp_mul(&V, &U1, &I);
    g1j out;
    fp_sqr(&out.x, &K);
    fp_sub(&out.x, &out.x, &J);
    fp_sub(&out.x, &out.x, &V);
    fp_sub(&out.x, &out.x, &V);
    fp_sub(&t, &V, &out.x);
    fp_mul(&t, &K, &t);
    fp_mul(&S1, &S1, &J);
    fp_add(&S1, &S1, &S1);
    fp_sub(&out.y, &t, &S1);
    fp_add(&t, &a->z, &b->z);
    fp_sqr(&t, &t);
    fp_sub(&t, &t, &Z1Z1);
    fp_sub(&t, &t, &Z2Z2);
    fp_mul(&out.z, &t, &H);
    *r = out;
}

static void g1_madd(g1j *r, const g1j *a, const g1a *b) {
    if (b->inf) { *r = *a; return; }
    if (fp_is_zero(&a->z)) {
        r->x = b->x;
        r->y = b->y;
        r->z = FP_ONE;
        return;
    }
    fp Z1Z1, U2, S2, t;
    fp_sqr(&Z1Z1, &a->z);
    fp_mul(&U2, &b->x, &Z1Z1);
    fp_mul(&t, &a->z, &Z1Z1);
    fp_mul(&S2, &b->y, &t);
    if (fp_eq(&a->x, &U2)) {
        if (fp_eq(&a->y, &S2)) { g1_dbl(r, a); return; }
        g1_set_identity(r);
        return;
    }
    fp H, HH, I, J, K, V;
    fp_sub(&H, &U2, &a->x);
    fp_sqr(&HH, &H);
    fp_add(&I, &HH, &HH);
    fp_add(&I, &I, &I);
    fp_mul(&J, &H, &I);
    fp_sub(&K, &S2, &a->y);
    fp_add(&K, &K, &K);
    fp_mul(&V, &a->x, &I);
    g1j out;
    fp_sqr(&out.x, &K);
    fp_sub(&out.x, &out.x, &J);
    fp_sub(&out.x, &out.x, &V);
    fp_sub(&out.x, &out.x, &V);
    fp_sub(&t, &V, &out.x);
    fp_mul(&t, &K, &t);
    fp_mul(&J, &a->y, &J);
    fp_add(&J, &J, &J);
    fp_sub(&out.y, &t, &J);
    fp_add(&t, &a->z, &H);
    fp_sqr(&t, &t);
    fp_sub(&t, &t, &Z1Z1);
    fp_sub(&out.z, &t, &HH);
    *r = out;
}

static void g1_neg(g1j *r, const g1j *p) {
    *r = *p;
    fp_neg(&r->y, &p->y);
}

static void g1a_to_j(g1j *r, const g1a *a) {
    if (a->inf) { g1_set_identity(r); return; }
    r->x = a->x;
    r->y = a->y;
    r->z = FP_ONE;
}

static void g1_to_affine(g1a *r, const g1j *p) {
    if (fp_is_zero(&p->z)) { memset(r, 0, sizeof(*r)); r->inf = 1; return; }
    fp zi, z2;
    fp_inv(&zi, &p->z);
    fp_sqr(&z2, &zi);
    fp_mul(&r->x, &p->x, &z2);
    fp_mul(&z2, &z2, &zi);
    fp_mul(&r->y, &p->y, &z2);
    r->inf = 0;
}

/* Montgomery's simultaneous inversion over n Jacobian points (k_normalize.comp:29-56). */
static void g1_batch_to_affine(g1a *out, const g1j *in, size_t n) {
    fp *pref = (fp *)malloc(sizeof(fp) * (n + 1));
    fp acc = FP_ONE;
    for (size_t k = 0; k < n; k++) {
        pref[k] = acc;
        if (!fp_is_zero(&in[k].z)) fp_mul(&acc, &acc, &in[k].z);
    }
    fp inv;
    fp_inv(&inv, &acc);
    for (size_t k = n; k-- > 0;) {
        if (fp_is_zero(&in[k].z)) { memset(&out[k], 0, sizeof(g1a)); out[k].inf = 1; continue; }
        fp zi, z2;
        fp_mul(&zi, &inv, &pref[k]);
        fp_mul(&inv, &inv, &in[k].z);
        fp_sqr(&z2, &zi);
        fp_mul(&out[k].x, &in[k].x, &z2);
        fp_mul(&z2, &z2, &zi);
        fp_mul(&out[k].y, &in[k].y, &z2);
        out[k].inf = 0;
    }
    free(pref);
}

/* 4-bit fixed-window scalar multiplication by a plain 256-bit scalar. */
static void g1_mul_plain(g1j *r, const g1j *p, const uint64_t k[4]) {
    g1j tab[16];
    g1_set_identity(&tab[0]);
    tab[1] = *p;
    for (int i = 2; i < 16; i++) g1_add(&tab[i], &tab[i - 1], p);
    g1j acc;
    g1_set_identity(&acc);
    for (int w = 63; w >= 0; w--) {
        for (int d = 0; d < 4; d++) g1_dbl(&acc, &acc);
        const unsigned nib = (unsigned)(k[w / 16] >> (4 * (w % 16))) & 15u;
        if (nib) g1_add(&acc, &acc, &tab[nib]);
    }
    *r = acc;
}

static void g1_mul_fr(g1j *r, const g1j *p, const fr *k) {
    uint64_t plain[4];
    fr_to_plain(plain, k);
    g1_mul_plain(r, p, plain);
}

static int fp_is_lex_largest(const fp *a) {
    uint64_t plain[6];
    fp_to_plain(plain, a);
    return cmp_n(plain, P_HALF, 6) > 0;
}

static void g1a_compress(uint8_t out[48], const g1a *a) {
    if (a->inf) {
        memset(out, 0, 48);
        out[0] = 0xc0;
        return;
    }
    uint64_t plain[6];
    fp_to_plain(plain, &a->x);
    limbs_to_be(out, plain, 6);
    out[0] |= 0x80;
    if (fp_is_lex_largest(&a->y)) out[0] |= 0x20;
}

static void g1_compress(uint8_t out[48], const g1j *p) {
    g1a a;
    g1_to_affine(&a, p);
    g1a_compress(out, &a);
}

static int fp_sqrt(fp *r, const fp *a) {
    fp s, chk;
    fp_pow(&s, a, P_SQRT_EXP, 6);
    fp_sqr(&chk, &s);
    if (!fp_eq(&chk, a)) return -1;
    *r = s;
    return 0;
}

static int g1a_decompress(g1a *r, const uint8_t in[48]) {
    const uint8_t b0 = in[0];
    if (!(b0 & 0x80)) return -1;
    if (b0 & 0x40) {
        if (b0 & 0x3f) return -1;
        for (int k = 1; k < 48; k++) if (in[k]) return -1;
        memset(r, 0, sizeof(*r));
        r->inf = 1;
        return 0;
    }
    uint8_t buf[48];
    memcpy(buf, in, 48);
    buf[0] &= 0x1f;
    uint64_t plain[6];
    be_to_limbs(plain, buf, 6);
    if (cmp_n(plain, P_MOD, 6) >= 0) return -1;
    fp x, rhs, y;
    fp_from_plain(&x, plain);
    fp_sqr(&rhs, &x);
    fp_mul(&rhs, &rhs, &x);
    fp_add(&rhs, &rhs, &FP_FOUR);
    if (fp_sqrt(&y, &rhs)) return -1;
    if (fp_is_lex_largest(&y) != !!(b0 & 0x20)) fp_neg(&y, &y);
    r->x = x;
    r->y = y;
    r->inf = 0;
    return 0;
}

/* ----------------------------------------------------------------- setup */

static void compute_constants(void) {
    P_N0 = neg_inv64(P_MOD[0]);
    R_N0 = neg_inv64(R_MOD[0]);
    /* R mod m and R^2 mod m by repeated doubling. */
    uint64_t x[6] = {1, 0, 0, 0, 0, 0};
    for (int i = 0; i < 384; i++) mod_add_n(x, x, x, P_MOD, 6);
    memcpy(FP_ONE.v, x, 48);
    for (int i = 0; i < 384; i++) mod_add_n(x, x, x, P_MOD, 6);
    memcpy(FP_R2.v, x, 48);
    uint64_t y[4] = {1, 0, 0, 0};
    for (int i = 0; i < 256; i++) mod_add_n(y, y, y, R_MOD, 4);
    memcpy(FR_ONE.v, y, 32);
    for (int i = 0; i < 256; i++) mod_add_n(y, y, y, R_MOD, 4);
    memcpy(FR_R2.v, y, 32);

    memcpy(P_MINUS_2, P_MOD, 48);
    P_MINUS_2[0] -= 2;
    memcpy(R_MINUS_2, R_MOD, 32);
    R_MINUS_2[0] -= 2;
    /* (p+1)/4 and (p-1)/2 */
    uint64_t t[6];
    memcpy(t, P_MOD, 48);
    t[0] += 1; /* p is odd and its low limb is not all-ones: no carry */
    for (int k = 0; k < 6; k++) P_SQRT_EXP[k] = (t[k] >> 2) | (k < 5 ? t[k + 1] << 62 : 0);
    memcpy(t, P_MOD, 48);
    t[0] -= 1;
    for (int k = 0; k < 6; k++) P_HALF[k] = (t[k] >> 1) | (k < 5 ? t[k + 1] << 63 : 0);

    uint64_t four[6] = {4, 0, 0, 0, 0, 0};
    fp_from_plain(&FP_FOUR, four);

    /* omega_8192 = 7^((r-1)/2^13): r-1 = 2^32 * odd, so shift right by 13. */
    uint64_t e[4];
    memcpy(e, R_MOD, 32);
    e[0] -= 1;
    for (int k = 0; k < 4; k++) e[k] = (e[k] >> 13) | (k < 3 ? e[k + 1] << 51 : 0);
    fr seven, w;
    fr_from_u64(&seven, 7);
    fr_pow(&w, &seven, e, 4);
    ROOTS[0] = FR_ONE;
    for (int k = 1; k < 8192; k++) fr_mul(&ROOTS[k], &ROOTS[k - 1], &w);
    fr inv7;
    fr_inv(&inv7, &seven);
    SHIFT_POW[0] = FR_ONE;
    SHIFT_INV_POW[0] = FR_ONE;
    for (int k = 1; k < 8192; k++) {
        fr_mul(&SHIFT_POW[k], &SHIFT_POW[k - 1], &seven);
        fr_mul(&SHIFT_INV_POW[k], &SHIFT_INV_POW[k - 1], &inv7);
    }

    /* The standard BLS12-381 G1 generator. */
    static const uint8_t gen[48] = {
        0x97, 0xf1, 0xd3, 0xa7, 0x31, 0x97, 0xd7, 0x94, 0x26, 0x95, 0x63, 0x8c,
        0x4f, 0xa9, 0xac, 0x0f, 0xc3, 0x68, 0x8c, 0x4f, 0x97, 0x74, 0xb9, 0x05,
        0xa1, 0x4e, 0x3a, 0x3f, 0x17, 0x1b, 0xac, 0x58, 0x6c, 0x55, 0xe8, 0x3f,
        0xf9, 0x7a, 0x1a, 0xef, 0xfb, 0x3a, 0xf0, 0x0a, 0xdb, 0x22, 0xc6, 0xbb};
    g1a ga;
    g1a_decompress(&ga, gen);
    g1a_to_j(&G1_GEN, &ga);
}

__attribute__((constructor)) static void oracle_init(void) { compute_constants(); }

static inline uint32_t brev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int b = 0; b < bits; b++) r |= ((x >> b) & 1u) << (bits - 1 - b);
    return r;
}

/* ------------------------------------------------------------------- NTT */

/* In-place radix-2 NTT of size n | 8192, natural order in and out.  The
 * inverse is unscaled here; callers scale by 1/n (src/vkzg.cpp:1052-1056). */
static void fr_ntt_inplace(fr *a, uint32_t n, int inverse) {
    int logn = 0;
    while ((1u << logn) < n) logn++;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t j = brev(i, logn);
        if (j > i) { fr t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (uint32_t len = 2; len <= n; len <<= 1) {
        const uint32_t step = 8192 / len;
        for (uint32_t i = 0; i < n; i += len) {
            for (uint32_t j = 0; j < len / 2; j++) {
                const uint32_t e = (step * j) & 8191u;
                const fr *w = &ROOTS[inverse ? ((8192 - e) & 8191u) : e];
                fr u = a[i + j], v;
                fr_mul(&v, &a[i + j + len / 2], w);
                fr_add(&a[i + j], &u, &v);
                fr_sub(&a[i + j + len / 2], &u, &v);
            }
        }
    }
}

static void fr_scale_inv_n(fr *a, uint32_t n) {
    fr nn, ninv;
    fr_from_u64(&nn, n);
    fr_inv(&ninv, &nn);
    for (uint32_t i = 0; i < n; i++) fr_mul(&a[i], &a[i], &ninv);
}

/* G1 FFT of size 128 in natural order; twiddles are scalar multiplications. */
static void g1_fft128(g1j *a, int inverse) {
    const uint32_t n = 128;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t j = brev(i, 7);
        if (j > i) { g1j t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (uint32_t len = 2; len <= n; len <<= 1) {
        const uint32_t step = 8192 / len;
        for (uint32_t i = 0; i < n; i += len) {
            for (uint32_t j = 0; j < len / 2; j++) {
                const uint32_t e = (step * j) & 8191u;
                g1j u = a[i + j], v = a[i + j + len / 2], nv;
                if (e) g1_mul_fr(&v, &v, &ROOTS[inverse ? ((8192 - e) & 8191u) : e]);
                g1_add(&a[i + j], &u, &v);
                g1_neg(&nv, &v);
                g1_add(&a[i + j + len / 2], &u, &nv);
            }
        }
    }
    if (inverse) {
        fr nn, ninv;
        fr_from_u64(&nn, n);
        fr_inv(&ninv, &nn);
        for (uint32_t i = 0; i < n; i++) g1_mul_fr(&a[i], &a[i], &ninv);
    }
}

/* ------------------------------------------------------------ byte helpers */

void or_modulus_r(uint8_t out[32]) { limbs_to_be(out, R_MOD, 4); }
void or_modulus_p(uint8_t out[48]) { limbs_to_be(out, P_MOD, 6); }

void or_fr_mul_be(uint8_t out[32], const uint8_t a[32], const uint8_t b[32]) {
    fr x, y;
    fr_from_be_reduce(&x, a);
    fr_from_be_reduce(&y, b);
    fr_mul(&x, &x, &y);
    fr_to_be(out, &x);
}

void or_fp_mul_be(uint8_t out[48], const uint8_t a[48], const uint8_t b[48]) {
    uint64_t pa[6], pb[6], pr[6];
    be_to_limbs(pa, a, 6);
    be_to_limbs(pb, b, 6);
    fp x, y;
    fp_from_plain(&x, pa);
    fp_from_plain(&y, pb);
    fp_mul(&x, &x, &y);
    fp_to_plain(pr, &x);
    limbs_to_be(out, pr, 6);
}

void or_fr_inv_be(uint8_t out[32], const uint8_t a[32]) {
    fr x;
    fr_from_be_reduce(&x, a);
    fr_inv(&x, &x);
    fr_to_be(out, &x);
}

void or_root_8192_be(uint8_t out[32], uint32_t k) { fr_to_be(out, &ROOTS[k & 8191u]); }

int or_g1_roundtrip(uint8_t out[48], const uint8_t in[48]) {
    g1a a;
    if (g1a_decompress(&a, in)) return -1;
    g1a_compress(out, &a);
    return 0;
}

void or_g1_mul_gen(uint8_t out[48], const uint8_t k[32]) {
    uint64_t s[4];
    be_to_limbs(s, k, 4);
    g1j r;
    g1_mul_plain(&r, &G1_GEN, s);
    g1_compress(out, &r);
}

int or_g1_mul(uint8_t out[48], const uint8_t p[48], const uint8_t k[32]) {
    g1a a;
    if (g1a_decompress(&a, p)) return -1;
    g1j pj, r;
    g1a_to_j(&pj, &a);
    uint64_t s[4];
    be_to_limbs(s, k, 4);
    g1_mul_plain(&r, &pj, s);
    g1_compress(out, &r);
    return 0;
}

int or_g1_add(uint8_t out[48], const uint8_t p[48], const uint8_t q[48]) {
    g1a a, b;
    if (g1a_decompress(&a, p) || g1a_decompress(&b, q)) return -1;
    g1j aj, bj, r;
    g1a_to_j(&aj, &a);
    g1a_to_j(&bj, &b);
    g1_add(&r, &aj, &bj);
    g1_compress(out, &r);
    return 0;
}

/* ------------------------------------------------------- blob -> polynomial */

/* blob[i] = p(omega_4096^brev12(i)); k_blob_to_fr.comp stores element i at
 * brev12(i), then src/vkzg.cpp:928-935 runs the scaled inverse NTT_4096. */
static int blob_to_coeffs(fr *coeffs, const uint8_t *blob) {
    for (uint32_t i = 0; i < 4096; i++) {
        if (fr_from_be_canonical(&coeffs[brev(i, 12)], blob + 32 * i)) return -1;
    }
    fr_ntt_inplace(coeffs, 4096, 1);
    fr_scale_inv_n(coeffs, 4096);
    return 0;
}

int or_blob_to_poly(uint8_t coeffs[4096 * 32], const uint8_t blob[131072]) {
    fr *c = (fr *)malloc(sizeof(fr) * 4096);
    int rc = blob_to_coeffs(c, blob);
    if (!rc) for (int i = 0; i < 4096; i++) fr_to_be(coeffs + 32 * i, &c[i]);
    free(c);
    return rc;
}

/* Cell c holds ext[brev13(64c + t)], ext[k] = p(omega_8192^k). */
int or_blob_to_cells(uint8_t cells[128 * 2048], const uint8_t blob[131072]) {
    fr *e = (fr *)calloc(8192, sizeof(fr));
    int rc = blob_to_coeffs(e, blob);
    if (!rc) {
        fr_ntt_inplace(e, 8192, 0);
        for (uint32_t n = 0; n < 8192; n++) fr_to_be(cells + 32 * n, &e[brev(n, 13)]);
    }
    free(e);
    return rc;
}

void or_fr_ntt(uint8_t *data, uint32_t n, int inverse) {
    fr *a = (fr *)malloc(sizeof(fr) * n);
    for (uint32_t i = 0; i < n; i++) fr_from_be_reduce(&a[i], data + 32 * i);
    fr_ntt_inplace(a, n, inverse);
    if (inverse) fr_scale_inv_n(a, n);
    for (uint32_t i = 0; i < n; i++) fr_to_be(data + 32 * i, &a[i]);
    free(a);
}

/* -------------------------------------------------------------- recovery */

/* One blob of src/vkzg.cpp:1086-1179, kernel by kernel. */
static void recover_one(uint8_t *out, const uint8_t *cells, const uint32_t *missing, uint32_t count) {
    fr *E = (fr *)calloc(8192, sizeof(fr));
    fr *Z = (fr *)calloc(8192, sizeof(fr));
    fr *T = (fr *)calloc(8192, sizeof(fr));
    fr *A = (fr *)calloc(8192, sizeof(fr));
    fr *B = (fr *)calloc(8192, sizeof(fr));

    /* k_cells_to_fr: wire element n -> natural index brev13(n), reduced mod r. */
    for (uint32_t n = 0; n < 8192; n++) fr_from_be_reduce(&E[brev(n, 13)], cells + 32 * n);

    /* k_vanishing_poly: prod (y - omega_8192^(64*m)) over the missing list,
     * coefficient i at slot 64i; Z = 1 when nothing is missing. */
    fr zc[65];
    memset(zc, 0, sizeof(zc));
    zc[0] = FR_ONE;
    for (uint32_t k = 0; k < count; k++) {
        fr neg_root, zero = {{0}};
        fr_sub(&neg_root, &zero, &ROOTS[(missing[k] * 64u) & 8191u]);
        for (uint32_t d = k + 1; d > 0; d--) {
            fr t;
            fr_mul(&t, &zc[d], &neg_root);
            fr_add(&zc[d], &t, &zc[d - 1]);
        }
        fr_mul(&zc[0], &zc[0], &neg_root);
    }
    for (uint32_t i = 0; i <= count; i++) Z[64 * i] = zc[i];

    /* NTT(Z); E*Z; iNTT; x 1/8192 */
    memcpy(T, Z, sizeof(fr) * 8192);
    fr_ntt_inplace(T, 8192, 0);
    for (uint32_t k = 0; k < 8192; k++) fr_mul(&A[k], &E[k], &T[k]);
    fr_ntt_inplace(A, 8192, 1);
    fr_scale_inv_n(A, 8192);
    /* coset evaluations of (E*Z) and Z */
    for (uint32_t k = 0; k < 8192; k++) fr_mul(&A[k], &A[k], &SHIFT_POW[k]);
    fr_ntt_inplace(A, 8192, 0);
    for (uint32_t k = 0; k < 8192; k++) fr_mul(&B[k], &Z[k], &SHIFT_POW[k]);
    fr_ntt_inplace(B, 8192, 0);
    /* k_ew_div: per-element Fermat inversion */
    for (uint32_t k = 0; k < 8192; k++) {
        fr inv;
        fr_inv(&inv, &B[k]);
        fr_mul(&A[k], &A[k], &inv);
    }
    /* coset iNTT, 1/8192, un-shift, NTT */
    fr_ntt_inplace(A, 8192, 1);
    fr_scale_inv_n(A, 8192);
    for (uint32_t k = 0; k < 8192; k++) fr_mul(&A[k], &A[k], &SHIFT_INV_POW[k]);
    fr_ntt_inplace(A, 8192, 0);
    /* k_fr_to_cells */
    for (uint32_t n = 0; n < 8192; n++) fr_to_be(out + 32 * n, &A[brev(n, 13)]);
    free(E);
    free(Z);
    free(T);
    free(A);
    free(B);
}

int or_recover_cells(uint8_t *cells_out, const uint8_t *cells, const uint8_t *present,
                     size_t num_blobs) {
    for (size_t b = 0; b < num_blobs; b++) {
        uint32_t missing[128];
        uint32_t count = 0;
        for (uint32_t i = 0; i < 128; i++) {
            if (!present[128 * b + i]) {
                if (count >= 64) return 1;
                missing[count++] = brev(i, 7); /* src/vkzg.cpp:1043-1050,1094-1104 */
            }
        }
        recover_one(cells_out + (size_t)262144 * b, cells + (size_t)262144 * b, missing, count);
    }
    return 0;
}

/* ------------------------------------------------------------------ FK20 */

struct or_setup {
    g1a srs[4096];
    g1a xext[128][64];  /* x_ext_fft[j][i] */
    g1a *table;         /* [128][64][32]: 2^(8d) * xext[j][i] */
};

or_setup *or_setup_new(const uint8_t *g1) {
    or_setup *s = (or_setup *)calloc(1, sizeof(or_setup));
    for (int k = 0; k < 4096; k++) {
        if (g1a_decompress(&s->srs[k], g1 + 48 * k)) { free(s); return NULL; }
    }
    /* x_i[t] = s[4096-64-1-i-64t] for t < 63, identity elsewhere; G1-FFT_128. */
    g1j *cols = (g1j *)malloc(sizeof(g1j) * 64 * 128);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < 64; i++) {
        g1j *x = cols + 128 * i;
        for (int t = 0; t < 128; t++) {
            if (t < 63) g1a_to_j(&x[t], &s->srs[4096 - 64 - 1 - i - 64 * t]);
            else g1_set_identity(&x[t]);
        }
        g1_fft128(x, 0);
    }
    g1j *xj = (g1j *)malloc(sizeof(g1j) * 128 * 64);
    for (int j = 0; j < 128; j++)
        for (int i = 0; i < 64; i++) xj[64 * j + i] = cols[128 * i + j];
    g1_batch_to_affine(&s->xext[0][0], xj, 128 * 64);
    free(cols);

    /* Position table: 2^(8d) multiples, d < 32. */
    g1j *tj = (g1j *)malloc(sizeof(g1j) * 128 * 64 * 32);
#pragma omp parallel for schedule(static)
    for (int ji = 0; ji < 128 * 64; ji++) {
        g1j acc = xj[ji];
        for (int d = 0; d < 32; d++) {
            tj[32 * ji + d] = acc;
            for (int k = 0; k < 8; k++) g1_dbl(&acc, &acc);
        }
    }
    s->table = (g1a *)malloc(sizeof(g1a) * 128 * 64 * 32);
#pragma omp parallel for schedule(static)
    for (int chunk = 0; chunk < 128; chunk++) {
        g1_batch_to_affine(s->table + (size_t)chunk * 64 * 32, tj + (size_t)chunk * 64 * 32, 64 * 32);
    }
    free(tj);
    free(xj);
    return s;
}

void or_setup_free(or_setup *s) {
    if (!s) return;
    free(s->table);
    free(s);
}

void or_setup_xext(const or_setup *s, uint32_t j, uint32_t i, uint8_t out[48]) {
    g1a_compress(out, &s->xext[j & 127][i & 63]);
}

/* Circulant columns (k_build_circulant.comp:30-72): tc_i[0] = c[4095-i],
 * tc_i[128-m] = c[4095-i-64m] for m = 1..62, zero elsewhere; NTT_128. */
static void build_circulant(fr coeffs[128][64], const fr *c) {
    for (int i = 0; i < 64; i++) {
        fr col[128];
        memset(col, 0, sizeof(col));
        col[0] = c[4095 - i];
        for (int m = 1; m <= 62; m++) col[128 - m] = c[4095 - i - 64 * m];
        fr_ntt_inplace(col, 128, 0);
        for (int j = 0; j < 128; j++) coeffs[j][i] = col[j];
    }
}

/* Phase A for output j: signed 8-bit digits (k_phase_a_sort.comp:31-40),
 * bucket accumulation against the position table (k_phase_a.comp:29-43) and
 * the running-sum reduction sum_{k=1..128} k*B[k-1] (field.glsl.inc:478-517). */
static void phase_a_output(g1j *u, const or_setup *s, const fr scalars[64], int j) {
    g1j buckets[128];
    for (int k = 0; k < 128; k++) g1_set_identity(&buckets[k]);
    for (int i = 0; i < 64; i++) {
        uint64_t plain[4];
        fr_to_plain(plain, &scalars[i]);
        int carry = 0;
        for (int dd = 0; dd < 32; dd++) {
            int v = (int)((plain[dd / 8] >> (8 * (dd % 8))) & 0xff) + carry;
            carry = 0;
            if (v > 128) { v -= 256; carry = 1; }
            if (v == 0) continue;
            g1a pt = s->table[((size_t)j * 64 + i) * 32 + dd];
            if (v < 0) fp_neg(&pt.y, &pt.y);
            g1_madd(&buckets[(v < 0 ? -v : v) - 1], &buckets[(v < 0 ? -v : v) - 1], &pt);
        }
    }
    g1j run, sum;
    g1_set_identity(&run);
    g1_set_identity(&sum);
    for (int k = 127; k >= 0; k--) {
        g1_add(&run, &run, &buckets[k]);
        g1_add(&sum, &sum, &run);
    }
    *u = sum;
}

static void fk20_one(const or_setup *s, uint8_t *proofs, const fr *c) {
    fr (*coeffs)[64] = (fr (*)[64])malloc(sizeof(fr) * 128 * 64);
    build_circulant(coeffs, c);
    g1j u[128];
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < 128; j++) phase_a_output(&u[j], s, coeffs[j], j);
    free(coeffs);
    /* Phase B: h = IFFT_128(u) truncated to 64, out = FFT_128(h). */
    g1_fft128(u, 1);
    for (int t = 64; t < 128; t++) g1_set_identity(&u[t]);
    g1_fft128(u, 0);
    g1a aff[128];
    g1_batch_to_affine(aff, u, 128);
    for (int k = 0; k < 128; k++) g1a_compress(proofs + 48 * brev(k, 7), &aff[k]);
}

int or_compute_proofs(const or_setup *s, uint8_t *proofs, const uint8_t *blobs, size_t num_blobs) {
    fr *c = (fr *)malloc(sizeof(fr) * 4096);
    for (size_t b = 0; b < num_blobs; b++) {
        if (blob_to_coeffs(c, blobs + (size_t)131072 * b)) { free(c); return 6; }
        fk20_one(s, proofs + (size_t)6144 * b, c);
    }
    free(c);
    return 0;
}

/* Pippenger (c = 8, unsigned windows) over affine bases, for the direct form. */
static void msm_pippenger(g1j *r, const g1a *bases, const fr *scalars, size_t n) {
    uint64_t (*plain)[4] = (uint64_t (*)[4])malloc(sizeof(uint64_t) * 4 * n);
    for (size_t k = 0; k < n; k++) fr_to_plain(plain[k], &scalars[k]);
    g1j acc;
    g1_set_identity(&acc);
    for (int w = 31; w >= 0; w--) {
        for (int d = 0; d < 8; d++) g1_dbl(&acc, &acc);
        g1j buckets[255];
        for (int k = 0; k < 255; k++) g1_set_identity(&buckets[k]);
        for (size_t k = 0; k < n; k++) {
            const unsigned dig = (unsigned)(plain[k][w / 8] >> (8 * (w % 8))) & 0xffu;
            if (dig) g1_madd(&buckets[dig - 1], &buckets[dig - 1], &bases[k]);
        }
        g1j run, sum;
        g1_set_identity(&run);
        g1_set_identity(&sum);
        for (int k = 254; k >= 0; k--) {
            g1_add(&run, &run, &buckets[k]);
            g1_add(&sum, &sum, &run);
        }
        g1_add(&acc, &acc, &sum);
    }
    free(plain);
    *r = acc;
}

int or_compute_proofs_direct(const or_setup *s, uint8_t *proofs, const uint8_t *blob) {
    fr *c = (fr *)malloc(sizeof(fr) * 4096);
    if (blob_to_coeffs(c, blob)) { free(c); return 6; }
    g1j h[128];
#pragma omp parallel for schedule(dynamic)
    for (int m = 1; m <= 63; m++) {
        msm_pippenger(&h[m - 1], s->srs, c + 64 * m, (size_t)(4096 - 64 * m));
    }
    for (int t = 63; t < 128; t++) g1_set_identity(&h[t]);
    g1_fft128(h, 0);
    g1a aff[128];
    g1_batch_to_affine(aff, h, 128);
    for (int k = 0; k < 128; k++) g1a_compress(proofs + 48 * brev(k, 7), &aff[k]);
    free(c);
    return 0;
}

/* ------------------------------------------------- seeded tau and closed form */

void or_setup_g1_from_tau(uint8_t *out, const uint8_t tau[32]) {
    fr t;
    fr_from_be_reduce(&t, tau);
    fr *pw = (fr *)malloc(sizeof(fr) * 4096);
    pw[0] = FR_ONE;
    for (int i = 1; i < 4096; i++) fr_mul(&pw[i], &pw[i - 1], &t);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < 4096; i++) {
        g1j r;
        g1_mul_fr(&r, &G1_GEN, &pw[i]);
        g1_compress(out + 48 * i, &r);
    }
    free(pw);
}

int or_proofs_from_tau(uint8_t *proofs, const uint8_t *blob, const uint8_t tau[32]) {
    fr *c = (fr *)malloc(sizeof(fr) * 4096);
    if (blob_to_coeffs(c, blob)) { free(c); return 6; }
    fr t;
    fr_from_be_reduce(&t, tau);
#pragma omp parallel for schedule(dynamic)
    for (int cell = 0; cell < 128; cell++) {
        /* q = p div (X^64 - a): q_k = c_{k+64} + a * q_{k+64}, from the top. */
        const fr a = ROOTS[(64u * brev((uint32_t)cell, 7)) & 8191u];
        fr q[4096 - 64];
        for (int k = 4095 - 64; k >= 0; k--) {
            fr v = c[k + 64];
            if (k + 64 <= 4095 - 64) {
                fr tmp;
                fr_mul(&tmp, &a, &q[k + 64]);
                fr_add(&v, &v, &tmp);
            }
            q[k] = v;
        }
        /* q(tau) by Horner */
        fr acc = {{0}};
        for (int k = 4095 - 64; k >= 0; k--) {
            fr_mul(&acc, &acc, &t);
            fr_add(&acc, &acc, &q[k]);
        }
        g1j r;
        g1_mul_fr(&r, &G1_GEN, &acc);
        g1_compress(proofs + 48 * cell, &r);
    }
    free(c);
    return 0;
}

/* ============================================================================
 * Pairing and batch cell-proof verification (test-side check of the proofs;
 * the reference has no verifier, include/vkzg.h:14-18).
 *
 * Written from the mathematics of BLS12-381, not from any library:
 *   Fp2 = Fp[u]/(u^2 + 1), Fp6 = Fp2[v]/(v^3 - xi) with xi = 1 + u,
 *   Fp12 = Fp6[w]/(w^2 - v), so w^6 = xi.
 *   G2 lives on the twist E': y^2 = x^3 + 4 xi over Fp2; the untwist
 *   psi(x, y) = (x w^-2, y w^-3) maps it into E(Fp12): y^2 = x^3 + 4.
 *   Miller loop over |z| = 0xd201000000010000 (the curve parameter z is
 *   negative; inverting every Miller value does not change whether a product
 *   of pairings is 1, nor bilinearity), with lines computed in affine
 *   coordinates over Fp12.  Vertical lines take values in Fp6 and are dropped:
 *   the final exponentiation (p^12 - 1)/r is a multiple of p^6 - 1.
 *   Final exponentiation: f^(p^6 - 1) = conj(f) / f (the p^6-power Frobenius
 *   negates w), then the power (p^6 + 1)/r by square-and-multiply.
 * Every part is checked by the tests through bilinearity, non-degeneracy and
 * group orders rather than by known answers, since no G2 data ships with the
 * reference.
 * ========================================================================== */

typedef struct { fp c0, c1; } fp2;
typedef struct { fp2 c0, c1, c2; } fp6;
typedef struct { fp6 d0, d1; } fp12;
typedef struct { fp2 x, y; int inf; } g2a;

/* (p^6 + 1) / r, little-endian 64-bit limbs. */
static const uint64_t FINAL_EXP_HARD[32] = {
    0x8739e1cdc0705d6aull, 0x09a5256de0381a16ull, 0x9cf0f70a61c791e2ull, 0x3a09c4497903f76eull, 0x2d7271563890f133ull, 0x224741b36fec7760ull, 0x338259c22a12bd40ull, 0x38ee1cd4778e0de7ull, 0xc3b5ef4b188a20b0ull, 0x1d615d49e2764d7bull, 0x816101ddd076117dull, 0xf007c01e7ebe3afcull, 0x27d7bd90935021c3ull, 0xc3b5e2f557c0b15full, 0x5e886c94c4f82384ull, 0xee6a95db11e63f56ull, 0x2b822f514a9c4f6full, 0x12d6a874d21b73daull, 0x1304275ef499dffbull, 0x967878febcb95d1full, 0x4744497f8b2f2922ull, 0x85a2e707f0841855ull, 0x9f0c50126c802eecull, 0xfb46e197bd2fa489ull, 0x548ce0809bc5f61aull, 0xcf56fb1573beaa8cull, 0xad7375a3763bdf7cull, 0xe0ec9031179bdeccull, 0x6579aea83c48c1daull, 0xdbf85ae664cf5bb3ull, 0x7b6f235c55ca7566ull, 0x000028b314877503ull};
/* G2 cofactor h2 = (z^8 - 4z^7 + 5z^6 - 4z^4 + 6z^3 - 4z^2 - 4z + 13) / 9, z = -0xd201000000010000. */
static const uint64_t G2_COFACTOR[8] = {
    0xcf1c38e31c7238e5ull, 0x1616ec6e786f0c70ull, 0x21537e293a6691aeull, 0xa628f1cb4d9e82efull, 0xa68a205b2e5a7ddfull, 0xcd91de4547085abaull, 0x091d50792876a202ull, 0x05d543a95414e7f1ull};
static const uint64_t ATE_LOOP = 0xd201000000010000ULL;

static void fp2_add(fp2 *r, const fp2 *a, const fp2 *b) { fp_add(&r->c0, &a->c0, &b->c0); fp_add(&r->c1, &a->c1, &b->c1); }
static void fp2_sub(fp2 *r, const fp2 *a, const fp2 *b) { fp_sub(&r->c0, &a->c0, &b->c0); fp_sub(&r->c1, &a->c1, &b->c1); }
static void fp2_neg(fp2 *r, const fp2 *a) { fp_neg(&r->c0, &a->c0); fp_neg(&r->c1, &a->c1); }
static int fp2_is_zero(const fp2 *a) { return fp_is_zero(&a->c0) && fp_is_zero(&a->c1); }
static int fp2_eq(const fp2 *a, const fp2 *b) { return fp_eq(&a->c0, &b->c0) && fp_eq(&a->c1, &b->c1); }
static void fp2_zero(fp2 *r) { memset(r, 0, sizeof(*r)); }
static void fp2_one(fp2 *r) { fp2_zero(r); r->c0 = FP_ONE; }

static void fp2_mul(fp2 *r, const fp2 *a, const fp2 *b) {
    fp t0, t1, t2, t3;
    fp_mul(&t0, &a->c0, &b->c0);
    fp_mul(&t1, &a->c1, &b->c1);
    fp_mul(&t2, &a->c0, &b->c1);
    fp_mul(&t3, &a->c1, &b->c0);
    fp_sub(&r->c0, &t0, &t1); /* u^2 = -1 */
    fp_add(&r->c1, &t2, &t3);
}
static void fp2_sqr(fp2 *r, const fp2 *a) { fp2_mul(r, a, a); }
static void fp2_mul_fp(fp2 *r, const fp2 *a, const fp *b) { fp_mul(&r->c0, &a->c0, b); fp_mul(&r->c1, &a->c1, b); }
/* a * xi = (a0 - a1) + (a0 + a1) u */
static void fp2_mul_xi(fp2 *r, const fp2 *a) {
    fp t0, t1;
    fp_sub(&t0, &a->c0, &a->c1);
    fp_add(&t1, &a->c0, &a->c1);
    r->c0 = t0;
    r->c1 = t1;
}
static void fp2_inv(fp2 *r, const fp2 *a) {
    fp n, t;
    fp_sqr(&n, &a->c0);
    fp_sqr(&t, &a->c1);
    fp_add(&n, &n, &t);
    fp_inv(&n, &n);
    fp_mul(&r->c0, &a->c0, &n);
    fp_mul(&t, &a->c1, &n);
    fp_neg(&r->c1, &t);
}
/* Square root through the norm: for a = a0 + a1 u with norm n^2 = a0^2 + a1^2,
 * x0^2 = (a0 +- n)/2 and x1 = a1 / (2 x0).  Returns 0 and the root, or -1. */
static int fp2_sqrt(fp2 *r, const fp2 *a) {
    fp two, inv2;
    fp_add(&two, &FP_ONE, &FP_ONE);
    fp_inv(&inv2, &two);
    fp2 cand;
    if (fp_is_zero(&a->c1)) {
        fp s, na;
        if (!fp_sqrt(&s, &a->c0)) { cand.c0 = s; memset(&cand.c1, 0, sizeof(fp)); }
        else {
            fp_neg(&na, &a->c0);
            if (fp_sqrt(&s, &na)) return -1;
            memset(&cand.c0, 0, sizeof(fp));
            cand.c1 = s;
        }
    } else {
        fp n2, t, n, x0sq, x0, d;
        fp_sqr(&n2, &a->c0);
        fp_sqr(&t, &a->c1);
        fp_add(&n2, &n2, &t);
        if (fp_sqrt(&n, &n2)) return -1;
        fp_add(&x0sq, &a->c0, &n);
        fp_mul(&x0sq, &x0sq, &inv2);
        if (fp_sqrt(&x0, &x0sq)) {
            fp_sub(&x0sq, &a->c0, &n);
            fp_mul(&x0sq, &x0sq, &inv2);
            if (fp_sqrt(&x0, &x0sq)) return -1;
        }
        fp_add(&d, &x0, &x0);
        fp_inv(&d, &d);
        cand.c0 = x0;
        fp_mul(&cand.c1, &a->c1, &d);
    }
    fp2 chk;
    fp2_sqr(&chk, &cand);
    if (!fp2_eq(&chk, a)) return -1;
    *r = cand;
    return 0;
}

static void fp6_add(fp6 *r, const fp6 *a, const fp6 *b) { fp2_add(&r->c0, &a->c0, &b->c0); fp2_add(&r->c1, &a->c1, &b->c1); fp2_add(&r->c2, &a->c2, &b->c2); }
static void fp6_sub(fp6 *r, const fp6 *a, const fp6 *b) { fp2_sub(&r->c0, &a->c0, &b->c0); fp2_sub(&r->c1, &a->c1, &b->c1); fp2_sub(&r->c2, &a->c2, &b->c2); }
static void fp6_neg(fp6 *r, const fp6 *a) { fp2_neg(&r->c0, &a->c0); fp2_neg(&r->c1, &a->c1); fp2_neg(&r->c2, &a->c2); }
static void fp6_zero(fp6 *r) { memset(r, 0, sizeof(*r)); }

static void fp6_mul(fp6 *r, const fp6 *a, const fp6 *b) {
    fp2 t, s, c0, c1, c2;
    /* c0 = a0 b0 + xi (a1 b2 + a2 b1) */
    fp2_mul(&t, &a->c1, &b->c2);
    fp2_mul(&s, &a->c2, &b->c1);
    fp2_add(&t, &t, &s);
    fp2_mul_xi(&t, &t);
    fp2_mul(&s, &a->c0, &b->c0);
    fp2_add(&c0, &s, &t);
    /* c1 = a0 b1 + a1 b0 + xi a2 b2 */
    fp2_mul(&t, &a->c2, &b->c2);
    fp2_mul_xi(&t, &t);
    fp2_mul(&s, &a->c0, &b->c1);
    fp2_add(&t, &t, &s);
    fp2_mul(&s, &a->c1, &b->c0);
    fp2_add(&c1, &t, &s);
    /* c2 = a0 b2 + a1 b1 + a2 b0 */
    fp2_mul(&t, &a->c0, &b->c2);
    fp2_mul(&s, &a->c1, &b->c1);
    fp2_add(&t, &t, &s);
    fp2_mul(&s, &a->c2, &b->c0);
    fp2_add(&c2, &t, &s);
    r->c0 = c0;
    r->c1 = c1;
    r->c2 = c2;
}
/* a * v = xi a2 + a0 v + a1 v^2 */
static void fp6_mul_v(fp6 *r, const fp6 *a) {
    fp2 t;
    fp2_mul_xi(&t, &a->c2);
    r->c2 = a->c1;
    r->c1 = a->c0;
    r->c0 = t;
}
static void fp6_inv(fp6 *r, const fp6 *a) {
    fp2 t0, t1, t2, s, n;
    fp2_sqr(&t0, &a->c0);
    fp2_mul(&s, &a->c1, &a->c2);
    fp2_mul_xi(&s, &s);
    fp2_sub(&t0, &t0, &s); /* a0^2 - xi a1 a2 */
    fp2_sqr(&t1, &a->c2);
    fp2_mul_xi(&t1, &t1);
    fp2_mul(&s, &a->c0, &a->c1);
    fp2_sub(&t1, &t1, &s); /* xi a2^2 - a0 a1 */
    fp2_sqr(&t2, &a->c1);
    fp2_mul(&s, &a->c0, &a->c2);
    fp2_sub(&t2, &t2, &s); /* a1^2 - a0 a2 */
    fp2_mul(&n, &a->c2, &t1);
    fp2_mul(&s, &a->c1, &t2);
    fp2_add(&n, &n, &s);
    fp2_mul_xi(&n, &n);
    fp2_mul(&s, &a->c0, &t0);
    fp2_add(&n, &n, &s);
    fp2_inv(&n, &n);
    fp2_mul(&r->c0, &t0, &n);
    fp2_mul(&r->c1, &t1, &n);
    fp2_mul(&r->c2, &t2, &n);
}

static void fp12_one(fp12 *r) { memset(r, 0, sizeof(*r)); r->d0.c0.c0 = FP_ONE; }
static int fp12_eq(const fp12 *a, const fp12 *b) { return memcmp(a, b, sizeof(fp12)) == 0; }
static void fp12_add(fp12 *r, const fp12 *a, const fp12 *b) { fp6_add(&r->d0, &a->d0, &b->d0); fp6_add(&r->d1, &a->d1, &b->d1); }
static void fp12_sub(fp12 *r, const fp12 *a, const fp12 *b) { fp6_sub(&r->d0, &a->d0, &b->d0); fp6_sub(&r->d1, &a->d1, &b->d1); }
static void fp12_mul(fp12 *r, const fp12 *a, const fp12 *b) {
    fp6 t0, t1, t2, s;
    fp6_mul(&t0, &a->d0, &b->d0);
    fp6_mul(&t1, &a->d1, &b->d1);
    fp6_mul(&t2, &a->d0, &b->d1);
    fp6_mul(&s, &a->d1, &b->d0);
    fp6_add(&r->d1, &t2, &s);
    fp6_mul_v(&t1, &t1);
    fp6_add(&r->d0, &t0, &t1);
}
static void fp12_sqr(fp12 *r, const fp12 *a) { fp12_mul(r, a, a); }
static void fp12_conj(fp12 *r, const fp12 *a) { r->d0 = a->d0; fp6_neg(&r->d1, &a->d1); }
/* 1/(a + b w) = (a - b w) / (a^2 - b^2 v) */
static void fp12_inv(fp12 *r, const fp12 *a) {
    fp6 t0, t1;
    fp6_mul(&t0, &a->d0, &a->d0);
    fp6_mul(&t1, &a->d1, &a->d1);
    fp6_mul_v(&t1, &t1);
    fp6_sub(&t0, &t0, &t1);
    fp6_inv(&t0, &t0);
    fp6_mul(&r->d0, &a->d0, &t0);
    fp6_mul(&t1, &a->d1, &t0);
    fp6_neg(&r->d1, &t1);
}
static void fp12_pow_limbs(fp12 *r, const fp12 *a, const uint64_t *e, int nl) {
    fp12 acc;
    fp12_one(&acc);
    for (int i = nl * 64 - 1; i >= 0; i--) {
        fp12_sqr(&acc, &acc);
        if ((e[i / 64] >> (i % 64)) & 1) fp12_mul(&acc, &acc, a);
    }
    *r = acc;
}
static void fp12_from_fp(fp12 *r, const fp *a) { memset(r, 0, sizeof(*r)); r->d0.c0.c0 = *a; }
static void fp12_from_fp2(fp12 *r, const fp2 *a) { memset(r, 0, sizeof(*r)); r->d0.c0 = *a; }

/* ---- G2 on the twist (affine; speed is irrelevant here) ---- */
static void g2_b(fp2 *b) { fp2 four; fp2_zero(&four); four.c0 = FP_FOUR; fp2_mul_xi(b, &four); }
static int g2_on_curve(const g2a *p) {
    if (p->inf) return 1;
    fp2 l, rr, b;
    fp2_sqr(&l, &p->y);
    fp2_sqr(&rr, &p->x);
    fp2_mul(&rr, &rr, &p->x);
    g2_b(&b);
    fp2_add(&rr, &rr, &b);
    return fp2_eq(&l, &rr);
}
static void g2_add(g2a *r, const g2a *a, const g2a *b) {
    if (a->inf) { *r = *b; return; }
    if (b->inf) { *r = *a; return; }
    fp2 lam, t, x3, y3;
    if (fp2_eq(&a->x, &b->x)) {
        fp2 sy;
        fp2_add(&sy, &a->y, &b->y);
        if (fp2_is_zero(&sy)) { memset(r, 0, sizeof(*r)); r->inf = 1; return; }
        /* doubling: lam = 3 x^2 / (2 y) */
        fp2_sqr(&lam, &a->x);
        fp2_add(&t, &lam, &lam);
        fp2_add(&lam, &lam, &t);
        fp2_add(&t, &a->y, &a->y);
    } else {
        fp2_sub(&lam, &b->y, &a->y);
        fp2_sub(&t, &b->x, &a->x);
    }
    fp2_inv(&t, &t);
    fp2_mul(&lam, &lam, &t);
    fp2_sqr(&x3, &lam);
    fp2_sub(&x3, &x3, &a->x);
    fp2_sub(&x3, &x3, &b->x);
    fp2_sub(&y3, &a->x, &x3);
    fp2_mul(&y3, &y3, &lam);
    fp2_sub(&y3, &y3, &a->y);
    r->x = x3;
    r->y = y3;
    r->inf = 0;
}
static void g2_mul_limbs(g2a *r, const g2a *p, const uint64_t *k, int nl) {
    g2a acc;
    memset(&acc, 0, sizeof(acc));
    acc.inf = 1;
    for (int i = nl * 64 - 1; i >= 0; i--) {
        g2_add(&acc, &acc, &acc);
        if ((k[i / 64] >> (i % 64)) & 1) g2_add(&acc, &acc, p);
    }
    *r = acc;
}

/* 192-byte encoding used by this oracle: x.c0, x.c1, y.c0, y.c1, each 48-byte
 * big-endian canonical; all-zero bytes = infinity. */
static void g2_to_bytes(uint8_t out[192], const g2a *p) {
    if (p->inf) { memset(out, 0, 192); return; }
    const fp *c[4] = {&p->x.c0, &p->x.c1, &p->y.c0, &p->y.c1};
    for (int k = 0; k < 4; k++) {
        uint64_t plain[6];
        fp_to_plain(plain, c[k]);
        limbs_to_be(out + 48 * k, plain, 6);
    }
}
static int g2_from_bytes(g2a *p, const uint8_t in[192]) {
    int zero = 1;
    for (int k = 0; k < 192; k++) if (in[k]) { zero = 0; break; }
    memset(p, 0, sizeof(*p));
    if (zero) { p->inf = 1; return 0; }
    fp *c[4] = {&p->x.c0, &p->x.c1, &p->y.c0, &p->y.c1};
    for (int k = 0; k < 4; k++) {
        uint64_t plain[6];
        be_to_limbs(plain, in + 48 * k, 6);
        if (cmp_n(plain, P_MOD, 6) >= 0) return -1;
        fp_from_plain(c[k], plain);
    }
    return g2_on_curve(p) ? 0 : -1;
}

/* A point of order r on E'(Fp2): the first x = i + u (i = 1, 2, ...) with a
 * square right-hand side, times the cofactor. */
static void g2_generator(g2a *q) {
    for (uint64_t i = 1;; i++) {
        uint64_t pi[6] = {i, 0, 0, 0, 0, 0};
        g2a p;
        memset(&p, 0, sizeof(p));
        fp_from_plain(&p.x.c0, pi);
        p.x.c1 = FP_ONE;
        fp2 rhs, b;
        fp2_sqr(&rhs, &p.x);
        fp2_mul(&rhs, &rhs, &p.x);
        g2_b(&b);
        fp2_add(&rhs, &rhs, &b);
        if (fp2_sqrt(&p.y, &rhs)) continue;
        g2_mul_limbs(q, &p, G2_COFACTOR, 8);
        if (!q->inf) return;
    }
}

/* ---- Miller loop ---- */
typedef struct { fp12 x, y; } e12pt;

/* line through t with slope lam, evaluated at P = (xp, yp): (yp - yt) - lam (xp - xt) */
static void line_eval(fp12 *l, const e12pt *t, const fp12 *lam, const fp12 *xp, const fp12 *yp) {
    fp12 a, b;
    fp12_sub(&a, yp, &t->y);
    fp12_sub(&b, xp, &t->x);
    fp12_mul(&b, &b, lam);
    fp12_sub(l, &a, &b);
}

static void miller(fp12 *f, const g1a *p, const g2a *q) {
    fp12_one(f);
    if (p->inf || q->inf) return;
    /* untwist Q into E(Fp12) */
    fp12 w, winv, w2inv, w3inv;
    memset(&w, 0, sizeof(w));
    w.d1.c0.c0 = FP_ONE;
    fp12_inv(&winv, &w);
    fp12_sqr(&w2inv, &winv);
    fp12_mul(&w3inv, &w2inv, &winv);
    e12pt Q, T;
    fp12_from_fp2(&Q.x, &q->x);
    fp12_mul(&Q.x, &Q.x, &w2inv);
    fp12_from_fp2(&Q.y, &q->y);
    fp12_mul(&Q.y, &Q.y, &w3inv);
    T = Q;
    fp12 xp, yp;
    fp12_from_fp(&xp, &p->x);
    fp12_from_fp(&yp, &p->y);
    int top = 63;
    while (!((ATE_LOOP >> top) & 1)) top--;
    for (int i = top - 1; i >= 0; i--) {
        /* doubling step: lam = 3 X^2 / (2 Y) */
        fp12 lam, t, l, x3, y3;
        fp12_sqr(&lam, &T.x);
        fp12_add(&t, &lam, &lam);
        fp12_add(&lam, &lam, &t);
        fp12_add(&t, &T.y, &T.y);
        fp12_inv(&t, &t);
        fp12_mul(&lam, &lam, &t);
        line_eval(&l, &T, &lam, &xp, &yp);
        fp12_sqr(f, f);
        fp12_mul(f, f, &l);
        fp12_sqr(&x3, &lam);
        fp12_sub(&x3, &x3, &T.x);
        fp12_sub(&x3, &x3, &T.x);
        fp12_sub(&y3, &T.x, &x3);
        fp12_mul(&y3, &y3, &lam);
        fp12_sub(&y3, &y3, &T.y);
        T.x = x3;
        T.y = y3;
        if ((ATE_LOOP >> i) & 1) {
            /* addition step with Q: lam = (YQ - Y) / (XQ - X) */
            fp12_sub(&lam, &Q.y, &T.y);
            fp12_sub(&t, &Q.x, &T.x);
            fp12_inv(&t, &t);
            fp12_mul(&lam, &lam, &t);
            line_eval(&l, &T, &lam, &xp, &yp);
            fp12_mul(f, f, &l);
            fp12_sqr(&x3, &lam);
            fp12_sub(&x3, &x3, &T.x);
            fp12_sub(&x3, &x3, &Q.x);
            fp12_sub(&y3, &T.x, &x3);
            fp12_mul(&y3, &y3, &lam);
            fp12_sub(&y3, &y3, &T.y);
            T.x = x3;
            T.y = y3;
        }
    }
}

static void final_exp(fp12 *r, const fp12 *f) {
    fp12 c, inv;
    fp12_conj(&c, f);
    fp12_inv(&inv, f);
    fp12_mul(&c, &c, &inv); /* f^(p^6 - 1) */
    fp12_pow_limbs(r, &c, FINAL_EXP_HARD, 32);
}

static void fp12_to_bytes(uint8_t out[576], const fp12 *a) {
    const fp *c[12] = {&a->d0.c0.c0, &a->d0.c0.c1, &a->d0.c1.c0, &a->d0.c1.c1, &a->d0.c2.c0, &a->d0.c2.c1,
                       &a->d1.c0.c0, &a->d1.c0.c1, &a->d1.c1.c0, &a->d1.c1.c1, &a->d1.c2.c0, &a->d1.c2.c1};
    for (int k = 0; k < 12; k++) {
        uint64_t plain[6];
        fp_to_plain(plain, c[k]);
        limbs_to_be(out + 48 * k, plain, 6);
    }
}

void or_g2_generator(uint8_t out[192]) {
    g2a q;
    g2_generator(&q);
    g2_to_bytes(out, &q);
}

void or_g2_cofactor(uint8_t out[64]) { limbs_to_be(out, G2_COFACTOR, 8); }
void or_final_exp_hard(uint8_t out[256]) { limbs_to_be(out, FINAL_EXP_HARD, 32); }

int or_g2_on_curve(const uint8_t p[192]) {
    g2a a;
    return g2_from_bytes(&a, p) == 0;
}

int or_g2_mul(uint8_t out[192], const uint8_t p[192], const uint8_t *k_be, uint32_t k_len) {
    g2a a, r;
    if (g2_from_bytes(&a, p)) return -1;
    const int nl = (int)((k_len + 7) / 8);
    uint64_t *k = (uint64_t *)calloc((size_t)nl, 8);
    for (uint32_t b = 0; b < k_len; b++) {
        const uint32_t pos = k_len - 1 - b;  /* byte index from the least significant end */
        k[pos / 8] |= (uint64_t)k_be[b] << (8 * (pos % 8));
    }
    g2_mul_limbs(&r, &a, k, nl);
    free(k);
    g2_to_bytes(out, &r);
    return 0;
}

int or_g2_add(uint8_t out[192], const uint8_t p[192], const uint8_t q[192]) {
    g2a a, b, r;
    if (g2_from_bytes(&a, p) || g2_from_bytes(&b, q)) return -1;
    g2_add(&r, &a, &b);
    g2_to_bytes(out, &r);
    return 0;
}

/* A point on E'(Fp2) that is not cleared by the cofactor (for order tests). */
int or_g2_raw_point(uint8_t out[192], uint64_t seed) {
    for (uint64_t i = seed + 1;; i++) {
        uint64_t pi[6] = {i, 0, 0, 0, 0, 0};
        g2a p;
        memset(&p, 0, sizeof(p));
        fp_from_plain(&p.x.c0, pi);
        p.x.c1 = FP_ONE;
        fp2 rhs, b;
        fp2_sqr(&rhs, &p.x);
        fp2_mul(&rhs, &rhs, &p.x);
        g2_b(&b);
        fp2_add(&rhs, &rhs, &b);
        if (fp2_sqrt(&p.y, &rhs)) continue;
        g2_to_bytes(out, &p);
        return 0;
    }
}

/* [tau^i] Q for i = 0..n-1, Q = or_g2_generator. */
void or_g2_powers_from_tau(uint8_t *out, const uint8_t tau[32], uint32_t n) {
    fr t;
    fr_from_be_reduce(&t, tau);
    g2a q;
    g2_generator(&q);
    fr pw = FR_ONE;
    fr *pws = (fr *)malloc(sizeof(fr) * n);
    for (uint32_t i = 0; i < n; i++) { pws[i] = pw; fr_mul(&pw, &pw, &t); }
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < (int)n; i++) {
        uint64_t plain[4];
        fr_to_plain(plain, &pws[i]);
        g2a r;
        g2_mul_limbs(&r, &q, plain, 4);
        g2_to_bytes(out + 192 * (size_t)i, &r);
    }
    free(pws);
}

/* e(P, Q) = final_exp(miller(P, Q)) as 576 bytes (12 Fp, tower order). */
int or_pairing(uint8_t out[576], const uint8_t p[48], const uint8_t q[192]) {
    g1a a;
    g2a b;
    if (g1a_decompress(&a, p) || g2_from_bytes(&b, q)) return -1;
    fp12 f, e;
    miller(&f, &a, &b);
    final_exp(&e, &f);
    fp12_to_bytes(out, &e);
    return 0;
}

static int pairing_product_is_one(const g1a *ps, const g2a *qs, size_t n) {
    fp12 *fs = (fp12 *)malloc(sizeof(fp12) * (n ? n : 1));
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < (int)n; k++) miller(&fs[k], &ps[k], &qs[k]);
    fp12 acc, e, one;
    fp12_one(&acc);
    for (size_t k = 0; k < n; k++) fp12_mul(&acc, &acc, &fs[k]);
    free(fs);
    final_exp(&e, &acc);
    fp12_one(&one);
    return fp12_eq(&e, &one);
}

/* 1 when prod_k e(P_k, Q_k) == 1, 0 when not, -1 on a bad encoding. */
int or_pairing_check(const uint8_t *g1s, const uint8_t *g2s, size_t n) {
    g1a *ps = (g1a *)malloc(sizeof(g1a) * (n ? n : 1));
    g2a *qs = (g2a *)malloc(sizeof(g2a) * (n ? n : 1));
    int rc = 0;
    for (size_t k = 0; k < n && !rc; k++)
        if (g1a_decompress(&ps[k], g1s + 48 * k) || g2_from_bytes(&qs[k], g2s + 192 * k)) rc = -1;
    if (!rc) rc = pairing_product_is_one(ps, qs, n);
    free(ps);
    free(qs);
    return rc;
}

/* ---- batch verification of cell proofs ---- */
struct or_vsetup {
    g1a srs[4096];
    g2a g2_gen, g2_tau64;
};

or_vsetup *or_vsetup_new(const uint8_t *g1_monomial, const uint8_t *g2_powers) {
    or_vsetup *s = (or_vsetup *)calloc(1, sizeof(or_vsetup));
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int k = 0; k < 4096; k++) bad |= g1a_decompress(&s->srs[k], g1_monomial + 48 * k) != 0;
    if (bad || g2_from_bytes(&s->g2_gen, g2_powers) || g2_from_bytes(&s->g2_tau64, g2_powers + 192 * 64)) {
        free(s);
        return NULL;
    }
    return s;
}

void or_vsetup_free(or_vsetup *s) { free(s); }

/* Commitment [p(tau)]G1 = sum_i c_i [tau^i]G1 from the monomial setup. */
int or_commit(const or_vsetup *s, uint8_t out[48], const uint8_t blob[131072]) {
    fr *c = (fr *)malloc(sizeof(fr) * 4096);
    if (blob_to_coeffs(c, blob)) { free(c); return 6; }
    g1j r;
    msm_pippenger(&r, s->srs, c, 4096);
    g1_compress(out, &r);
    free(c);
    return 0;
}

/* Commitments of n blobs, in parallel over blobs; 0, or 6 if any blob is non-canonical. */
int or_commit_batch(const or_vsetup *s, uint8_t *out, const uint8_t *blobs, size_t n) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int b = 0; b < (int)n; b++) bad |= or_commit(s, out + 48 * (size_t)b, blobs + (size_t)131072 * b) != 0;
    return bad ? 6 : 0;
}

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* Cell proof k claims p(X) = q(X)(X^64 - z_k) + I_k(X), where cell c holds p
 * on the coset h_c H_64 (h_c = omega_8192^brev7(c), element j at
 * h_c omega_64^brev6(j)), z_k = h_c^64 and I_k interpolates the cell.  With
 * random weights r_k the 128-per-blob equations
 *   e(pi_k, [tau^64]Q - z_k Q) = e(C_k - [I_k(tau)]G1, Q)
 * fold into one product of two pairings:
 *   e(sum r_k pi_k, [tau^64]Q) * e(-(sum r_k (C_k + z_k pi_k) - [sum r_k I_k](tau)), Q) == 1.
 * This is the check of the consensus specs' verify_cell_kzg_proof_batch with
 * seeded weights instead of Fiat-Shamir ones.  Returns 1 (valid), 0 (invalid),
 * -1 (bad encoding, non-canonical cell element or cell index >= 128). */
int or_verify_cell_proofs(const or_vsetup *s, const uint8_t *commitments, const uint32_t *cell_indices,
                          const uint8_t *cells, const uint8_t *proofs, size_t n, uint64_t seed) {
    if (n == 0) return 1;
    g1a *pi = (g1a *)malloc(sizeof(g1a) * n);
    fr *w = (fr *)malloc(sizeof(fr) * n);      /* r_k */
    fr *wz = (fr *)malloc(sizeof(fr) * n);     /* r_k z_k */
    fr (*ik)[64] = (fr (*)[64])malloc(sizeof(fr) * 64 * n);
    int bad = 0;
    uint64_t st = seed;
    for (size_t k = 0; k < n; k++) {
        uint64_t plain[4];
        for (int q = 0; q < 4; q++) plain[q] = splitmix64(&st);
        plain[3] &= (1ULL << 62) - 1;  /* < 2^254 < r */
        fr_from_plain(&w[k], plain);
        if (cell_indices[k] >= 128) bad = 1;
    }
    if (bad) goto out;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int k = 0; k < (int)n; k++) {
        if (g1a_decompress(&pi[k], proofs + 48 * (size_t)k)) { bad |= 1; continue; }
        const uint32_t hc = brev(cell_indices[k], 7);
        const fr z = ROOTS[(64u * hc) & 8191u];
        fr_mul(&wz[k], &w[k], &z);
        /* natural-order values on h H_64, inverse NTT_64 -> coefficients of I(h X) */
        fr y[64];
        for (uint32_t j = 0; j < 64; j++)
            if (fr_from_be_canonical(&y[brev(j, 6)], cells + 2048 * (size_t)k + 32 * j)) bad |= 1;
        fr_ntt_inplace(y, 64, 1);
        fr_scale_inv_n(y, 64);
        const fr hinv = ROOTS[(8192u - hc) & 8191u];
        fr hp = w[k];  /* r_k h^-i */
        for (int i = 0; i < 64; i++) {
            fr_mul(&ik[k][i], &y[i], &hp);
            fr_mul(&hp, &hp, &hinv);
        }
    }
    if (bad) goto out;
    {
        fr agg[64];
        memset(agg, 0, sizeof(agg));
        for (size_t k = 0; k < n; k++)
            for (int i = 0; i < 64; i++) fr_add(&agg[i], &agg[i], &ik[k][i]);
        /* commitments: runs of equal bytes share one point and a summed weight */
        g1a *cs = (g1a *)malloc(sizeof(g1a) * n);
        fr *cw = (fr *)malloc(sizeof(fr) * n);
        size_t nc = 0;
        for (size_t k = 0; k < n; k++) {
            if (k && !memcmp(commitments + 48 * k, commitments + 48 * (k - 1), 48)) {
                fr_add(&cw[nc - 1], &cw[nc - 1], &w[k]);
                continue;
            }
            if (g1a_decompress(&cs[nc], commitments + 48 * k)) { bad = 1; break; }
            cw[nc++] = w[k];
        }
        if (!bad) {
            g1j A, B, t;
            msm_pippenger(&A, pi, w, n);
            msm_pippenger(&B, pi, wz, n);
            msm_pippenger(&t, cs, cw, nc);
            g1_add(&B, &B, &t);
            msm_pippenger(&t, s->srs, agg, 64);
            g1_neg(&t, &t);
            g1_add(&B, &B, &t);
            g1_neg(&B, &B);
            g1a ps[2];
            g2a qs[2] = {s->g2_tau64, s->g2_gen};
            g1_to_affine(&ps[0], &A);
            g1_to_affine(&ps[1], &B);
            bad = pairing_product_is_one(ps, qs, 2) ? 0 : 2;
        }
        free(cs);
        free(cw);
    }
out:
    free(pi);
    free(w);
    free(wz);
    free(ik);
    if (bad == 1) return -1;
    return bad ? 0 : 1;
}


/* ---- spec-weighted verification, G2 compression, MSM (test-side checks of
 * the device verifier) ---- */

/* As or_verify_cell_proofs, with the weights r^k of the consensus specs'
 * verify_cell_kzg_proof_batch for a given challenge r (32 bytes big-endian,
 * below r) instead of seeded ones.  The commitment sum is taken over cells
 * (sum r^k C_k), which is the specs' sum over deduplicated commitments with
 * summed weights.  out_a / out_b (may be NULL) receive, compressed,
 * A = sum r^k pi_k and RL = sum r^k (C_k + z_k pi_k) - [sum r^k I_k](tau).
 * Returns 1 (valid), 0 (invalid), -1 (bad encoding, non-canonical element or
 * cell index >= 128). */
int or_verify_cell_proofs_r(const or_vsetup *s, const uint8_t *commitments, const uint32_t *cell_indices,
                            const uint8_t *cells, const uint8_t *proofs, size_t n, const uint8_t r_be[32],
                            uint8_t out_a[48], uint8_t out_b[48]) {
    fr r;
    if (fr_from_be_canonical(&r, r_be)) return -1;
    size_t m = n ? n : 1;
    g1a *pi = (g1a *)malloc(sizeof(g1a) * m);
    g1a *cm = (g1a *)malloc(sizeof(g1a) * m);
    fr *w = (fr *)malloc(sizeof(fr) * m);
    fr *wz = (fr *)malloc(sizeof(fr) * m);
    fr (*ik)[64] = (fr (*)[64])malloc(sizeof(fr) * 64 * m);
    int bad = 0;
    fr pw = FR_ONE;
    for (size_t k = 0; k < n; k++) {
        w[k] = pw;
        fr_mul(&pw, &pw, &r);
        if (cell_indices[k] >= 128) bad = 1;
    }
    if (!bad) {
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (int k = 0; k < (int)n; k++) {
            if (g1a_decompress(&pi[k], proofs + 48 * (size_t)k) ||
                g1a_decompress(&cm[k], commitments + 48 * (size_t)k)) { bad |= 1; continue; }
            const uint32_t hc = brev(cell_indices[k], 7);
            const fr z = ROOTS[(64u * hc) & 8191u];
            fr_mul(&wz[k], &w[k], &z);
            fr y[64];
            for (uint32_t j = 0; j < 64; j++)
                if (fr_from_be_canonical(&y[brev(j, 6)], cells + 2048 * (size_t)k + 32 * j)) bad |= 1;
            fr_ntt_inplace(y, 64, 1);
            fr_scale_inv_n(y, 64);
            const fr hinv = ROOTS[(8192u - hc) & 8191u];
            fr hp = w[k];
            for (int i = 0; i < 64; i++) {
                fr_mul(&ik[k][i], &y[i], &hp);
                fr_mul(&hp, &hp, &hinv);
            }
        }
    }
    int rc = -1;
    if (!bad) {
        fr agg[64];
        memset(agg, 0, sizeof(agg));
        for (size_t k = 0; k < n; k++)
            for (int i = 0; i < 64; i++) fr_add(&agg[i], &agg[i], &ik[k][i]);
        g1j A, B, t;
        msm_pippenger(&A, pi, w, n);
        msm_pippenger(&B, pi, wz, n);
        msm_pippenger(&t, cm, w, n);
        g1_add(&B, &B, &t);
        msm_pippenger(&t, s->srs, agg, 64);
        g1_neg(&t, &t);
        g1_add(&B, &B, &t);
        if (out_a) g1_compress(out_a, &A);
        if (out_b) g1_compress(out_b, &B);
        g1_neg(&B, &B);
        g1a ps[2];
        g2a qs[2] = {s->g2_tau64, s->g2_gen};
        g1_to_affine(&ps[0], &A);
        g1_to_affine(&ps[1], &B);
        rc = pairing_product_is_one(ps, qs, 2) ? 1 : 0;
    }
    free(pi);
    free(cm);
    free(w);
    free(wz);
    free(ik);
    return rc;
}

/* ZCash compressed G2 (96 bytes: x.c1 then x.c0, big-endian, flags in the
 * first byte; y is "larger" when y.c1 > (p-1)/2, or y.c1 == 0 and
 * y.c0 > (p-1)/2) from / to this oracle's 192-byte encoding. */
static int fp2_is_lex_largest(const fp2 *y) {
    if (!fp_is_zero(&y->c1)) return fp_is_lex_largest(&y->c1);
    return fp_is_lex_largest(&y->c0);
}

int or_g2_compress(uint8_t out[96], const uint8_t in[192]) {
    g2a q;
    if (g2_from_bytes(&q, in)) return -1;
    memset(out, 0, 96);
    if (q.inf) { out[0] = 0xc0; return 0; }
    uint64_t plain[6];
    fp_to_plain(plain, &q.x.c1);
    limbs_to_be(out, plain, 6);
    fp_to_plain(plain, &q.x.c0);
    limbs_to_be(out + 48, plain, 6);
    out[0] |= 0x80;
    if (fp2_is_lex_largest(&q.y)) out[0] |= 0x20;
    return 0;
}

int or_g2_decompress(uint8_t out[192], const uint8_t in[96]) {
    const uint8_t b0 = in[0];
    if (!(b0 & 0x80)) return -1;
    g2a q;
    memset(&q, 0, sizeof(q));
    if (b0 & 0x40) {
        if (b0 & 0x3f) return -1;
        for (int k = 1; k < 96; k++) if (in[k]) return -1;
        q.inf = 1;
        g2_to_bytes(out, &q);
        return 0;
    }
    uint8_t buf[48];
    uint64_t plain[6];
    memcpy(buf, in, 48);
    buf[0] &= 0x1f;
    be_to_limbs(plain, buf, 6);
    if (cmp_n(plain, P_MOD, 6) >= 0) return -1;
    fp_from_plain(&q.x.c1, plain);
    be_to_limbs(plain, in + 48, 6);
    if (cmp_n(plain, P_MOD, 6) >= 0) return -1;
    fp_from_plain(&q.x.c0, plain);
    fp2 rhs, b;
    fp2_sqr(&rhs, &q.x);
    fp2_mul(&rhs, &rhs, &q.x);
    g2_b(&b);
    fp2_add(&rhs, &rhs, &b);
    if (fp2_sqrt(&q.y, &rhs)) return -1;
    if (fp2_is_lex_largest(&q.y) != !!(b0 & 0x20)) fp2_neg(&q.y, &q.y);
    g2_to_bytes(out, &q);
    return 0;
}

/* sum [s_i] P_i (points compressed, scalars 32 bytes big-endian below r). */
int or_g1_msm(uint8_t out[48], const uint8_t *points, const uint8_t *scalars, size_t n) {
    size_t m = n ? n : 1;
    g1a *ps = (g1a *)malloc(sizeof(g1a) * m);
    fr *ks = (fr *)malloc(sizeof(fr) * m);
    int bad = 0;
    for (size_t i = 0; i < n && !bad; i++)
        bad = g1a_decompress(&ps[i], points + 48 * i) || fr_from_be_canonical(&ks[i], scalars + 32 * i);
    if (!bad) {
        g1j r;
        msm_pippenger(&r, ps, ks, n);
        g1_compress(out, &r);
    }
    free(ps);
    free(ks);
    return bad ? -1 : 0;
}
