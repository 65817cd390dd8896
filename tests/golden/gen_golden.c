/*
 * Golden-vector generator (runs in the build container only; needs GMP).
 * Computes c = a*b mod (X^N+1, q) by Kronecker substitution with GMP's
 * big-integer multiply -- the independent oracle algorithm of the
 * reference's tests/validate.py:29-37 -- on inputs built from the seeded
 * streams of oracle/oracle.c and the case families of tests/validate.py:46-59.
 *
 * usage: gen_golden <outdir> <n> <seed:S | delta:D> <family> <input seed>
 * writes outdir/{q,a,b,c}.bin (little-endian uint32 limbs, 28 per coefficient)
 */
#include <gmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define L ORC_LIMBS

static void write_file(const char* dir, const char* name, const void* data, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE* f = fopen(path, "wb");
    if (!f || fwrite(data, 1, bytes, f) != bytes) { perror(path); exit(1); }
    fclose(f);
}

static void set_small(uint32_t* x, uint64_t v) {
    memset(x, 0, L * 4);
    x[0] = (uint32_t)v;
    x[1] = (uint32_t)(v >> 32);
}

static void q_minus(uint32_t* x, const uint32_t* q, uint32_t d) {
    uint64_t borrow = d;
    for (int i = 0; i < L; ++i) {
        uint64_t t = (uint64_t)q[i] - borrow;
        x[i] = (uint32_t)t;
        borrow = (t >> 63) & 1;
    }
}

static void build(const char* family, uint32_t n, const uint32_t* q, uint64_t seed,
                  uint32_t* a, uint32_t* b) {
    size_t bytes = (size_t)n * L * 4;
    uint32_t *r0 = malloc(bytes), *r1 = malloc(bytes);
    orc_gen_poly(seed, 0, n, q, r0);
    orc_gen_poly(seed, 1, n, q, r1);
    memset(a, 0, bytes);
    memset(b, 0, bytes);
    if (!strcmp(family, "dense-random-1")) { memcpy(a, r0, bytes); memcpy(b, r1, bytes); }
    else if (!strcmp(family, "dense-random-2")) { orc_gen_poly(seed, 2, n, q, a); orc_gen_poly(seed, 3, n, q, b); }
    else if (!strcmp(family, "zero")) { memcpy(a, r0, bytes); }
    else if (!strcmp(family, "identity")) { memcpy(a, r0, bytes); b[0] = 1; }
    else if (!strcmp(family, "wrap-sign")) {
        q_minus(a + (size_t)(n - 1) * L, q, 1);
        if (n > 1) q_minus(b + L, q, 1);
    } else if (!strcmp(family, "all-q-minus-one")) {
        for (uint32_t i = 0; i < n; ++i) { q_minus(a + (size_t)i * L, q, 1); q_minus(b + (size_t)i * L, q, 1); }
    } else if (!strcmp(family, "alternating")) {
        for (uint32_t i = 0; i < n; ++i) if (i & 1) q_minus(a + (size_t)i * L, q, 1);
        memcpy(b, r0, bytes);
    } else if (!strcmp(family, "limb-boundaries")) {
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t* x = a + (size_t)i * L;
            switch (i % 8) {
                case 0: set_small(x, 0); break;
                case 1: set_small(x, 1); break;
                case 2: q_minus(x, q, 1); break;
                case 3: q_minus(x, q, 2); break;
                case 4: set_small(x, 0); x[27] = 1u; break;               /* 2^864 */
                case 5: for (int w = 0; w < 27; ++w) x[w] = 0xFFFFFFFFu; break; /* 2^864-1 */
                case 6: set_small(x, 1ull << 32); break;
                case 7: set_small(x, 0xFFFFFFFFull); break;
            }
        }
        memcpy(b, r1, bytes);
    } else if (!strcmp(family, "random-after-edges")) { memcpy(a, r1, bytes); memcpy(b, r0, bytes); }
    else { fprintf(stderr, "unknown family %s\n", family); exit(2); }
    free(r0); free(r1);
}

int main(int argc, char** argv) {
    if (argc != 6) { fprintf(stderr, "usage\n"); return 2; }
    const char* out = argv[1];
    uint32_t n = (uint32_t)strtoul(argv[2], NULL, 10);
    uint32_t q[L];
    if (!strncmp(argv[3], "seed:", 5)) orc_gen_modulus(strtoull(argv[3] + 5, NULL, 10), q);
    else if (!strncmp(argv[3], "delta:", 6)) orc_special_modulus((uint32_t)strtoul(argv[3] + 6, NULL, 10), q);
    else return 2;
    uint64_t seed = strtoull(argv[5], NULL, 10);
    size_t bytes = (size_t)n * L * 4;
    uint32_t *a = malloc(bytes), *b = malloc(bytes), *c = malloc(bytes);
    build(argv[4], n, q, seed, a, b);

    mpz_t Q, x, y, z, s0, s1;
    mpz_inits(Q, x, y, z, s0, s1, NULL);
    mpz_import(Q, L, -1, 4, 0, 0, q);
    /* slot width > N*(q-1)^2 in bytes */
    mpz_sub_ui(s0, Q, 1);
    mpz_mul(s0, s0, s0);
    mpz_mul_ui(s0, s0, n);
    size_t width = (mpz_sizeinbase(s0, 2) + 7) / 8 + 1;
    unsigned char* buf = calloc((size_t)n * width, 1);
    for (uint32_t i = 0; i < n; ++i) memcpy(buf + (size_t)i * width, a + (size_t)i * L, L * 4);
    mpz_import(x, (size_t)n * width, -1, 1, 0, 0, buf);
    memset(buf, 0, (size_t)n * width);
    for (uint32_t i = 0; i < n; ++i) memcpy(buf + (size_t)i * width, b + (size_t)i * L, L * 4);
    mpz_import(y, (size_t)n * width, -1, 1, 0, 0, buf);
    free(buf);
    mpz_mul(z, x, y);
    unsigned char* raw = calloc((size_t)2 * n * width + 16, 1);
    size_t got = 0;
    mpz_export(raw, &got, -1, 1, 0, 0, z);
    for (uint32_t k = 0; k < n; ++k) {
        mpz_import(s0, width, -1, 1, 0, 0, raw + (size_t)k * width);
        mpz_import(s1, width, -1, 1, 0, 0, raw + (size_t)(k + n) * width);
        mpz_sub(s0, s0, s1);
        mpz_mod(s0, s0, Q);
        uint32_t* dst = c + (size_t)k * L;
        memset(dst, 0, L * 4);
        size_t words = 0;
        mpz_export(dst, &words, -1, 4, 0, 0, s0);
    }
    free(raw);
    write_file(out, "q.bin", q, L * 4);
    write_file(out, "a.bin", a, bytes);
    write_file(out, "b.bin", b, bytes);
    write_file(out, "c.bin", c, bytes);
    mpz_clears(Q, x, y, z, s0, s1, NULL);
    free(a); free(b); free(c);
    return 0;
}
