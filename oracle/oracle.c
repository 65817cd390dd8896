/*
 * oracle.c -- CPU restatement of the reference's algorithm for the exact
 * 868-bit negacyclic product.  TEST INFRASTRUCTURE ONLY: it is the checker
 * that tests/, smoke() and bench.py's cpu_baseline leg compare against; the
 * product path never calls it.  Plain C99 (+ POSIX threads), no other
 * dependencies.
 *
 * Reference anchors (paths relative to the reference tree):
 *   prime rule, tables, CRT bases ........ rns/host_setup.h:61-126
 *   Shoup / Barrett small-modulus ops ..... rns/solve.cu:109-125
 *   limb-power residue conversion ......... rns/solve.cu:138-157
 *   twisted cyclic NTT, pointwise, inverse  rns/solve.cu:225-264, 327-333
 *   nearest-integer CRT ................... rns/solve.cu:336-420
 *   odd-root CT / GS (Harvey) order ....... research/HARVEY_NTT_RU.md:5-14
 *   schoolbook definition ................. tests/validate.py:39-44
 * The special-q fold of rns/solve.cu:336-357 is replaced by an exact
 * reduction modulo any odd 868-bit q.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define L ORC_LIMBS

/* ------------------------------------------------------------------------ */
/* PRNG                                                                     */
/* ------------------------------------------------------------------------ */
uint64_t orc_mix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_value64(uint64_t seed, uint64_t tag, uint64_t index) {
    return orc_mix64(orc_mix64(seed ^ tag) + index);
}

#define TAG_MODULUS 0x51ull
#define TAG_POLY 0xA0000000ull

static void words_from_u64(const uint64_t* v, uint32_t* out) {
    for (int w = 0; w < L / 2; ++w) {
        out[2 * w] = (uint32_t)v[w];
        out[2 * w + 1] = (uint32_t)(v[w] >> 32);
    }
}

void orc_gen_modulus(uint64_t seed, uint32_t q[L]) {
    uint64_t v[L / 2];
    for (int w = 0; w < L / 2; ++w) v[w] = orc_value64(seed, TAG_MODULUS, (uint64_t)w);
    v[13] &= (1ull << 36) - 1; /* 868 = 13*64 + 36 */
    v[13] |= 1ull << 35;       /* bit 867 */
    v[0] |= 1ull;
    words_from_u64(v, q);
}

void orc_special_modulus(uint32_t delta, uint32_t q[L]) {
    /* 2^868 - delta = (2^868 - 1) - (delta - 1) for 0 < delta < 2^32 */
    for (int i = 0; i < L; ++i) q[i] = 0xFFFFFFFFu;
    q[L - 1] = 0xFu;
    q[0] -= delta - 1u;
}

static int cmp_words(const uint32_t* a, const uint32_t* b, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

void orc_gen_poly(uint64_t seed, uint32_t poly, uint32_t n, const uint32_t q[L],
                  uint32_t* out) {
    const uint64_t tag = TAG_POLY + poly;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t* x = out + (size_t)i * L;
        int accepted = 0;
        for (uint32_t t = 0; t < 64 && !accepted; ++t) {
            uint64_t v[L / 2];
            for (uint32_t w = 0; w < L / 2; ++w)
                v[w] = orc_value64(seed, tag, (((uint64_t)i * 64 + t) * 16) + w);
            v[13] &= (1ull << 36) - 1;
            words_from_u64(v, x);
            accepted = cmp_words(x, q, L) < 0;
        }
        if (!accepted) { /* value in [q, 2^868) < 2q: one subtraction */
            uint64_t borrow = 0;
            for (int w = 0; w < L; ++w) {
                uint64_t d = (uint64_t)x[w] - q[w] - borrow;
                x[w] = (uint32_t)d;
                borrow = (d >> 63) & 1;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* big integers: little-endian uint32 words                                 */
/* ------------------------------------------------------------------------ */
static uint32_t bits_of(const uint32_t* a, int n) {
    for (int i = n - 1; i >= 0; --i)
        if (a[i]) return (uint32_t)(i * 32 + 32 - __builtin_clz(a[i]));
    return 0;
}

static int bit_at(const uint32_t* a, uint32_t bit) { return (a[bit / 32] >> (bit % 32)) & 1; }

/* r = x mod q for any x of nx words; q has bit 867 as its top bit. */
static void mod_q(const uint32_t* x, int nx, const uint32_t* q, uint32_t* r) {
    const uint32_t qbits = bits_of(q, L);
    uint32_t xb = bits_of(x, nx);
    uint32_t acc[L + 1];
    memset(acc, 0, sizeof(acc));
    if (xb < qbits) {
        for (int i = 0; i < L; ++i) r[i] = i < nx ? x[i] : 0;
        return;
    }
    /* acc = x >> s with s = xb - (qbits-1): acc < 2^(qbits-1) <= q */
    uint32_t s = xb - (qbits - 1);
    for (uint32_t bit = 0; bit < qbits - 1; ++bit)
        if (bit_at(x, bit + s)) acc[bit / 32] |= 1u << (bit % 32);
    for (uint32_t k = s; k-- > 0;) {
        uint32_t carry = (uint32_t)bit_at(x, k);
        for (int i = 0; i <= L; ++i) {
            uint32_t next = acc[i] >> 31;
            acc[i] = (acc[i] << 1) | carry;
            carry = next;
        }
        /* acc < 2q: subtract once if acc >= q */
        int ge = acc[L] != 0 || cmp_words(acc, q, L) >= 0;
        if (ge) {
            uint64_t borrow = 0;
            for (int i = 0; i <= L; ++i) {
                uint64_t d = (uint64_t)acc[i] - (i < L ? q[i] : 0) - borrow;
                acc[i] = (uint32_t)d;
                borrow = (d >> 63) & 1;
            }
        }
    }
    memcpy(r, acc, L * sizeof(uint32_t));
}

/* acc[0..na) += a[0..L) * b[0..L); na large enough */
static void mul_add(uint32_t* acc, int na, const uint32_t* a, const uint32_t* b) {
    for (int i = 0; i < L; ++i) {
        uint64_t carry = 0;
        uint64_t ai = a[i];
        if (!ai) continue;
        int j = 0;
        for (; j < L; ++j) {
            uint64_t t = ai * b[j] + acc[i + j] + carry;
            acc[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        for (int k = i + j; carry && k < na; ++k) {
            uint64_t t = (uint64_t)acc[k] + carry;
            acc[k] = (uint32_t)t;
            carry = t >> 32;
        }
    }
}

/* acc += scalar * a (na words of a), into nacc words */
static void mul_small_add(uint32_t* acc, int nacc, const uint32_t* a, int na, uint32_t s) {
    uint64_t carry = 0;
    int i = 0;
    for (; i < na; ++i) {
        uint64_t t = (uint64_t)a[i] * s + acc[i] + carry;
        acc[i] = (uint32_t)t;
        carry = t >> 32;
    }
    for (; carry && i < nacc; ++i) {
        uint64_t t = (uint64_t)acc[i] + carry;
        acc[i] = (uint32_t)t;
        carry = t >> 32;
    }
}

/* r = (a - b) mod q for a, b in [0,q) */
static void sub_mod_q(const uint32_t* a, const uint32_t* b, const uint32_t* q, uint32_t* r) {
    uint64_t borrow = 0;
    for (int i = 0; i < L; ++i) {
        uint64_t d = (uint64_t)a[i] - b[i] - borrow;
        r[i] = (uint32_t)d;
        borrow = (d >> 63) & 1;
    }
    if (borrow) {
        uint64_t carry = 0;
        for (int i = 0; i < L; ++i) {
            uint64_t t = (uint64_t)r[i] + q[i] + carry;
            r[i] = (uint32_t)t;
            carry = t >> 32;
        }
    }
}

int orc_schoolbook(uint32_t n, const uint32_t q[L], const uint32_t* a, const uint32_t* b,
                   uint32_t* c) {
    enum { ACC = 2 * L + 4 };
    uint32_t pos[ACC], neg[ACC], rp[L], rn[L];
    if (!n) return 0;
    for (uint32_t k = 0; k < n; ++k) {
        memset(pos, 0, sizeof(pos));
        memset(neg, 0, sizeof(neg));
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t* ai = a + (size_t)i * L;
            if (i <= k) mul_add(pos, ACC, ai, b + (size_t)(k - i) * L);
            else mul_add(neg, ACC, ai, b + (size_t)(k + n - i) * L);
        }
        mod_q(pos, ACC, q, rp);
        mod_q(neg, ACC, q, rn);
        sub_mod_q(rp, rn, q, c + (size_t)k * L);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* small-modulus arithmetic (rns/solve.cu:109-125)                          */
/* ------------------------------------------------------------------------ */
uint32_t orc_pow_mod(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1 % p, b = a % p;
    while (e) {
        if (e & 1) r = r * b % p;
        e >>= 1;
        if (e) b = b * b % p;
    }
    return (uint32_t)r;
}

static uint32_t mulmod(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)((uint64_t)a * b % p); }
static uint32_t addmod(uint32_t a, uint32_t b, uint32_t p) { uint32_t s = a + b; return s >= p ? s - p : s; }
static uint32_t submod(uint32_t a, uint32_t b, uint32_t p) { return a >= b ? a - b : a + (p - b); }

typedef struct { uint32_t value, shoup; } tw_t;
static tw_t twiddle(uint32_t value, uint32_t p) {
    tw_t t = {value, (uint32_t)(((uint64_t)value << 32) / p)};
    return t;
}
/* Shoup: exact residual of a*w below 2p for any a < 2^32, one subtraction. */
static uint32_t shoup(uint32_t a, tw_t w, uint32_t p) {
    uint32_t quotient = (uint32_t)(((uint64_t)a * w.shoup) >> 32);
    uint32_t r = a * w.value - quotient * p;
    return r >= p ? r - p : r;
}
/* Barrett with floor((2^64-1)/p), one subtraction. */
static uint32_t barrett(uint32_t a, uint32_t b, uint32_t p, uint64_t reciprocal) {
    uint64_t product = (uint64_t)a * b;
    uint64_t quotient = (uint64_t)(((unsigned __int128)product * reciprocal) >> 64);
    uint32_t r = (uint32_t)(product - quotient * p);
    return r >= p ? r - p : r;
}

static int is_prime32(uint32_t p) {
    if (p < 2) return 0;
    if (!(p & 1)) return p == 2;
    for (uint32_t d = 3; (uint64_t)d * d <= p; d += 2)
        if (p % d == 0) return 0;
    return 1;
}

int orc_primes_reference(uint32_t k, uint32_t* primes) {
    uint32_t candidate = (1u << 31) - 65535u, found = 0;
    while (found < k) {
        if (candidate <= 0xFFFFFFFFu / 3) return -1;
        if (is_prime32(candidate)) primes[found++] = candidate;
        candidate -= 65536u;
    }
    return 0;
}

uint32_t orc_product_bits(uint32_t k, const uint32_t* primes) {
    uint32_t prod[ORC_MAX_PRIMES + 2];
    memset(prod, 0, sizeof(prod));
    prod[0] = 1;
    for (uint32_t i = 0; i < k; ++i) {
        uint32_t tmp[ORC_MAX_PRIMES + 2];
        memset(tmp, 0, sizeof(tmp));
        mul_small_add(tmp, ORC_MAX_PRIMES + 2, prod, ORC_MAX_PRIMES + 1, primes[i]);
        memcpy(prod, tmp, sizeof(prod));
    }
    return bits_of(prod, ORC_MAX_PRIMES + 2);
}

int orc_crt_bound_ok(uint32_t n, const uint32_t q[L], uint32_t k, const uint32_t* primes) {
    enum { W = ORC_MAX_PRIMES + 4 };
    uint32_t prod[W], qm1[L], bound[W], tmp[W];
    memset(prod, 0, sizeof(prod));
    prod[0] = 1;
    for (uint32_t i = 0; i < k; ++i) {
        memset(tmp, 0, sizeof(tmp));
        mul_small_add(tmp, W, prod, W, primes[i]);
        memcpy(prod, tmp, sizeof(prod));
    }
    memcpy(qm1, q, sizeof(qm1));
    qm1[0] -= 1; /* q odd */
    memset(bound, 0, sizeof(bound));
    mul_add(bound, W, qm1, qm1);
    memset(tmp, 0, sizeof(tmp));
    mul_small_add(tmp, W, bound, W, 4u * n);
    return cmp_words(prod, tmp, W) > 0;
}

uint32_t orc_primes_general(uint32_t max_n, const uint32_t q[L], uint32_t* primes) {
    uint32_t k = 0;
    for (uint32_t c = (1u << 12) - 1; c > 0 && k < ORC_MAX_PRIMES; --c) {
        uint32_t p = (c << 18) + 1;
        if (!is_prime32(p)) continue;
        primes[k++] = p;
        if (orc_crt_bound_ok(max_n, q, k, primes)) return k;
    }
    return 0;
}

uint32_t orc_find_psi(uint32_t p, uint32_t n) {
    for (uint32_t g = 2; g < 10000; ++g) {
        uint32_t root = orc_pow_mod(g, (p - 1) / (2 * n), p);
        if (orc_pow_mod(root, n, p) == p - 1) return root;
    }
    return 0;
}

static uint32_t brv(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

/* ------------------------------------------------------------------------ */
/* stage functions                                                          */
/* ------------------------------------------------------------------------ */
/* one prime's row: x_j mod p for every coefficient (rns/solve.cu:138-157) */
static void residues_row(uint32_t n, const uint32_t* x, uint32_t p, uint32_t* row) {
    tw_t powers[L];
    uint32_t base = (uint32_t)((1ull << 32) % p), power = 1;
    for (int limb = 0; limb < L; ++limb) {
        powers[limb] = twiddle(power, p);
        power = mulmod(power, base, p);
    }
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t* w = x + (size_t)j * L;
        uint32_t even = 0, odd = 0;
        for (int limb = 0; limb < L; ++limb) {
            uint32_t term = shoup(w[limb], powers[limb], p);
            if (limb & 1) odd = addmod(odd, term, p);
            else even = addmod(even, term, p);
        }
        row[j] = addmod(even, odd, p);
    }
}

void orc_residues(uint32_t n, const uint32_t* x, uint32_t k, const uint32_t* primes,
                  uint32_t* out) {
    for (uint32_t i = 0; i < k; ++i) residues_row(n, x, primes[i], out + (size_t)i * n);
}

/* tw[i] = psi^brv(i) for i in [0,n) */
static void bitrev_powers(uint32_t p, uint32_t psi, uint32_t logn, tw_t* tw) {
    const uint32_t n = 1u << logn;
    uint32_t* pw = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t v = 1;
    for (uint32_t e = 0; e < n; ++e) { pw[e] = v; v = mulmod(v, psi, p); }
    for (uint32_t i = 0; i < n; ++i) tw[i] = twiddle(pw[brv(i, logn)], p);
    free(pw);
}

void orc_ntt_forward(uint32_t p, uint32_t psi, uint32_t logn, uint32_t* x) {
    const uint32_t n = 1u << logn;
    tw_t* tw = (tw_t*)malloc(sizeof(tw_t) * n);
    bitrev_powers(p, psi, logn, tw);
    for (uint32_t h = n / 2; h >= 1; h /= 2) {
        const uint32_t m = n / (2 * h);
        for (uint32_t g = 0; g < m; ++g) {
            const tw_t w = tw[m + g];
            for (uint32_t i = 2 * h * g; i < 2 * h * g + h; ++i) {
                uint32_t u = x[i], v = shoup(x[i + h], w, p);
                x[i] = addmod(u, v, p);
                x[i + h] = submod(u, v, p);
            }
        }
        if (h == 1) break;
    }
    free(tw);
}

void orc_ntt_inverse(uint32_t p, uint32_t psi, uint32_t logn, uint32_t* x) {
    const uint32_t n = 1u << logn;
    tw_t* tw = (tw_t*)malloc(sizeof(tw_t) * n);
    bitrev_powers(p, orc_pow_mod(psi, p - 2, p), logn, tw);
    for (uint32_t h = 1; h < n; h *= 2) {
        const uint32_t m = n / (2 * h);
        for (uint32_t g = 0; g < m; ++g) {
            const tw_t w = tw[m + g];
            for (uint32_t i = 2 * h * g; i < 2 * h * g + h; ++i) {
                uint32_t u = x[i], v = x[i + h];
                x[i] = addmod(u, v, p);
                x[i + h] = shoup(submod(u, v, p), w, p);
            }
        }
    }
    free(tw);
}

/* ------------------------------------------------------------------------ */
/* CRT (rns/solve.cu:336-420, general-q reduction)                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t k;
    uint32_t primes[ORC_MAX_PRIMES];
    uint32_t basis_inverse[ORC_MAX_PRIMES];            /* (M_i mod p_i)^-1 */
    uint32_t bases_mod_q[ORC_MAX_PRIMES][L];           /* M_i mod q */
    uint64_t reciprocal[ORC_MAX_PRIMES];               /* floor((2^64-1)/p) */
    uint32_t product_mod_q[L];                         /* P mod q */
} crt_t;

static void crt_setup(const uint32_t q[L], uint32_t k, const uint32_t* primes, crt_t* s) {
    enum { W = ORC_MAX_PRIMES + 4 };
    uint32_t prod[W], tmp[W];
    s->k = k;
    memset(prod, 0, sizeof(prod));
    prod[0] = 1;
    for (uint32_t i = 0; i < k; ++i) {
        memset(tmp, 0, sizeof(tmp));
        mul_small_add(tmp, W, prod, W, primes[i]);
        memcpy(prod, tmp, sizeof(prod));
        s->primes[i] = primes[i];
        s->reciprocal[i] = UINT64_MAX / primes[i];
    }
    mod_q(prod, W, q, s->product_mod_q);
    for (uint32_t i = 0; i < k; ++i) {
        /* M_i = P / p_i by long division */
        uint32_t basis[W];
        uint64_t rem = 0;
        for (int w = W - 1; w >= 0; --w) {
            uint64_t v = (rem << 32) | prod[w];
            basis[w] = (uint32_t)(v / primes[i]);
            rem = v % primes[i];
        }
        mod_q(basis, W, q, s->bases_mod_q[i]);
        uint64_t r = 0;
        for (int w = W - 1; w >= 0; --w) r = ((r << 32) | basis[w]) % primes[i];
        s->basis_inverse[i] = orc_pow_mod((uint32_t)r, primes[i] - 2, primes[i]);
    }
}

int orc_crt_constants(const uint32_t q[L], uint32_t k, const uint32_t* primes,
                      uint32_t* bases_mod_q, uint32_t* basis_inverse, uint32_t* product_mod_q) {
    crt_t* s = (crt_t*)malloc(sizeof(crt_t));
    if (!s || k > ORC_MAX_PRIMES) { free(s); return -1; }
    crt_setup(q, k, primes, s);
    for (uint32_t i = 0; i < k; ++i) {
        memcpy(bases_mod_q + (size_t)i * L, s->bases_mod_q[i], L * sizeof(uint32_t));
        basis_inverse[i] = s->basis_inverse[i];
    }
    memcpy(product_mod_q, s->product_mod_q, L * sizeof(uint32_t));
    free(s);
    return 0;
}

/* t[i] = y mod p_i times (M_i mod p_i)^-1, canonical; out = y mod q */
static void crt_one(const crt_t* s, const uint32_t* t, const uint32_t* q, uint32_t* out) {
    enum { A = L + 4 };
    uint32_t acc[A], answer[L], corr_full[A], correction[L];
    uint64_t fraction = 0;
    uint32_t alpha = 0;
    memset(acc, 0, sizeof(acc));
    for (uint32_t i = 0; i < s->k; ++i) {
        uint64_t term = (uint64_t)t[i] * s->reciprocal[i];
        uint64_t next = fraction + term;
        alpha += next < fraction;
        fraction = next;
        mul_small_add(acc, A, s->bases_mod_q[i], L, t[i]);
    }
    /* |y| < P/4 and the fixed-point error is below 1/4: rounding is exact */
    uint32_t nearest = alpha + (uint32_t)(fraction >> 63);
    mod_q(acc, A, q, answer);
    memset(corr_full, 0, sizeof(corr_full));
    mul_small_add(corr_full, A, s->product_mod_q, L, nearest);
    mod_q(corr_full, A, q, correction);
    sub_mod_q(answer, correction, q, out);
}

int orc_crt(uint32_t n, const uint32_t q[L], uint32_t k, const uint32_t* primes,
            const uint32_t* residues, uint32_t* c) {
    crt_t* s = (crt_t*)malloc(sizeof(crt_t));
    uint32_t t[ORC_MAX_PRIMES];
    if (!s || k > ORC_MAX_PRIMES) { free(s); return -1; }
    crt_setup(q, k, primes, s);
    for (uint32_t j = 0; j < n; ++j) {
        for (uint32_t i = 0; i < k; ++i)
            t[i] = mulmod(residues[(size_t)i * n + j], s->basis_inverse[i], primes[i]);
        crt_one(s, t, q, c + (size_t)j * L);
    }
    free(s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* the reference RNS product                                                */
/* ------------------------------------------------------------------------ */
/* Cyclic radix-2 DIT from bit-reversed input to natural output, twiddles
 * psi^(j*n/half) (rns/solve.cu:248-262 with tables[j*stride]). */
static void cyclic_dit(uint32_t* x, uint32_t n, uint32_t p, const tw_t* powers /* psi^e, e<n */) {
    for (uint32_t half = 1; half < n; half *= 2) {
        const uint32_t step = n / half;
        for (uint32_t start = 0; start < n; start += 2 * half)
            for (uint32_t j = 0; j < half; ++j) {
                uint32_t u = x[start + j];
                uint32_t v = shoup(x[start + j + half], powers[j * step], p);
                x[start + j] = addmod(u, v, p);
                x[start + j + half] = submod(u, v, p);
            }
    }
}

/* The product as two kinds of independent units: per prime, conversion of
 * a and b, twist, forward NTTs, pointwise product, inverse, scale (row i of
 * ra); then per block of coefficients, the CRT.  One thread runs them in
 * order; orc_rns_multiply_threads spreads each kind over threads. */
typedef struct {
    uint32_t n, logn, k, threads, index;
    const uint32_t* q;
    const uint32_t* primes;
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* c;
    const crt_t* s;
    uint32_t* ra;
    uint32_t* rb;
    int failed;
} rns_job_t;

static int rns_prime_unit(const rns_job_t* job, uint32_t i, uint32_t* tmp, tw_t* fwd, tw_t* inv, tw_t* scale) {
    const uint32_t n = job->n, logn = job->logn, p = job->primes[i];
    const uint32_t psi = orc_find_psi(p, n);
    const uint32_t ipsi = orc_pow_mod(psi, p - 2, p);
    const uint32_t inv_n = orc_pow_mod(n, p - 2, p);
    const uint64_t recip = UINT64_MAX / p;
    uint32_t f = 1, bk = 1, sc = mulmod(inv_n, job->s->basis_inverse[i], p);
    uint32_t* xa = job->ra + (size_t)i * n;
    uint32_t* xb = job->rb + (size_t)i * n;
    residues_row(n, job->a, p, xa);
    residues_row(n, job->b, p, xb);
    for (uint32_t e = 0; e < n; ++e) {
        fwd[e] = twiddle(f, p);
        inv[e] = twiddle(bk, p);
        scale[e] = twiddle(sc, p);
        f = mulmod(f, psi, p);
        bk = mulmod(bk, ipsi, p);
        sc = mulmod(sc, ipsi, p);
    }
    /* twist by psi^j and store bit-reversed (rns/solve.cu:156) */
    for (uint32_t j = 0; j < n; ++j) tmp[brv(j, logn)] = shoup(xa[j], fwd[j], p);
    memcpy(xa, tmp, sizeof(uint32_t) * n);
    for (uint32_t j = 0; j < n; ++j) tmp[brv(j, logn)] = shoup(xb[j], fwd[j], p);
    memcpy(xb, tmp, sizeof(uint32_t) * n);
    cyclic_dit(xa, n, p, fwd);
    cyclic_dit(xb, n, p, fwd);
    /* pointwise Barrett product scattered bit-reversed (rns/solve.cu:327-333) */
    for (uint32_t j = 0; j < n; ++j) tmp[brv(j, logn)] = barrett(xa[j], xb[j], p, recip);
    cyclic_dit(tmp, n, p, inv);
    /* scale = N^-1 (M_i mod p_i)^-1 psi^-j (rns/host_setup.h:111-118) */
    for (uint32_t j = 0; j < n; ++j) xa[j] = shoup(tmp[j], scale[j], p);
    return 0;
}

/* primes index, index + threads, ... */
static void* rns_prime_units(void* arg) {
    rns_job_t* job = (rns_job_t*)arg;
    const uint32_t n = job->n;
    uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * n);
    tw_t* tw = (tw_t*)malloc(sizeof(tw_t) * 3 * (size_t)n);
    if (!tmp || !tw) {
        job->failed = 1;
    } else {
        for (uint32_t i = job->index; i < job->k; i += job->threads)
            rns_prime_unit(job, i, tmp, tw, tw + n, tw + 2 * (size_t)n);
    }
    free(tmp);
    free(tw);
    return NULL;
}

/* block `index` of `threads` equal coefficient ranges */
static void* rns_crt_units(void* arg) {
    rns_job_t* job = (rns_job_t*)arg;
    const uint32_t n = job->n, per = (n + job->threads - 1) / job->threads;
    const uint32_t j0 = job->index * per, j1 = j0 + per < n ? j0 + per : n;
    uint32_t t[ORC_MAX_PRIMES];
    for (uint32_t j = j0; j < j1; ++j) {
        for (uint32_t i = 0; i < job->k; ++i) t[i] = job->ra[(size_t)i * n + j];
        crt_one(job->s, t, job->q, job->c + (size_t)j * L);
    }
    return NULL;
}

/* runs fn over `threads` copies of base (index 0..threads-1): thread 0 on
 * the caller, the rest on their own threads; 0 if every unit ran */
static int run_units(void* (*fn)(void*), const rns_job_t* base, uint32_t threads) {
    rns_job_t jobs[ORC_MAX_THREADS];
    pthread_t ids[ORC_MAX_THREADS];
    int started[ORC_MAX_THREADS] = {0};
    int failed = 0;
    for (uint32_t t = 0; t < threads; ++t) {
        jobs[t] = *base;
        jobs[t].index = t;
        jobs[t].failed = 0;
    }
    for (uint32_t t = 1; t < threads; ++t) started[t] = pthread_create(&ids[t], NULL, fn, &jobs[t]) == 0;
    fn(&jobs[0]);
    for (uint32_t t = 1; t < threads; ++t) {
        if (started[t]) pthread_join(ids[t], NULL);
        else fn(&jobs[t]); /* no thread: run the unit here */
    }
    for (uint32_t t = 0; t < threads; ++t) failed |= jobs[t].failed;
    return failed ? -3 : 0;
}

int orc_rns_multiply_threads(uint32_t n, const uint32_t q[L], uint32_t k, const uint32_t* primes,
                             const uint32_t* a, const uint32_t* b, uint32_t* c, uint32_t threads) {
    uint32_t logn = 0;
    while ((1u << logn) < n) ++logn;
    if (!n || (1u << logn) != n || k > ORC_MAX_PRIMES || threads < 1 || threads > ORC_MAX_THREADS) return -1;
    if (!orc_crt_bound_ok(n, q, k, primes)) return -2;
    crt_t* s = (crt_t*)malloc(sizeof(crt_t));
    uint32_t* ra = (uint32_t*)malloc(sizeof(uint32_t) * n * k);
    uint32_t* rb = (uint32_t*)malloc(sizeof(uint32_t) * n * k);
    int rc = -3;
    if (s && ra && rb) {
        crt_setup(q, k, primes, s);
        rns_job_t job = {n, logn, k, 0, 0, q, primes, a, b, c, s, ra, rb, 0};
        const uint32_t prime_threads = threads < k ? threads : k;
        const uint32_t crt_threads = threads < n ? threads : n;
        job.threads = prime_threads;
        rc = run_units(rns_prime_units, &job, prime_threads);
        if (rc == 0) {
            job.threads = crt_threads;
            rc = run_units(rns_crt_units, &job, crt_threads);
        }
    }
    free(s); free(ra); free(rb);
    return rc;
}

int orc_rns_multiply(uint32_t n, const uint32_t q[L], uint32_t k, const uint32_t* primes,
                     const uint32_t* a, const uint32_t* b, uint32_t* c) {
    return orc_rns_multiply_threads(n, q, k, primes, a, b, c, 1);
}

/* ------------------------------------------------------------------------ */
/* the special-modulus quartic product (quartic/host_setup.h:41-110,         */
/* quartic/solve.cu:289-316 and 906-1040, research/QUARTIC_RNS_RU.md:3-47)   */
/* ------------------------------------------------------------------------ */
enum { QW = 20 }; /* words of the quartic CRT integers (P < 2^527) */

int orc_quartic_delta(const uint32_t q[L], uint32_t* delta) {
    /* q = 2^868 - delta, 0 < delta < 2^28: words 1..26 all ones, word 27 = 0xF */
    for (int i = 1; i < L - 1; ++i)
        if (q[i] != 0xFFFFFFFFu) return -1;
    if (q[L - 1] != 0xFu || q[0] == 0) return -1;
    const uint32_t d = 0u - q[0];
    if (d == 0 || d >= (1u << 28)) return -1;
    *delta = d;
    return 0;
}

/* bound = n (2^217 - 1)^2 (1 + 3 delta) (quartic/host_setup.h:48-51), as words */
static void quartic_bound(uint32_t n, uint32_t delta, uint32_t* bound /* QW */) {
    uint32_t part[7], sq[QW];
    memset(sq, 0, sizeof(sq));
    for (int i = 0; i < 7; ++i) part[i] = 0xFFFFFFFFu;
    part[6] = (1u << 25) - 1;
    for (int i = 0; i < 7; ++i) mul_small_add(sq + i, QW - i, part, 7, part[i]);
    memset(bound, 0, QW * sizeof(uint32_t));
    uint32_t t[QW];
    memset(t, 0, sizeof(t));
    mul_small_add(t, QW, sq, QW, n);
    mul_small_add(bound, QW, t, QW, 1u + 3u * delta);
}

/* the reference's quartic prime rule (quartic/host_setup.h:52-66): 16 primes
 * 2^bits - 65535 - 65536 j with delta a fourth power mod p, first with bits =
 * 30, then 31, until P > 4 bound */
uint32_t orc_quartic_primes_reference(uint32_t n, uint32_t delta, uint32_t* primes, uint32_t* product_bits,
                                      uint32_t* bound_bits) {
    uint32_t bound[QW], four[QW];
    quartic_bound(n, delta, bound);
    memset(four, 0, sizeof(four));
    mul_small_add(four, QW, bound, QW, 4);
    if (bound_bits) *bound_bits = bits_of(bound, QW);
    for (uint32_t bits = 30; bits <= 31; ++bits) {
        uint32_t k = 0, product[QW];
        memset(product, 0, sizeof(product));
        product[0] = 1;
        for (uint32_t cand = (1u << bits) - 65535u; k < 16 && cand > (1u << (bits - 1)); cand -= 65536u) {
            if (!is_prime32(cand) || orc_pow_mod(delta, (cand - 1) / 4, cand) != 1) continue;
            primes[k++] = cand;
            uint32_t t[QW];
            memset(t, 0, sizeof(t));
            mul_small_add(t, QW, product, QW, cand);
            memcpy(product, t, sizeof(t));
        }
        if (k == 16 && cmp_words(product, four, QW) > 0) {
            if (product_bits) *product_bits = bits_of(product, QW);
            return 16;
        }
    }
    return 0;
}

/* square root of a quadratic residue mod p (Tonelli-Shanks, as
 * quartic/host_setup.h:20-38 computes it); 0 if none */
static uint32_t sqrt_mod_p(uint32_t a, uint32_t p) {
    a %= p;
    if (!a || orc_pow_mod(a, (p - 1) / 2, p) != 1) return 0;
    uint32_t odd = p - 1, order = 0;
    while (!(odd & 1)) { odd >>= 1; ++order; }
    uint32_t z = 2;
    while (orc_pow_mod(z, (p - 1) / 2, p) != p - 1) ++z;
    uint32_t c = orc_pow_mod(z, odd, p), x = orc_pow_mod(a, (odd + 1) / 2, p), t = orc_pow_mod(a, odd, p);
    while (t != 1) {
        uint32_t i = 0, v = t;
        while (v != 1 && i < order) { v = mulmod(v, v, p); ++i; }
        if (i >= order) return 0;
        const uint32_t b = orc_pow_mod(c, 1ull << (order - i - 1), p);
        x = mulmod(x, b, p);
        c = mulmod(b, b, p);
        t = mulmod(t, c, p);
        order = i;
    }
    return x;
}

/* negacyclic product mod p of xa and xb (length n) into xa, unnormalised by
 * N (the twisted cyclic transform of orc_rns_multiply) */
static void negacyclic_mod_p(uint32_t n, uint32_t logn, uint32_t p, uint32_t psi, uint32_t* xa, uint32_t* xb,
                             uint32_t* tmp, tw_t* fwd, tw_t* inv) {
    const uint32_t ipsi = orc_pow_mod(psi, p - 2, p);
    const uint64_t recip = UINT64_MAX / p;
    uint32_t f = 1, bk = 1;
    for (uint32_t e = 0; e < n; ++e) {
        fwd[e] = twiddle(f, p);
        inv[e] = twiddle(bk, p);
        f = mulmod(f, psi, p);
        bk = mulmod(bk, ipsi, p);
    }
    for (uint32_t j = 0; j < n; ++j) tmp[brv(j, logn)] = shoup(xa[j], fwd[j], p) % p;
    memcpy(xa, tmp, sizeof(uint32_t) * n);
    for (uint32_t j = 0; j < n; ++j) tmp[brv(j, logn)] = shoup(xb[j], fwd[j], p) % p;
    memcpy(xb, tmp, sizeof(uint32_t) * n);
    cyclic_dit(xa, n, p, fwd);
    cyclic_dit(xb, n, p, fwd);
    for (uint32_t j = 0; j < n; ++j) tmp[brv(j, logn)] = barrett(xa[j], xb[j], p, recip);
    cyclic_dit(tmp, n, p, inv);
    for (uint32_t j = 0; j < n; ++j) xa[j] = mulmod(tmp[j], inv[j].value, p); /* untwist by psi^-j */
}

/* bits [217 j, 217 j + 217) of a 28-word coefficient as 7 words
 * (quartic/solve.cu:293-302) */
static void quartic_component(const uint32_t* w, uint32_t j, uint32_t part[7]) {
    const uint32_t shift = 217u * j, base = shift / 32, s = shift % 32;
    for (uint32_t l = 0; l < 7; ++l) {
        uint32_t v = base + l < L ? w[base + l] >> s : 0;
        if (s && base + l + 1 < L) v |= w[base + l + 1] << (32 - s);
        part[l] = v;
    }
    part[6] &= (1u << 25) - 1;
}

int orc_quartic_multiply(uint32_t n, const uint32_t q[L], uint32_t k, const uint32_t* primes, const uint32_t* a,
                         const uint32_t* b, uint32_t* c) {
    uint32_t logn = 0, delta = 0;
    while ((1u << logn) < n) ++logn;
    if (n < 4 || (1u << logn) != n || k == 0 || k > 32 || orc_quartic_delta(q, &delta)) return -1;
    /* P > 4 bound */
    uint32_t bound[QW], four[QW], P[QW];
    quartic_bound(n, delta, bound);
    memset(four, 0, sizeof(four));
    mul_small_add(four, QW, bound, QW, 4);
    memset(P, 0, sizeof(P));
    P[0] = 1;
    for (uint32_t i = 0; i < k; ++i) {
        uint32_t t[QW];
        memset(t, 0, sizeof(t));
        mul_small_add(t, QW, P, QW, primes[i]);
        memcpy(P, t, sizeof(t));
    }
    if (cmp_words(P, four, QW) <= 0) return -2;
    const size_t rows = (size_t)4 * k;
    uint32_t* ra = (uint32_t*)malloc(sizeof(uint32_t) * n * rows);
    uint32_t* rb = (uint32_t*)malloc(sizeof(uint32_t) * n * rows);
    uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * n);
    tw_t* fwd = (tw_t*)malloc(sizeof(tw_t) * n);
    tw_t* inv = (tw_t*)malloc(sizeof(tw_t) * n);
    uint32_t* comp = (uint32_t*)malloc(sizeof(uint32_t) * n * rows); /* R_j mod p_i, [i][j][n] */
    if (!ra || !rb || !tmp || !fwd || !inv || !comp) {
        free(ra); free(rb); free(tmp); free(fwd); free(inv); free(comp);
        return -3;
    }
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t p = primes[i];
        const uint32_t psi = orc_find_psi(p, n);
        const uint32_t r = sqrt_mod_p(sqrt_mod_p(delta, p), p);
        const uint32_t io = orc_pow_mod(psi, n / 2, p); /* i^2 = -1 */
        if (!psi || !r || orc_pow_mod(r, 4, p) != delta % p || mulmod(io, io, p) != p - 1) {
            free(ra); free(rb); free(tmp); free(fwd); free(inv); free(comp);
            return -4;
        }
        /* channel k of prime i: A(r i^k) = sum_j a_j r^j i^(jk) (quartic/solve.cu:303-316) */
        for (int poly = 0; poly < 2; ++poly) {
            const uint32_t* x = poly ? b : a;
            uint32_t* out = poly ? rb : ra;
            for (uint32_t m = 0; m < n; ++m) {
                uint32_t z[4];
                for (uint32_t j = 0; j < 4; ++j) {
                    uint32_t part[7];
                    quartic_component(x + (size_t)m * L, j, part);
                    uint32_t acc = 0, pw = 1, base = (uint32_t)((1ull << 32) % p);
                    for (uint32_t l = 0; l < 7; ++l) {
                        acc = addmod(acc, mulmod(part[l] % p, pw, p), p);
                        pw = mulmod(pw, base, p);
                    }
                    z[j] = mulmod(acc, orc_pow_mod(r, j, p), p);
                }
                const uint32_t u = addmod(z[0], z[2], p), v = submod(z[0], z[2], p);
                const uint32_t s = addmod(z[1], z[3], p), t = mulmod(submod(z[1], z[3], p), io, p);
                out[(4 * i + 0) * (size_t)n + m] = addmod(u, s, p);
                out[(4 * i + 1) * (size_t)n + m] = addmod(v, t, p);
                out[(4 * i + 2) * (size_t)n + m] = submod(u, s, p);
                out[(4 * i + 3) * (size_t)n + m] = submod(v, t, p);
            }
        }
        for (uint32_t ch = 0; ch < 4; ++ch)
            negacyclic_mod_p(n, logn, p, psi, ra + (4 * i + ch) * (size_t)n, rb + (4 * i + ch) * (size_t)n, tmp, fwd,
                             inv);
        /* inverse mixing and (4 N)^-1 r^-j (quartic/solve.cu:920-935) */
        const uint32_t inv4n = orc_pow_mod((uint32_t)(((uint64_t)4 * n) % p), p - 2, p);
        const uint32_t rinv = orc_pow_mod(r, p - 2, p);
        for (uint32_t m = 0; m < n; ++m) {
            const uint32_t f0 = ra[(4 * i + 0) * (size_t)n + m], f1 = ra[(4 * i + 1) * (size_t)n + m];
            const uint32_t f2 = ra[(4 * i + 2) * (size_t)n + m], f3 = ra[(4 * i + 3) * (size_t)n + m];
            const uint32_t u = addmod(f0, f2, p), v = addmod(f1, f3, p);
            const uint32_t d0 = submod(f0, f2, p), d1 = mulmod(submod(f1, f3, p), io, p);
            const uint32_t mixed[4] = {addmod(u, v, p), submod(d0, d1, p), submod(u, v, p), addmod(d0, d1, p)};
            for (uint32_t j = 0; j < 4; ++j)
                comp[(4 * i + j) * (size_t)n + m] = mulmod(mulmod(mixed[j], inv4n, p), orc_pow_mod(rinv, j, p), p);
        }
    }
    /* signed CRT per component, then sum_j R_j 2^(217 j) mod q
     * (quartic/solve.cu:936-1040: magnitude, shift, fold, sign, add) */
    uint32_t M[32][QW], Minv[32], half[QW];
    for (uint32_t i = 0; i < k; ++i) {
        uint64_t rem = 0;
        for (int w = QW - 1; w >= 0; --w) {
            const uint64_t v = (rem << 32) | P[w];
            M[i][w] = (uint32_t)(v / primes[i]);
            rem = v % primes[i];
        }
        uint64_t mm = 0;
        for (int w = QW - 1; w >= 0; --w) mm = ((mm << 32) | M[i][w]) % primes[i];
        Minv[i] = orc_pow_mod((uint32_t)mm, primes[i] - 2, primes[i]);
    }
    for (int w = 0; w < QW; ++w) half[w] = (P[w] >> 1) | (w + 1 < QW ? P[w + 1] << 31 : 0);
    for (uint32_t m = 0; m < n; ++m) {
        uint32_t total[L];
        memset(total, 0, sizeof(total));
        for (uint32_t j = 0; j < 4; ++j) {
            uint32_t x[QW + 1];
            memset(x, 0, sizeof(x));
            for (uint32_t i = 0; i < k; ++i) {
                const uint32_t t = mulmod(comp[(4 * i + j) * (size_t)n + m], Minv[i], primes[i]);
                mul_small_add(x, QW + 1, M[i], QW, t);
            }
            /* x mod P (x < k P), then the signed value: negative above P/2 */
            while (x[QW] || cmp_words(x, P, QW) >= 0) {
                uint64_t borrow = 0;
                for (int w = 0; w <= QW; ++w) {
                    const uint64_t d = (uint64_t)x[w] - (w < QW ? P[w] : 0) - borrow;
                    x[w] = (uint32_t)d;
                    borrow = (d >> 63) & 1;
                }
            }
            int negative = cmp_words(x, half, QW) > 0;
            if (negative) { /* |R| = P - x */
                uint64_t borrow = 0;
                for (int w = 0; w < QW; ++w) {
                    const uint64_t d = (uint64_t)P[w] - x[w] - borrow;
                    x[w] = (uint32_t)d;
                    borrow = (d >> 63) & 1;
                }
            }
            /* |R_j| 2^(217 j) mod q */
            uint32_t shifted[QW + 24], red[L];
            memset(shifted, 0, sizeof(shifted));
            const uint32_t sh = 217u * j;
            for (uint32_t bit = 0; bit < 32u * QW; ++bit)
                if (bit_at(x, bit)) shifted[(bit + sh) / 32] |= 1u << ((bit + sh) % 32);
            mod_q(shifted, QW + 24, q, red);
            uint32_t zero[L];
            memset(zero, 0, sizeof(zero));
            if (negative) sub_mod_q(zero, red, q, red);
            /* total = (total + red) mod q */
            uint32_t nred[L];
            sub_mod_q(zero, red, q, nred);
            sub_mod_q(total, nred, q, total);
        }
        memcpy(c + (size_t)m * L, total, sizeof(total));
    }
    free(ra); free(rb); free(tmp); free(fwd); free(inv); free(comp);
    return 0;
}
