/*
 * oracle.h -- CPU restatement of the reference's exact negacyclic product
 * c = a*b mod (X^N+1, q) for an 868-bit q, used ONLY as the checker:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product path never links, calls or falls back to anything here.
 *
 * Coefficients are 28 little-endian uint32 limbs (868 bits used), the wire
 * layout of fherma.h:31-40 (tensor<N x L x u32>, row-major).
 *
 * Parity pinning (what fixes each function independently of its own code):
 *   orc_schoolbook        -- brute-force definition; checked against GMP
 *                            Kronecker fixtures (tests/golden) and against
 *                            the reference's published validator families.
 *   orc_rns_multiply      -- equals orc_schoolbook on small N, equals the
 *                            GMP Kronecker SHA-256 fixtures at N up to 2^17.
 *   orc_rns_multiply_threads -- equals orc_rns_multiply word for word, and
 *                            orc_schoolbook and a GMP fixture directly.
 *   orc_primes_reference  -- reproduces results/rns-arithmetic.json and the
 *                            prime list of rns/host_setup.h (oracle/_ref).
 *   orc_ntt_forward/inverse -- direct evaluation at odd powers of psi for
 *                            small N, inverse(forward(x)) = N*x, convolution
 *                            theorem against schoolbook mod p.
 *   orc_residues          -- Python big-int `a % p` in the tests.
 */
#ifndef POLYMUL_ORACLE_H
#define POLYMUL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LIMBS 28       /* uint32 limbs per ABI coefficient (fherma.h:35) */
#define ORC_WIDTH 868      /* bits of q (W in the manifest) */
#define ORC_MAX_PRIMES 64

/* ---- seeded synthetic inputs (SplitMix64 streams) ------------------------ */
uint64_t orc_mix64(uint64_t x);
uint64_t orc_value64(uint64_t seed, uint64_t tag, uint64_t index);
/* q: 868-bit, bit 867 and bit 0 set, other bits from the stream. */
void orc_gen_modulus(uint64_t seed, uint32_t q[ORC_LIMBS]);
/* q = 2^868 - delta (the reference's special modulus form). */
void orc_special_modulus(uint32_t delta, uint32_t q[ORC_LIMBS]);
/* n coefficients uniform in [0,q) by rejection; poly selects the stream. */
void orc_gen_poly(uint64_t seed, uint32_t poly, uint32_t n,
                  const uint32_t q[ORC_LIMBS], uint32_t* out);

/* ---- exact big-integer definition --------------------------------------- */
/* c_k = sum_{i+j=k} a_i b_j - sum_{i+j=k+N} a_i b_j  (mod q). O(N^2). */
int orc_schoolbook(uint32_t n, const uint32_t q[ORC_LIMBS],
                   const uint32_t* a, const uint32_t* b, uint32_t* c);

/* ---- RNS parameters ------------------------------------------------------ */
/* rns/host_setup.h:64-75: p = 2^31 - 65535 - 65536*j, prime, k of them. */
int orc_primes_reference(uint32_t k, uint32_t* primes);
/* Primes p = c*2^18 + 1 < 2^30, largest first, as few as make
 * prod(p) > 4*max_n*(q-1)^2.  Returns k (0 on failure). */
uint32_t orc_primes_general(uint32_t max_n, const uint32_t q[ORC_LIMBS],
                            uint32_t* primes);
/* rns/host_setup.h:103-107: first g>=2 with g^((p-1)/2n) of order 2n. */
uint32_t orc_find_psi(uint32_t p, uint32_t n);
/* 1 if prod(primes) > 4*n*(q-1)^2 (rns/host_setup.h:76-78). */
int orc_crt_bound_ok(uint32_t n, const uint32_t q[ORC_LIMBS],
                     uint32_t k, const uint32_t* primes);
/* bits of prod(primes) */
uint32_t orc_product_bits(uint32_t k, const uint32_t* primes);

uint32_t orc_pow_mod(uint32_t a, uint64_t e, uint32_t p);

/* ---- stage functions (one prime, N = 2^logn) ----------------------------- */
/* out[i*n + j] = (coefficient j) mod primes[i] (closed form). */
void orc_residues(uint32_t n, const uint32_t* x, uint32_t k,
                  const uint32_t* primes, uint32_t* out);
/* Odd-root Cooley-Tukey negacyclic NTT: natural order in, bit-reversed out;
 * out[r] = x(psi^(2*brv(r)+1)) mod p.  Values canonical in [0,p). */
void orc_ntt_forward(uint32_t p, uint32_t psi, uint32_t logn, uint32_t* x);
/* Gentleman-Sande inverse of the above, unnormalised (returns N*x). */
void orc_ntt_inverse(uint32_t p, uint32_t psi, uint32_t logn, uint32_t* x);

/* Exact CRT of the signed convolution from canonical residues
 * r[i*n+j] = y_j mod p_i, with |y_j| < prod(p)/4, then y_j mod q. */
int orc_crt(uint32_t n, const uint32_t q[ORC_LIMBS], uint32_t k,
            const uint32_t* primes, const uint32_t* residues, uint32_t* c);

/* CRT constants of rns/host_setup.h:84-101: M_i mod q (28 words each),
 * (M_i mod p_i)^-1, and P mod q. */
int orc_crt_constants(const uint32_t q[ORC_LIMBS], uint32_t k, const uint32_t* primes,
                      uint32_t* bases_mod_q, uint32_t* basis_inverse,
                      uint32_t* product_mod_q);

/* ---- the reference's RNS algorithm end to end ---------------------------- */
/* rns/solve.cu: limb-power conversion, psi^i twist, cyclic NTT per prime,
 * pointwise product, inverse, scale, nearest-integer CRT, mod q.
 * primes/k as chosen by the caller (orc_primes_reference for N <= 2^15 is
 * the reference's own parameter set). */
int orc_rns_multiply(uint32_t n, const uint32_t q[ORC_LIMBS], uint32_t k,
                     const uint32_t* primes, const uint32_t* a,
                     const uint32_t* b, uint32_t* c);
/* The same product with its units spread over `threads` threads (1 to
 * ORC_MAX_THREADS): per prime (conversion, NTTs, pointwise, inverse, scale),
 * then per block of coefficients (CRT).  Word-for-word the result of
 * orc_rns_multiply. */
#define ORC_MAX_THREADS 256
int orc_rns_multiply_threads(uint32_t n, const uint32_t q[ORC_LIMBS], uint32_t k,
                             const uint32_t* primes, const uint32_t* a,
                             const uint32_t* b, uint32_t* c, uint32_t threads);

/* special modulus q = 2^868 - delta, 0 < delta < 2^28: the quartic RNS
 * product (quartic/host_setup.h:41-110, quartic/solve.cu:289-316,906-1040).
 * orc_quartic_delta: 0 and delta when q has that form.
 * orc_quartic_primes_reference: the reference's prime rule (16 primes
 *   2^bits - 65535 - 65536 j, bits 30 then 31, delta a fourth power, P > 4
 *   bound); returns 16 or 0; product and bound bit lengths.
 * orc_quartic_multiply: c = a*b mod (X^N+1, q) through four 217-bit
 *   components, four channels A(r i^k) per prime, negacyclic products,
 *   inverse mixing, signed CRT per component and sum_j R_j 2^(217 j) mod q;
 *   any primes with delta a fourth power, 2N | p - 1 and P > 4 bound. */
int orc_quartic_delta(const uint32_t q[ORC_LIMBS], uint32_t* delta);
uint32_t orc_quartic_primes_reference(uint32_t n, uint32_t delta, uint32_t* primes, uint32_t* product_bits,
                                      uint32_t* bound_bits);
int orc_quartic_multiply(uint32_t n, const uint32_t q[ORC_LIMBS], uint32_t k, const uint32_t* primes,
                         const uint32_t* a, const uint32_t* b, uint32_t* c);

#ifdef __cplusplus
}
#endif
#endif
