// Host-side parameter math for the RNS negacyclic product (input-independent,
// computed once per modulus q).  Plain C++17, no HIP.
//
// Follows the role of rns/host_setup.h:61-126 (prime search, CRT bases,
// P mod q, per-prime roots and Shoup twiddles) with three deliberate changes:
//   * primes are p = c*2^18 + 1 < 2^30 (30-bit, so Harvey lazy butterflies keep
//     values in [0,4p) inside 32 bits, and 2N | p-1 for every N <= 2^17);
//   * one primitive 2^18-th root per prime: the bit-reversed twiddle table of
//     N = 2^17 has every smaller N as a prefix, so tables are shared by all N;
//   * the CRT accumulates M_i mod q in 30-bit limbs and reduces modulo any odd
//     868-bit q (the reference folds only q = 2^868 - c, rns/solve.cu:336-357).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pm {

constexpr uint32_t kLimbs = 28;          // u32 limbs per ABI coefficient (fherma.h:35)
constexpr uint32_t kWidth = 868;         // W
constexpr uint32_t kMinModulusBits = 865; // q > 2^864: the CRT's Barrett finish then needs one conditional subtraction
constexpr uint32_t kMaxLogN = 17;        // largest supported N = 2^17
constexpr uint32_t kMaxPrimes = 64;
constexpr uint32_t kCrtLimbs = 29;       // 30-bit limbs of M_i mod q (870 bits)
constexpr uint32_t kConvBytes = 112;     // bytes of an ABI coefficient
constexpr uint32_t kRoundShift = 42;     // fixed point of the CRT rounding estimate
constexpr uint32_t kCrtDigits = 110;     // balanced base-256 digits of M_i mod q (< 2^868)
constexpr uint32_t kCrtRowE = 116;       // first row of the rounding-estimate digits

struct Twiddle {
    uint32_t value;
    uint32_t shoup;  // floor(value * 2^32 / p)
};

struct Params {
    uint32_t max_logn = 0;
    uint32_t k = 0;                              // number of primes
    std::vector<uint32_t> q;                     // 28 words
    std::vector<uint32_t> primes;                // k
    std::vector<uint32_t> psi_max;               // k, primitive 2^(max_logn+1)-th roots
    std::vector<uint32_t> mont_neg_inv;          // k, -p^-1 mod 2^32
    std::vector<Twiddle> forward;                // k * 2^max_logn, psi^brv(i)
    std::vector<Twiddle> inverse;                // k * 2^max_logn, psi^-brv(i)
    // post-inverse scale per (logn, prime): N^-1 * (M_i mod p_i)^-1 * 2^32 mod p
    // (the 2^32 undoes the Montgomery factor of the pointwise product)
    std::vector<Twiddle> scale;                  // (max_logn+1) * k
    // residue conversion: balanced base-256 digits of 256^j mod p_i
    std::vector<int8_t> conv_digits;             // k * 112 * 4  [prime][byte j][digit b]
    std::vector<int32_t> conv_correction;        // k * 4: 128 * sum_j digit
    std::vector<uint32_t> conv_shoup16;          // k: floor(2^16 * 2^32 / p)
    std::vector<Twiddle> limb_pow;               // k * 28: 2^(32j) mod p (scalar conversion)
    // CRT
    std::vector<uint32_t> crt_bases30;           // k * 29: M_i mod q in 30-bit limbs
    std::vector<uint32_t> crt_round;             // k: floor(2^42 / p_i)
    std::vector<uint32_t> crt_neg_kp;            // (2k+1) * 28: (-K * P) mod q
    std::vector<uint32_t> barrett_mu;            // 2 words: floor(2^906 / q)
    // CRT on int8 MFMA: the 128 x 256 digit matrix A[j][4i+u] (row j = byte
    // position of the sum; balanced base-256 digit j-u of M_i mod q for
    // j < 113, of floor(2^42/p_i) at rows 116 + (j-u)) laid out as the A
    // operand of v_mfma_i32_16x16x64_i8: [row tile 8][k step 4][lane 64][16]
    std::vector<int8_t> crt_mfma_a;              // 8*4*64*16
    std::vector<int32_t> crt_mfma_corr;          // 128: 128 * sum_k A[j][k] (unsigned-byte correction)
    uint32_t product_bits = 0;
};

// Returns an empty string on success, else the reason the modulus/size is
// unsupported (mirrors the coverage checks of rns/solve.cu:581-590).
std::string build_params(uint32_t max_logn, const uint32_t* q, uint32_t limbs, Params& out);

// ---- special modulus q = 2^868 - delta, 0 < delta < 2^28: the quartic RNS
// (role of quartic/host_setup.h:41-110).  With B = 2^217, q = B^4 - delta, a
// coefficient is a0 + a1 B + a2 B^2 + a3 B^3 (a_j < 2^217) and the product is
// taken in Z[X, Y]/(X^N + 1, Y^4 - delta): four signed integer components R_j,
// |R_j| <= N (B-1)^2 (1 + 3 delta).  Modulo each prime p (delta a fourth
// power mod p, fourth root r, i = psi^(N/2) with i^2 = -1), Y^4 - delta splits
// into Y - r i^k, k = 0..3: four independent negacyclic channels per prime.
// Same prime family as the general RNS (c 2^18 + 1 < 2^30); as many primes
// as make P > 4 max|R_j| for N = max_n.
constexpr uint32_t kQuarticPartBits = 217;
constexpr uint32_t kQuarticPartWords = 7;   // 32-bit limbs of a component
constexpr uint32_t kQuarticLimbs30 = 18;    // 30-bit limbs of M_i, K*P and the CRT sums (540 bits)
constexpr uint32_t kQuarticMaxDelta = 1u << 28;
constexpr uint32_t kQuarticMaxPrimes = 17;  // 68 rows: enough for every delta < 2^28 at N = 2^17

struct QuarticParams {
    uint32_t max_logn = 0;
    uint32_t k = 0;                    // primes; 4k residue rows (row 4i + j: channel j of prime i)
    uint32_t delta = 0;
    uint32_t product_bits = 0;
    std::vector<uint32_t> primes;      // k
    std::vector<uint32_t> psi_max;     // k, primitive 2^(max_logn+1)-th roots
    std::vector<uint32_t> mont_neg_inv;
    std::vector<uint32_t> root4;       // k: r, r^4 = delta mod p
    std::vector<Twiddle> limb_pow;     // k * 7: 2^(32 l) mod p
    std::vector<Twiddle> rpow;         // k * 4: r^j (forward mixing)
    std::vector<Twiddle> rinv;         // k * 4: r^-j (inverse mixing)
    std::vector<Twiddle> iota;         // k: i, i^2 = -1
    std::vector<Twiddle> forward;      // k * 2^max_logn, psi^brv(i) (shared by the 4 channels)
    std::vector<Twiddle> inverse;
    // post-inverse scale per (logn, prime): N^-1 (M_i mod p_i)^-1 4^-1 2^32 mod p
    std::vector<Twiddle> scale;        // (max_logn+1) * k
    std::vector<uint32_t> m30;         // k * 18: M_i = P / p_i in 30-bit limbs
    std::vector<uint32_t> round;       // k: floor(2^42 / p_i)
    std::vector<uint32_t> kp30;        // (k+1) * 18: K * P in 30-bit limbs
    std::vector<uint32_t> negc;        // 28 words: (-k P (1 + B + B^2 + B^3)) mod q
};

// Empty on success, else why q (28 words) is not a supported special modulus.
std::string build_quartic_params(uint32_t max_logn, const uint32_t* q, QuarticParams& out);

// (p - w, floor((p - w) 2^32 / p)): the twiddle whose Shoup product is
// congruent to -w times the operand (w in [1, p)).
Twiddle complement_twiddle(Twiddle w, uint32_t p);

// Constants of the tile's block product for one prime p.  For N = 2^logn >= 4
// the tile's forward transform stops at blocks of 4 positions, block G
// holding its operand modulo X^4 - zeta_G, and its inverse runs logn - 2
// stages, so the product carries 4 times the post-inverse scale:
//   scale4[logn] = 4 scale[logn * scale_stride] mod p  (logn >= 2; as is below),
//   zeta[2^(logn-2) + G] = zeta_G scale4[logn] mod p   (G < N/4),
// zeta_G = +w (G even) or -w (G odd) with w the forward twiddle (psi^brv) of
// index 2^(logn-3) + G/2, and zeta = -1 for N = 4.  zeta has 2^(max_logn-1)
// entries (entry 0 unused).
void block_product_constants(uint32_t p, uint32_t max_logn, const Twiddle* forward, const Twiddle* scale,
                             size_t scale_stride, Twiddle* scale4, Twiddle* zeta);

// Shared helpers (also used by tests through the C-ABI).
uint32_t pow_mod(uint32_t a, uint64_t e, uint32_t p);
bool is_prime(uint32_t p);
uint32_t bit_reverse(uint32_t x, uint32_t bits);

}  // namespace pm
