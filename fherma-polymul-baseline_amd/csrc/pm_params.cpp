// Host parameter math; see pm_params.h for what differs from the reference's
// rns/host_setup.h:61-126 and why.
#include "pm_params.h"

#include <algorithm>
#include <stdexcept>

namespace pm {
namespace {

using Big = std::vector<uint32_t>;  // little-endian u32 words

void trim(Big& a) {
    while (!a.empty() && a.back() == 0) a.pop_back();
}

int cmp(Big a, Big b) {
    trim(a);
    trim(b);
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

Big mul_small(const Big& a, uint32_t s) {
    Big r(a.size() + 1, 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t t = uint64_t(a[i]) * s + carry;
        r[i] = uint32_t(t);
        carry = t >> 32;
    }
    r[a.size()] = uint32_t(carry);
    trim(r);
    return r;
}

Big mul(const Big& a, const Big& b) {
    Big r(a.size() + b.size() + 1, 0);
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t carry = 0;
        for (size_t j = 0; j < b.size(); ++j) {
            uint64_t t = uint64_t(a[i]) * b[j] + r[i + j] + carry;
            r[i + j] = uint32_t(t);
            carry = t >> 32;
        }
        for (size_t k = i + b.size(); carry; ++k) {
            uint64_t t = uint64_t(r[k]) + carry;
            r[k] = uint32_t(t);
            carry = t >> 32;
        }
    }
    trim(r);
    return r;
}

Big sub(const Big& a, const Big& b) {  // a >= b
    Big r(a.size(), 0);
    int64_t borrow = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        int64_t t = int64_t(a[i]) - (i < b.size() ? int64_t(b[i]) : 0) - borrow;
        borrow = t < 0;
        r[i] = uint32_t(t + (borrow ? (int64_t(1) << 32) : 0));
    }
    trim(r);
    return r;
}

uint32_t bits(const Big& a) {
    for (size_t i = a.size(); i-- > 0;)
        if (a[i]) return uint32_t(i * 32 + 32 - __builtin_clz(a[i]));
    return 0;
}

Big shl1_or(const Big& a, uint32_t bit) {
    Big r(a.size() + 1, 0);
    uint32_t carry = bit;
    for (size_t i = 0; i < a.size(); ++i) {
        r[i] = (a[i] << 1) | carry;
        carry = a[i] >> 31;
    }
    r[a.size()] = carry;
    trim(r);
    return r;
}

// floor(a / d) and a mod d by binary long division (init time only).
void divmod(const Big& a, const Big& d, Big* quotient, Big* remainder) {
    Big r, qt(a.size() + 1, 0);
    for (uint32_t bit = bits(a); bit-- > 0;) {
        r = shl1_or(r, (a[bit / 32] >> (bit % 32)) & 1u);
        if (cmp(r, d) >= 0) {
            r = sub(r, d);
            qt[bit / 32] |= 1u << (bit % 32);
        }
    }
    trim(qt);
    if (quotient) *quotient = qt;
    if (remainder) *remainder = r;
}

uint32_t mod_small(const Big& a, uint32_t p) {
    uint64_t r = 0;
    for (size_t i = a.size(); i-- > 0;) r = ((r << 32) | a[i]) % p;
    return uint32_t(r);
}

Big div_small(const Big& a, uint32_t p, uint32_t* rem) {
    Big r(a.size(), 0);
    uint64_t carry = 0;
    for (size_t i = a.size(); i-- > 0;) {
        uint64_t v = (carry << 32) | a[i];
        r[i] = uint32_t(v / p);
        carry = v % p;
    }
    if (rem) *rem = uint32_t(carry);
    trim(r);
    return r;
}

Big pow2(uint32_t e) {
    Big r(e / 32 + 1, 0);
    r[e / 32] = 1u << (e % 32);
    return r;
}

Twiddle make_twiddle(uint32_t value, uint32_t p) {
    return Twiddle{value, uint32_t((uint64_t(value) << 32) / p)};
}

uint32_t mulmod(uint32_t a, uint32_t b, uint32_t p) { return uint32_t(uint64_t(a) * b % p); }

}  // namespace

uint32_t pow_mod(uint32_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1 % p, b = a % p;
    while (e) {
        if (e & 1) r = r * b % p;
        e >>= 1;
        if (e) b = b * b % p;
    }
    return uint32_t(r);
}

bool is_prime(uint32_t p) {
    if (p < 2) return false;
    if (p % 2 == 0) return p == 2;
    // deterministic Miller-Rabin for 32-bit inputs
    uint32_t d = p - 1, s = 0;
    while (d % 2 == 0) { d /= 2; ++s; }
    for (uint32_t a : {2u, 3u, 5u, 7u}) {
        if (a % p == 0) continue;
        uint64_t x = pow_mod(a, d, p);
        if (x == 1 || x == p - 1) continue;
        bool composite = true;
        for (uint32_t r = 1; r < s; ++r) {
            x = x * x % p;
            if (x == p - 1) { composite = false; break; }
        }
        if (composite) return false;
    }
    return true;
}

uint32_t bit_reverse(uint32_t x, uint32_t nbits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < nbits; ++i) r |= ((x >> i) & 1u) << (nbits - 1 - i);
    return r;
}

std::string build_params(uint32_t max_logn, const uint32_t* q_words, uint32_t limbs, Params& s) {
    if (limbs != kLimbs) return "L must be 28 u32 limbs";
    if (max_logn < 1 || max_logn > kMaxLogN) return "N must be a power of two in [2, 2^17]";
    Big q(q_words, q_words + kLimbs);
    trim(q);
    if (q.empty() || (q[0] & 1u) == 0) return "q must be odd";
    if (bits(q) > kWidth) return "q must be below 2^868";
    // The CRT's Barrett step estimates the quotient of X < 2^905 from
    // floor(X / 2^863) and floor(2^906 / q); its error is below
    // 1 + X/2^906 + 2^863/q, under 2 (one conditional subtraction) when q > 2^864.
    if (bits(q) < kMinModulusBits) return "q must have 865 to 868 bits (q > 2^864)";
    s = Params{};
    s.max_logn = max_logn;
    s.q.assign(q_words, q_words + kLimbs);
    const uint32_t max_n = 1u << max_logn;
    const uint32_t root_order_log = max_logn + 1;

    // prime search: largest p = c*2^18 + 1 < 2^30 first, until P > 4*max_n*(q-1)^2
    Big qm1 = sub(q, Big{1});
    Big bound = mul_small(mul(qm1, qm1), 4u * max_n);
    Big product{1};
    for (uint32_t c = (1u << 12) - 1; c > 0 && s.primes.size() < kMaxPrimes; --c) {
        const uint32_t p = (c << 18) + 1u;
        if (!is_prime(p)) continue;
        s.primes.push_back(p);
        product = mul_small(product, p);
        if (cmp(product, bound) > 0) break;
    }
    if (cmp(product, bound) <= 0) return "not enough RNS primes for this modulus";
    s.k = uint32_t(s.primes.size());
    s.product_bits = bits(product);
    const uint32_t k = s.k;

    // roots: psi_max of order 2^(max_logn+1)
    for (uint32_t p : s.primes) {
        uint32_t psi = 0;
        for (uint32_t g = 2; g < 100000 && !psi; ++g) {
            uint32_t r = pow_mod(g, (p - 1) >> root_order_log, p);
            if (pow_mod(r, 1ull << (root_order_log - 1), p) == p - 1) psi = r;
        }
        if (!psi) return "RNS root not found";
        s.psi_max.push_back(psi);
        // -p^-1 mod 2^32 by Newton iteration
        uint32_t inv = p;  // correct to 3 bits for odd p
        for (int i = 0; i < 5; ++i) inv *= 2u - p * inv;
        s.mont_neg_inv.push_back(0u - inv);
    }

    // twiddles: forward psi^brv(i), inverse psi^-brv(i), i < max_n
    s.forward.resize(size_t(k) * max_n);
    s.inverse.resize(size_t(k) * max_n);
    std::vector<uint32_t> pw(max_n), ipw(max_n);
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t p = s.primes[i], psi = s.psi_max[i], ipsi = pow_mod(psi, p - 2, p);
        uint32_t f = 1, b = 1;
        for (uint32_t e = 0; e < max_n; ++e) {
            pw[e] = f;
            ipw[e] = b;
            f = mulmod(f, psi, p);
            b = mulmod(b, ipsi, p);
        }
        for (uint32_t j = 0; j < max_n; ++j) {
            const uint32_t r = bit_reverse(j, max_logn);
            s.forward[size_t(i) * max_n + j] = make_twiddle(pw[r], p);
            s.inverse[size_t(i) * max_n + j] = make_twiddle(ipw[r], p);
        }
    }

    // CRT constants
    s.crt_bases30.assign(size_t(k) * kCrtLimbs, 0);
    s.crt_round.resize(k);
    std::vector<uint32_t> basis_inverse(k);
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t p = s.primes[i];
        uint32_t rem = 0;
        Big basis = div_small(product, p, &rem);
        if (rem) return "CRT basis division not exact";
        Big reduced;
        divmod(basis, q, nullptr, &reduced);
        // 30-bit limbs
        for (uint32_t l = 0; l < kCrtLimbs; ++l) {
            uint32_t v = 0;
            for (uint32_t b = 0; b < 30; ++b) {
                const uint32_t bit = 30 * l + b;
                if (bit / 32 < reduced.size() && ((reduced[bit / 32] >> (bit % 32)) & 1u)) v |= 1u << b;
            }
            s.crt_bases30[size_t(i) * kCrtLimbs + l] = v;
        }
        basis_inverse[i] = pow_mod(mod_small(basis, p), p - 2, p);
        s.crt_round[i] = uint32_t((uint64_t(1) << kRoundShift) / p);
        if (i >= 64) return "too many primes for the MFMA CRT (K = 4 bytes x 64 primes)";
        // balanced base-256 digits of M_i mod q and of floor(2^42/p_i)
        int8_t dig[kCrtDigits];
        int64_t carry = 0;
        for (uint32_t v = 0; v < kCrtDigits; ++v) {
            const uint32_t word = v / 4 < reduced.size() ? reduced[v / 4] : 0u;
            const int64_t x = int64_t((word >> (8 * (v % 4))) & 255u) + carry;
            const int64_t d = ((x + 128) & 255) - 128;
            dig[v] = int8_t(d);
            carry = (x - d) / 256;
        }
        if (carry != 0) return "balanced digits of a CRT basis overflow";
        int64_t rv = s.crt_round[i];
        const int64_t e0 = ((rv + 128) & 255) - 128, e1 = (rv - e0) / 256;
        if (e1 < -128 || e1 > 127) return "rounding constant out of digit range";
        if (s.crt_mfma_a.empty()) {
            s.crt_mfma_a.assign(8 * 4 * 64 * 16, 0);
            s.crt_mfma_corr.assign(128, 0);
        }
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t kk = 4 * i + u;
            auto put = [&](uint32_t row, int8_t value) {
                const uint32_t rt = row / 16, s_ = kk / 64, lane = (row % 16) + 16 * ((kk % 64) / 16);
                s.crt_mfma_a[((size_t(rt) * 4 + s_) * 64 + lane) * 16 + kk % 16] = value;
                s.crt_mfma_corr[row] += 128 * int32_t(value);
            };
            for (uint32_t v = 0; v < kCrtDigits; ++v) put(u + v, dig[v]);
            put(kCrtRowE + u, int8_t(e0));
            put(kCrtRowE + u + 1, int8_t(e1));
        }
    }
    Big product_mod_q;
    divmod(product, q, nullptr, &product_mod_q);
    s.crt_neg_kp.assign(size_t(2 * k + 1) * kLimbs, 0);
    for (uint32_t K = 0; K <= 2 * k; ++K) {  // residues below 2p_i: K <= 2k
        // (-K * P) mod q = (q - (K*P mod q)) mod q
        Big v;
        divmod(mul_small(product_mod_q, K), q, nullptr, &v);
        if (!v.empty()) v = sub(q, v);
        for (size_t w = 0; w < v.size() && w < kLimbs; ++w) s.crt_neg_kp[size_t(K) * kLimbs + w] = v[w];
    }
    Big mu;
    divmod(pow2(906), q, &mu, nullptr);  // floor(2^906 / q)
    mu.resize(2, 0);
    s.barrett_mu = {mu[0], mu[1]};

    // post-inverse scale: N^-1 (M_i mod p_i)^-1 2^32 mod p, per logn
    s.scale.resize(size_t(max_logn + 1) * k);
    for (uint32_t logn = 0; logn <= max_logn; ++logn)
        for (uint32_t i = 0; i < k; ++i) {
            const uint32_t p = s.primes[i];
            uint32_t v = mulmod(pow_mod(1u << logn, p - 2, p), basis_inverse[i], p);
            v = mulmod(v, uint32_t((uint64_t(1) << 32) % p), p);
            s.scale[size_t(logn) * k + i] = make_twiddle(v, p);
        }

    // residue conversion digits
    s.conv_digits.assign(size_t(k) * kConvBytes * 4, 0);
    s.conv_correction.assign(size_t(k) * 4, 0);
    s.conv_shoup16.resize(k);
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t p = s.primes[i];
        uint32_t r = 1;
        for (uint32_t j = 0; j < kConvBytes; ++j) {
            int64_t v = r;  // r < p < 2^30 has a 4-digit balanced form
            for (uint32_t b = 0; b < 4; ++b) {
                int64_t d = ((v + 128) & 255) - 128;
                s.conv_digits[(size_t(i) * kConvBytes + j) * 4 + b] = int8_t(d);
                s.conv_correction[size_t(i) * 4 + b] += int32_t(128 * d);
                v = (v - d) / 256;
            }
            if (v != 0) return "balanced digit decomposition failed";
            r = mulmod(r, 256u, p);
        }
        s.conv_shoup16[i] = uint32_t((uint64_t(1) << 48) / p);
        // The conversion keeps v = u + p + ((w + 2p) 2^16 mod p, in [0,2p)) with
        // u = T0 + 256 T1, w = T2 + 256 T3 (T_d the byte-weighted digit sums)
        // and reduces it at most once by 2p: exact ranges of u and w over any
        // 112 bytes must give w + 2p in [0, 2^32), v in [0, 2^32) and v - 2p
        // below 4p.  (Wide margins for every prime of this form: u >= -0.31 p,
        // u + 3p <= 0.81 * 2^32.)
        int64_t u_lo = 0, u_hi = 0, w_lo = 0, w_hi = 0;
        for (uint32_t j = 0; j < kConvBytes; ++j) {
            const int8_t* d = &s.conv_digits[(size_t(i) * kConvBytes + j) * 4];
            const int64_t e = int64_t(d[0]) + 256 * int64_t(d[1]), f = int64_t(d[2]) + 256 * int64_t(d[3]);
            (e < 0 ? u_lo : u_hi) += 255 * e;
            (f < 0 ? w_lo : w_hi) += 255 * f;
        }
        const int64_t P = p, two32 = int64_t(1) << 32;
        if (u_lo < -P || u_hi + 3 * P > two32 || u_hi >= 3 * P || w_lo < -2 * P || w_hi + 2 * P >= two32)
            return "residue conversion bounds do not hold for a prime";
    }
    s.limb_pow.resize(size_t(k) * kLimbs);
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t p = s.primes[i], base = uint32_t((uint64_t(1) << 32) % p);
        uint32_t v = 1;
        for (uint32_t j = 0; j < kLimbs; ++j) {
            s.limb_pow[size_t(i) * kLimbs + j] = make_twiddle(v, p);
            v = mulmod(v, base, p);
        }
    }
    return std::string();
}

namespace {

// a (any size) as n limbs of 30 bits; false when it does not fit
bool to_limbs30(const Big& a, uint32_t n, uint32_t* out) {
    for (uint32_t l = 0; l < n; ++l) {
        uint32_t v = 0;
        for (uint32_t b = 0; b < 30; ++b) {
            const uint32_t bit = 30 * l + b;
            if (bit / 32 < a.size() && ((a[bit / 32] >> (bit % 32)) & 1u)) v |= 1u << b;
        }
        out[l] = v;
    }
    return bits(a) <= 30 * n;
}

// square root of a quadratic residue a mod an odd prime p (Tonelli-Shanks);
// 0 when a is not a residue
uint32_t sqrt_mod(uint32_t a, uint32_t p) {
    a %= p;
    if (a == 0) return 0;
    if (pow_mod(a, (p - 1) / 2, p) != 1) return 0;
    uint32_t q = p - 1, s = 0;
    while ((q & 1u) == 0) { q >>= 1; ++s; }
    uint32_t z = 2;
    while (pow_mod(z, (p - 1) / 2, p) != p - 1) ++z;
    uint32_t m = s, c = pow_mod(z, q, p), t = pow_mod(a, q, p), r = pow_mod(a, (q + 1) / 2, p);
    while (t != 1) {
        uint32_t i = 0, t2 = t;
        while (t2 != 1 && i < m) { t2 = mulmod(t2, t2, p); ++i; }
        if (i == m) return 0;
        const uint32_t b = pow_mod(c, uint64_t(1) << (m - i - 1), p);
        m = i;
        c = mulmod(b, b, p);
        t = mulmod(t, c, p);
        r = mulmod(r, b, p);
    }
    return r;
}

}  // namespace

std::string build_quartic_params(uint32_t max_logn, const uint32_t* q_words, QuarticParams& s) {
    if (max_logn < 2 || max_logn > kMaxLogN) return "N must be a power of two in [4, 2^17]";
    // q = 2^868 - delta with 0 < delta < 2^28
    Big q(q_words, q_words + kLimbs);
    trim(q);
    Big two868 = pow2(kWidth);
    if (cmp(q, two868) >= 0) return "q must be below 2^868";
    Big d = sub(two868, q);
    if (d.empty() || bits(d) > 28 || d[0] >= kQuarticMaxDelta) return "q is not 2^868 - delta with 0 < delta < 2^28";
    const uint32_t delta = d[0];
    if ((delta & 1u) == 0) return "q must be odd";
    s = QuarticParams{};
    s.max_logn = max_logn;
    s.delta = delta;
    const uint32_t max_n = 1u << max_logn, root_order_log = max_logn + 1;
    // |R_j| <= N (B-1)^2 (1 + 3 delta); P > 4 of that
    Big part = sub(pow2(kQuarticPartBits), Big{1});
    Big bound = mul_small(mul_small(mul(part, part), max_n), 1u + 3u * delta);
    Big four_bound = mul_small(bound, 4u);
    Big product{1};
    for (uint32_t c = (1u << 12) - 1; c > 0 && s.primes.size() < kQuarticMaxPrimes; --c) {
        const uint32_t p = (c << 18) + 1u;
        if (!is_prime(p) || pow_mod(delta, (p - 1) / 4, p) != 1) continue;  // delta must be a fourth power
        s.primes.push_back(p);
        product = mul_small(product, p);
        if (cmp(product, four_bound) > 0) break;
    }
    if (cmp(product, four_bound) <= 0) return "not enough quartic primes for this modulus";
    s.k = uint32_t(s.primes.size());
    s.product_bits = bits(product);
    const uint32_t k = s.k;
    std::vector<uint32_t> basis_inverse(k);
    s.m30.assign(size_t(k) * kQuarticLimbs30, 0);
    s.round.resize(k);
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t p = s.primes[i];
        uint32_t psi = 0;
        for (uint32_t g = 2; g < 100000 && !psi; ++g) {
            const uint32_t r = pow_mod(g, (p - 1) >> root_order_log, p);
            if (pow_mod(r, 1ull << (root_order_log - 1), p) == p - 1) psi = r;
        }
        if (!psi) return "quartic RNS root not found";
        s.psi_max.push_back(psi);
        uint32_t inv = p;
        for (int t = 0; t < 5; ++t) inv *= 2u - p * inv;
        s.mont_neg_inv.push_back(0u - inv);
        // fourth root of delta: a square root of a square root (both exist:
        // delta is a fourth power and -1 = i^2 is a square)
        const uint32_t r = sqrt_mod(sqrt_mod(delta, p), p);
        if (r == 0 || pow_mod(r, 4, p) != delta % p) return "fourth root of delta not found";
        s.root4.push_back(r);
        const uint32_t iota = pow_mod(psi, uint64_t(1) << (max_logn - 1), p);  // psi has order 2^(max_logn+1)
        if (mulmod(iota, iota, p) != p - 1) return "fourth root of unity check failed";
        s.iota.push_back(make_twiddle(iota, p));
        const uint32_t rinv = pow_mod(r, p - 2, p);
        for (uint32_t j = 0; j < 4; ++j) {
            s.rpow.push_back(make_twiddle(pow_mod(r, j, p), p));
            s.rinv.push_back(make_twiddle(pow_mod(rinv, j, p), p));
        }
        const uint32_t base = uint32_t((uint64_t(1) << 32) % p);
        uint32_t v = 1;
        for (uint32_t l = 0; l < kQuarticPartWords; ++l) {
            s.limb_pow.push_back(make_twiddle(v, p));
            v = mulmod(v, base, p);
        }
        uint32_t rem = 0;
        Big basis = div_small(product, p, &rem);
        if (rem) return "quartic CRT basis division not exact";
        if (!to_limbs30(basis, kQuarticLimbs30, &s.m30[size_t(i) * kQuarticLimbs30]))
            return "quartic CRT basis exceeds its limbs";
        basis_inverse[i] = pow_mod(mod_small(basis, p), p - 2, p);
        s.round[i] = uint32_t((uint64_t(1) << kRoundShift) / p);
    }
    s.kp30.assign(size_t(k + 1) * kQuarticLimbs30, 0);
    for (uint32_t K = 0; K <= k; ++K)
        if (!to_limbs30(mul_small(product, K), kQuarticLimbs30, &s.kp30[size_t(K) * kQuarticLimbs30]))
            return "quartic K*P exceeds its limbs";
    // (-k P (1 + B + B^2 + B^3)) mod q
    Big w;
    for (uint32_t j = 0; j < 4; ++j) {
        Big t = pow2(kQuarticPartBits * j);
        Big sum(std::max(w.size(), t.size()) + 1, 0);
        uint64_t carry = 0;
        for (size_t i = 0; i < sum.size(); ++i) {
            const uint64_t x = (i < w.size() ? w[i] : 0u) + uint64_t(i < t.size() ? t[i] : 0u) + carry;
            sum[i] = uint32_t(x);
            carry = x >> 32;
        }
        trim(sum);
        w = sum;
    }
    Big kpw;
    divmod(mul(mul_small(product, k), w), q, nullptr, &kpw);
    if (!kpw.empty()) kpw = sub(q, kpw);
    s.negc.assign(kLimbs, 0);
    for (size_t i = 0; i < kpw.size() && i < kLimbs; ++i) s.negc[i] = kpw[i];
    // twiddles (shared by a prime's four channels) and the post-inverse scale
    s.forward.resize(size_t(k) * max_n);
    s.inverse.resize(size_t(k) * max_n);
    std::vector<uint32_t> pw(max_n), ipw(max_n);
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t p = s.primes[i], psi = s.psi_max[i], ipsi = pow_mod(psi, p - 2, p);
        uint32_t f = 1, b = 1;
        for (uint32_t e = 0; e < max_n; ++e) {
            pw[e] = f;
            ipw[e] = b;
            f = mulmod(f, psi, p);
            b = mulmod(b, ipsi, p);
        }
        for (uint32_t j = 0; j < max_n; ++j) {
            const uint32_t r = bit_reverse(j, max_logn);
            s.forward[size_t(i) * max_n + j] = make_twiddle(pw[r], p);
            s.inverse[size_t(i) * max_n + j] = make_twiddle(ipw[r], p);
        }
    }
    s.scale.resize(size_t(max_logn + 1) * k);
    for (uint32_t logn = 0; logn <= max_logn; ++logn)
        for (uint32_t i = 0; i < k; ++i) {
            const uint32_t p = s.primes[i];
            uint32_t v = mulmod(pow_mod(4u << logn, p - 2, p), basis_inverse[i], p);  // (4N)^-1 (M_i)^-1
            v = mulmod(v, uint32_t((uint64_t(1) << 32) % p), p);
            s.scale[size_t(logn) * k + i] = make_twiddle(v, p);
        }
    return std::string();
}

Twiddle complement_twiddle(Twiddle w, uint32_t p) { return make_twiddle(w.value ? p - w.value : 0u, p); }

void block_product_constants(uint32_t p, uint32_t max_logn, const Twiddle* forward, const Twiddle* scale,
                             size_t scale_stride, Twiddle* scale4, Twiddle* zeta) {
    zeta[0] = make_twiddle(0, p);
    for (uint32_t logn = 0; logn <= max_logn; ++logn) {
        const uint32_t s = scale[size_t(logn) * scale_stride].value;
        const uint32_t S = logn >= 2 ? mulmod(4, s, p) : s;
        scale4[logn] = make_twiddle(S, p);
        if (logn < 2) continue;
        for (uint32_t g = 0; g < (1u << (logn - 2)); ++g) {
            uint32_t z = p - 1;  // N = 4: X^4 + 1
            if (logn >= 3) {
                const uint32_t w = forward[(1u << (logn - 3)) + g / 2].value;
                z = (g & 1u) && w ? p - w : w;
            }
            zeta[(1u << (logn - 2)) + g] = make_twiddle(mulmod(z, S, p), p);
        }
    }
}

}  // namespace pm
