// Dumps, for the first two primes the planner picks for a fixed 868-bit q at
// max N = 2^6, what the tile's block product and complemented inverse read:
// the forward and inverse twiddles, the post-inverse scale per logn, and
// block_product_constants / complement_twiddle of pm_params.cpp.  One JSON
// object on stdout; tests/test_block_constants.py checks it against direct
// polynomial arithmetic.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "pm_params.h"

int main() {
    std::vector<uint32_t> q(28, 0);
    q[27] = 0x8u;  // 2^867 + 12345: an odd 868-bit modulus
    q[0] = 12345u;
    const uint32_t max_logn = 6, max_n = 1u << max_logn;
    pm::Params s;
    const std::string why = pm::build_params(max_logn, q.data(), 28, s);
    if (!why.empty()) {
        std::printf("{\"error\": \"%s\"}\n", why.c_str());
        return 1;
    }
    std::printf("{\"primes\": [");
    for (uint32_t i = 0; i < 2; ++i) {
        const uint32_t p = s.primes[i];
        std::vector<pm::Twiddle> s4(max_logn + 1), z(1u << (max_logn - 1));
        pm::block_product_constants(p, max_logn, s.forward.data() + size_t(i) * max_n, s.scale.data() + i, s.k,
                                    s4.data(), z.data());
        std::printf("%s{\"p\": %u, \"forward\": [", i ? ", " : "", p);
        for (uint32_t x = 0; x < max_n; ++x)
            std::printf("%s%u", x ? ", " : "", s.forward[size_t(i) * max_n + x].value);
        std::printf("], \"inverse\": [");
        for (uint32_t x = 0; x < max_n; ++x)
            std::printf("%s%u", x ? ", " : "", s.inverse[size_t(i) * max_n + x].value);
        std::printf("], \"complement\": [");
        for (uint32_t x = 0; x < max_n; ++x) {
            const pm::Twiddle c = pm::complement_twiddle(s.inverse[size_t(i) * max_n + x], p);
            std::printf("%s[%u, %u]", x ? ", " : "", c.value, c.shoup);
        }
        std::printf("], \"scale\": [");
        for (uint32_t l = 0; l <= max_logn; ++l) std::printf("%s%u", l ? ", " : "", s.scale[size_t(l) * s.k + i].value);
        std::printf("], \"scale4\": [");
        for (uint32_t l = 0; l <= max_logn; ++l) std::printf("%s[%u, %u]", l ? ", " : "", s4[l].value, s4[l].shoup);
        std::printf("], \"zeta\": [");
        for (uint32_t x = 0; x < z.size(); ++x) std::printf("%s[%u, %u]", x ? ", " : "", z[x].value, z[x].shoup);
        std::printf("]}");
    }
    std::printf("]}\n");
    return 0;
}
