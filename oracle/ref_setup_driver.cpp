// Driver (own code) around the reference's parameter setup, rns/host_setup.h:61-126,
// compiled from the reference tree by oracle/Makefile into oracle/_ref/.
// Usage: ref_setup <n> <q as 28 hex words, most significant first>
// Prints the prime list, psi per prime, M_i mod q, P mod q as JSON on stdout.
#include "rns/host_setup.h"
#include <cstdio>
#include <cstdlib>
#include <string>
int main(int argc, char** argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: ref_setup n qhex\n"); return 2; }
    unsigned n = unsigned(std::strtoul(argv[1], nullptr, 10));
    std::string hex = argv[2];
    if (hex.size() != 28 * 8) { std::fprintf(stderr, "q must be 224 hex digits\n"); return 2; }
    rns::Words q(28);
    for (int w = 0; w < 28; ++w)
        q[27 - w] = uint32_t(std::strtoul(hex.substr(size_t(w) * 8, 8).c_str(), nullptr, 16));
    rns::Setup s = rns::setup(n, q);
    std::printf("{\"primes\":[");
    for (size_t i = 0; i < s.mods.size(); ++i) std::printf("%s%u", i ? "," : "", s.mods[i].p);
    std::printf("],\"reciprocal\":[");
    for (size_t i = 0; i < s.mods.size(); ++i) std::printf("%s%llu", i ? "," : "", (unsigned long long)s.mods[i].reciprocal);
    // psi = forward[prime*n + 1].value (psi^1)
    std::printf("],\"psi\":[");
    for (size_t i = 0; i < s.mods.size(); ++i) std::printf("%s%u", i ? "," : "", s.forward[i * n + 1].value);
    std::printf("],\"scale0\":[");
    for (size_t i = 0; i < s.mods.size(); ++i) std::printf("%s%u", i ? "," : "", s.scale[i * n].value);
    std::printf("],\"bases_mod_q\":[");
    for (size_t i = 0; i < s.mods.size(); ++i) {
        std::printf("%s[", i ? "," : "");
        for (unsigned w = 0; w < 28; ++w) std::printf("%s%u", w ? "," : "", s.bases_mod_q[i * rns::AccumWords + w]);
        std::printf("]");
    }
    std::printf("],\"product_mod_q\":[");
    for (unsigned w = 0; w < 28; ++w) std::printf("%s%u", w ? "," : "", s.product_mod_q[w]);
    std::printf("]}\n");
    return 0;
}
