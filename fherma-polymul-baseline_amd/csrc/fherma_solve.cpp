// fherma.h entry points (fherma.h:46-48) on top of the polymul C-ABI.
//
// fherma_init  -- coverage checks and every input-independent step: RNS
//                 parameters, resident tables, device workspace
//                 (role of rns/solve.cu:581-718);
// fherma_run   -- the timed product: inputs are read, not retained; the
//                 returned Outputs owns a fresh vector (rns/solve.cu:721-853);
// fherma_free  -- releases the state (rns/solve.cu:854).
// Failures throw std::runtime_error naming the operation, as the reference
// does (rns/solve.cu:421-423); the harness records them per case.
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "fherma.h"
#include "pm_placement.h"
#include "polymul.h"

namespace {

[[noreturn]] void raise(const char* operation, const pm_context* ctx, int status) {
    throw std::runtime_error(std::string(operation) + " failed (status " + std::to_string(status) +
                             "): " + pm_last_error(ctx));
}

}  // namespace

void* fherma_init(const fherma::Point& p) {
    if (p.N < 2 || (p.N & (p.N - 1)) || p.N > 131072 || p.W != PM_WIDTH || p.L != PM_LIMBS ||
        p.q.data.size() != PM_LIMBS)
        throw std::runtime_error("coverage: power-of-two 2<=N<=131072, W=868, L=28");
    // the caller on the GPU's NUMA node before staging is allocated, as the
    // reference's init does (host_affinity.h:14-46); unknown topology: as is
    if (pm_placement_enabled()) (void)pm_place_near_device(0, nullptr, nullptr);
    pm_context* ctx = nullptr;
    int status = pm_create(p.N, p.W, p.L, p.q.data.data(), 0, &ctx);
    if (status != PM_OK) raise("fherma_init", nullptr, status);
    // PM_BACKEND=quartic selects the special-modulus backend (q = 2^868 -
    // delta, delta < 2^28), as the reference's build picks quartic/solve.cu
    // (CMakeLists.txt:29-38); any other q keeps the general one
    const char* backend = std::getenv("PM_BACKEND");
    if (backend && std::strcmp(backend, "quartic") == 0) {
        status = pm_set_backend(ctx, PM_BACKEND_QUARTIC);
        if (status != PM_OK && status != PM_EUNSUPPORTED) {
            pm_destroy(ctx);
            throw std::runtime_error(std::string("fherma_init failed: quartic backend (status ") +
                                     std::to_string(status) + ")");
        }
    }
    return ctx;
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto* ctx = static_cast<pm_context*>(state);
    if (!ctx) throw std::runtime_error("fherma_run: null state");
    // the Point's N is the context's max_n; a and b must hold exactly N x L
    // words, as the reference checks (rns/solve.cu:723-724)
    pm_info_t info{};
    int status = pm_get_info(ctx, &info);
    if (status != PM_OK) raise("fherma_run", ctx, status);
    const uint32_t n = info.max_n;
    const size_t words = size_t(n) * PM_LIMBS;
    if (in.a.data.size() != words || in.b.data.size() != words) throw std::runtime_error("RNS input size");
    status = pm_multiply_submit(ctx, n, in.a.data.data(), in.b.data.data());
    if (status != PM_OK) raise("fherma_run", ctx, status);
    // the fresh output vector is allocated (and zero-filled) while the
    // device computes
    fherma::Outputs out;
    out.c.shape = {int64_t(n), int64_t(PM_LIMBS)};
    out.c.data.resize(words);
    status = pm_multiply_collect(ctx, out.c.data.data());
    if (status != PM_OK) raise("fherma_run", ctx, status);
    return out;
}

void fherma_free(void* state) { pm_destroy(static_cast<pm_context*>(state)); }
