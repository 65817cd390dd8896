// GPU check of the fherma.h entry points' input contract: fherma_run accepts
// only a and b of exactly N x L words for the N given to fherma_init and
// throws "RNS input size" otherwise, as the reference does
// (rns/solve.cu:723-724); a failed call leaves the state usable.  Prints "ok"
// and exits 0 when every check passes.
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "fherma.h"

static int failures = 0;

static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

// fherma_run must throw std::runtime_error whose text contains `needle`
static void expect_throw(void* state, const fherma::Inputs& in, const char* needle, const char* what) {
    try {
        (void)fherma_run(state, in);
        expect(false, what);
    } catch (const std::runtime_error& e) {
        expect(std::string(e.what()).find(needle) != std::string::npos, what);
    }
}

static fherma::Tensor<uint32_t> words(size_t count) {
    fherma::Tensor<uint32_t> t;
    t.shape = {int64_t(count / 28), 28};
    t.data.assign(count, 0u);
    return t;
}

int main() {
    const uint32_t n = 1024, L = 28;
    fherma::Point point;
    point.N = n;
    point.W = 868;
    point.L = L;
    // q = 2^868 - 11993087: words 0..26 = 2^32 - 1 except word 0, word 27 = 2^4 - 1
    point.q.shape = {int64_t(L)};
    point.q.data.assign(L, 0xFFFFFFFFu);
    point.q.data[0] = 0xFFFFFFFFu - 11993087u + 1u;
    point.q.data[27] = 0xFu;
    void* state = fherma_init(point);

    fherma::Inputs shorter{words(size_t(n / 2) * L), words(size_t(n / 2) * L)};
    expect_throw(state, shorter, "RNS input size", "a and b of N/2 x L words are refused");
    fherma::Inputs uneven{words(size_t(n) * L), words(size_t(n) * L + 1)};
    expect_throw(state, uneven, "RNS input size", "b one word longer than a is refused");
    fherma::Inputs longer{words(size_t(2 * n) * L), words(size_t(2 * n) * L)};
    expect_throw(state, longer, "RNS input size", "a and b of 2N x L words are refused");

    // the state still answers: 1 * b = b
    fherma::Inputs ok{words(size_t(n) * L), words(size_t(n) * L)};
    ok.a.data[0] = 1u;
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i < ok.b.data.size(); ++i) {
        x ^= x << 13, x ^= x >> 7, x ^= x << 17;
        ok.b.data[i] = (i % L == L - 1) ? uint32_t(x) & 0x7u : uint32_t(x);  // below 2^867 < q
    }
    const fherma::Outputs out = fherma_run(state, ok);
    expect(out.c.data == ok.b.data, "1 * b == b after the refused calls");
    expect(out.c.shape.size() == 2 && out.c.shape[0] == int64_t(n) && out.c.shape[1] == int64_t(L),
           "output shape N x L");
    fherma_free(state);
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
