// Host check of the copy pool of the host-buffer path: many calls in a row
// with different chunk counts and sizes; every chunk must land whole and its
// flag must carry the call's value.  Built with -fsanitize=thread by
// tests/test_copy_pool.py so a data race fails the run.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pm_copy_pool.h"

int main(int argc, char** argv) {
    const unsigned workers = argc > 1 ? unsigned(std::atoi(argv[1])) : 3u;
    const int calls = argc > 2 ? std::atoi(argv[2]) : 300;
    PmCopyPool pool(workers);
    std::vector<uint32_t> src(1 << 20), dst(1 << 20);
    volatile uint32_t flags[PmCopyPool::kMaxChunks] = {};
    for (int call = 1; call <= calls; ++call) {
        const uint32_t chunks = 1u + uint32_t(call * 7) % 16u;
        const size_t words = size_t(1 + call % 5) * 4096;  // per chunk
        for (size_t i = 0; i < chunks * words; ++i) src[i] = uint32_t(i * 2654435761u) ^ uint32_t(call);
        std::vector<PmCopyPool::Chunk> list(chunks);
        for (uint32_t j = 0; j < chunks; ++j)
            list[j] = {dst.data() + j * words, src.data() + j * words, words * sizeof(uint32_t), &flags[j],
                       uint32_t(call)};
        pool.run(list.data(), chunks);
        for (uint32_t j = 0; j < chunks; ++j)
            if (flags[j] != uint32_t(call)) {
                std::printf("call %d: flag %u not raised\n", call, j);
                return 1;
            }
        if (std::memcmp(dst.data(), src.data(), chunks * words * sizeof(uint32_t)) != 0) {
            std::printf("call %d: data differs\n", call);
            return 1;
        }
    }
    std::printf("ok: %d calls, %u workers\n", calls, pool.workers());
    return 0;
}
