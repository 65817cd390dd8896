// Host copy pool of the host-buffer path (see pm_copy_pool.h).
#include "pm_copy_pool.h"

#include "pm_placement.h"

#include <immintrin.h>

#include <chrono>
#include <cstring>

namespace {

// how long an idle worker keeps spinning before it sleeps: back-to-back
// products then find it awake, a pause between calls costs one wake-up
constexpr auto kSpin = std::chrono::microseconds(200);

}  // namespace

PmCopyPool::PmCopyPool(unsigned workers, const std::vector<int>& cpus) {
    for (unsigned i = 0; i < workers; ++i) {
        const int cpu = cpus.empty() ? -1 : cpus[i % cpus.size()];
        try {
            threads_.emplace_back([this, cpu] { worker_loop(cpu); });
        } catch (...) {
            break;  // fewer workers: the caller copies the rest
        }
    }
}

PmCopyPool::~PmCopyPool() {
    {
        std::lock_guard<std::mutex> lock(m_);
        stop_ = true;
    }
    cv_.notify_all();
    for (std::thread& t : threads_) t.join();
}

// Claims chunks in order until none is left.
void PmCopyPool::drain() {
    for (;;) {
        const uint32_t j = next_.fetch_add(1, std::memory_order_acq_rel);
        if (j >= kClosed || j >= count_.load(std::memory_order_relaxed)) return;
        const Chunk& c = chunks_[j];
        std::memcpy(c.dst, c.src, c.bytes);
        if (c.flag) {
            _mm_sfence();  // the copy (possibly streaming stores) lands before the flag
            __atomic_store_n(c.flag, c.value, __ATOMIC_RELEASE);
        }
        finished_.fetch_add(1, std::memory_order_acq_rel);
    }
}

void PmCopyPool::worker_loop(int cpu) {
    if (cpu >= 0) (void)pm_pin_thread({cpu});  // best effort: unpinned if refused
    uint64_t seen = 0;
    for (;;) {
        // spin a little, then sleep until the next generation (or stop)
        const auto until = std::chrono::steady_clock::now() + kSpin;
        while (generation_.load(std::memory_order_acquire) == seen && std::chrono::steady_clock::now() < until)
            _mm_pause();
        if (generation_.load(std::memory_order_acquire) == seen) {
            std::unique_lock<std::mutex> lock(m_);
            cv_.wait(lock, [&] { return stop_ || generation_.load(std::memory_order_acquire) != seen; });
            if (stop_) return;
        }
        seen = generation_.load(std::memory_order_acquire);
        drain();
    }
}

void PmCopyPool::run(const Chunk* chunks, uint32_t count) {
    if (count > kMaxChunks) count = kMaxChunks;
    next_.store(kClosed, std::memory_order_relaxed);  // every chunk of the last call is done
    for (uint32_t j = 0; j < count; ++j) chunks_[j] = chunks[j];
    count_.store(count, std::memory_order_relaxed);
    finished_.store(0, std::memory_order_relaxed);
    next_.store(0, std::memory_order_release);
    {
        // publish the chunks, then wake any sleeping worker
        std::lock_guard<std::mutex> lock(m_);
        generation_.fetch_add(1, std::memory_order_acq_rel);
    }
    cv_.notify_all();
    drain();
    while (finished_.load(std::memory_order_acquire) < count) _mm_pause();
}
