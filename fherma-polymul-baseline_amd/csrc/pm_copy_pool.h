// Host copy pool of the host-buffer path (role of the reference's
// host_copy_pool.h:202-234,384-440): a few resident worker threads that, with
// the calling thread, copy the input chunks of one product into page-locked
// staging and raise each chunk's ready flag as soon as that chunk has landed.
// A single core copies cold (not cached) memory at ~35 GB/s, below the PCIe
// rate the device pulls the staging at; several cores keep ahead of it.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <thread>
#include <vector>

class PmCopyPool {
public:
    struct Chunk {
        void* dst;
        const void* src;
        size_t bytes;
        volatile uint32_t* flag;  // set to `value` once the chunk is in place (none if null)
        uint32_t value;
    };
    static constexpr uint32_t kMaxChunks = 64;

    // Worker i is pinned to cpus[i % cpus.size()] (no pinning when empty).
    explicit PmCopyPool(unsigned workers, const std::vector<int>& cpus = {});
    ~PmCopyPool();
    PmCopyPool(const PmCopyPool&) = delete;
    PmCopyPool& operator=(const PmCopyPool&) = delete;

    // Copies every chunk (workers and the caller claim them in order) and
    // returns once all of them are in place.  One call at a time.
    void run(const Chunk* chunks, uint32_t count);
    unsigned workers() const { return unsigned(threads_.size()); }

private:
    void worker_loop(int cpu);
    void drain();

    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_;
    bool stop_ = false;
    std::atomic<uint64_t> generation_{0};
    // chunk claims: kClosed while run() publishes a new set of chunks, so a
    // worker still leaving the previous set claims nothing; the release store
    // of 0 publishes chunks_ and count_ to every later claim
    static constexpr uint32_t kClosed = 1u << 30;
    std::atomic<uint32_t> next_{kClosed};
    std::atomic<uint32_t> finished_{0};
    std::atomic<uint32_t> count_{0};
    Chunk chunks_[kMaxChunks];
};
