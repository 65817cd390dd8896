// Internal definition of the opaque pm_batch.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pm_device.h"

struct pm_context;

struct pm_batch {
    pm_context* ctx = nullptr;
    uint32_t count = 0, n_items = 0, blocks = 0, reset_count = 0;
    std::vector<uint32_t> logn;
    std::vector<pmd::Job> jobs;     // host copy (device pointers)
    pmd::BatchArgs args{};
    uint32_t* d_words = nullptr;    // claim counter, error flag, per-job counters
    std::vector<void*> allocations;
    uint64_t device_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void* alloc(size_t bytes);
    ~pm_batch();
};
