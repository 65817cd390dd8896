// Internal definition of the opaque pm_batch.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pm_device.h"
#include "pm_kernels.h"

struct pm_context;

// Dependency-ordered work list of a set of products (shared by the batch and
// the serving loop).
bool pm_phase_ranges(const std::vector<pmd::Item>& items, pmd::PhaseRanges& ranges);
// tiles_per_block[tlog]: tiles (units) per tile item of each size class,
// NULL for the default (pm_default_batch_tiles); residues_per_thread[tlog]:
// 16 or 32 (NULL, 0: 32)
void pm_build_worklist(uint32_t k, const pmd::Job* jobs, uint32_t count, bool largest_first,
                       std::vector<pmd::Item>& items, std::vector<pmd::Unit>& units,
                       const uint32_t* tiles_per_block = nullptr, const uint32_t* residues_per_thread = nullptr);

struct pm_batch {
    pm_context* ctx = nullptr;
    uint32_t count = 0, n_items = 0;
    std::vector<uint32_t> logn;
    std::vector<pmd::Job> jobs;     // host copy (device pointers)
    pmd::Tables tables{};           // the context's backend when the batch was created
    pmd::BatchArgs args{};
    pmd::PhaseRanges ranges{};
    bool scalar_conversion = false; // phased only: scalar limb reduction instead of MFMA
    bool scalar_crt = false;        // phased only: vector-ALU CRT sums instead of MFMA
    bool separate_outer = false;    // outer NTT stages as passes of their own, not inside conversion / CRT
    uint32_t crt_waves = 0;         // waves per block of the MFMA CRT (the context's at creation)
    uint32_t parts = 1;             // staggered parts of the graph (1: one part)
    bool persistent = false;        // one launch for the whole work list (grid barriers between phases)
    uint32_t* sync = nullptr;       // persistent: [2] barrier counters, [2] error word (device memory)
    uint32_t persistent_blocks = 0;
    std::vector<void*> allocations;
    uint64_t device_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t last_run = nullptr;  // recorded behind every replay, on whichever stream it ran

    void* alloc(size_t bytes);
    ~pm_batch();
};
