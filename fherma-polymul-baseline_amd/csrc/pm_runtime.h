// Internal definition of the opaque pm_context and shared runtime helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "pm_copy_pool.h"
#include "pm_device.h"
#include "pm_params.h"

struct pm_context {
    int device = 0;
    uint32_t max_logn = 0;
    pm::Params params;
    pmd::Tables tables{};
    std::vector<void*> allocations;
    uint64_t table_bytes = 0, workspace_bytes = 0;
    hipStream_t stream = nullptr;
    // plain-path workspace
    uint32_t *d_a = nullptr, *d_b = nullptr, *d_c = nullptr, *d_ra = nullptr, *d_rb = nullptr;
    // recorded behind every plain-path product on the stream it ran on; the
    // next product (device or host-buffer, on any stream) waits for it, so two
    // products never share the residue workspace at once
    hipEvent_t workspace_free = nullptr;
    bool workspace_used = false;
    // this device's static residue scratch (pm_multiply_device_noalloc)
    uint32_t* scratch = nullptr;
    uint32_t workspace_rows = 0;    // residue rows of the plain-path workspace
    // quartic backend (q = 2^868 - delta): parameters and device tables, built
    // by the first pm_set_backend(PM_BACKEND_QUARTIC); backend selects which
    // tables every path of the context uses
    pm::QuarticParams qparams;
    pmd::Tables qtables{};
    bool quartic_built = false;
    uint32_t backend = 0;
    // outer radix-2 stages run as separate passes, per log2 N: plain path,
    // and batches / served launches
    uint32_t outer[pm::kMaxLogN + 1] = {};
    // batches: one table per backend (index PM_BACKEND_*), since only the
    // general backend runs its outer stages inside conversion and CRT
    uint32_t batch_outer[2][pm::kMaxLogN + 1] = {};
    // tiles per block of each tile size class 2^tlog in batches and served
    // launches (blocks of tiles x 2^tlog / batch_rpt[tlog] threads), and the
    // residues each thread holds (32, or 16 for 2^6 .. 2^14)
    uint32_t batch_tiles[16] = {};
    uint32_t batch_rpt[16] = {};
    // waves per block of the batches' MFMA CRT (2, 4 or 8)
    uint32_t batch_crt_waves = pmd::kCrtMfmaWaves;
    // staggered parts of a batch's graph (1-4; 0: chosen per batch)
    uint32_t batch_parts = 0;
    // plain path: slices of the primes in the conversion, per log2 N (0:
    // chosen at launch)
    uint32_t conv_slices[pm::kMaxLogN + 1] = {};
    // host-buffer path: page-locked staging (inputs a|b and the output) with
    // its device views, output chunk events, and the product between submit
    // and collect
    static constexpr uint32_t kMaxIoChunks = pmd::kMaxIoChunks;
    uint32_t* h_in = nullptr;    // [2][max_n][28]
    uint32_t* h_out = nullptr;   // [max_n][28]
    uint32_t* h_flags = nullptr; // ready[16] | done[16] | error
    uint32_t* hd_in = nullptr;   // device views of the same memory
    uint32_t* hd_out = nullptr;
    uint32_t* hd_flags = nullptr;
    uint32_t* d_io_counters = nullptr;
    uint32_t io_seq = 0;
    uint64_t wait_ticks = 0;        // bound of one device wait on a host flag
    uint64_t clock_khz = 100000;    // device wall clock (ticks per ms)
    int numa_node = -1;             // NUMA node of the device (-1: unknown)
    int32_t stall_chunk = -1;       // input chunk whose ready flag the next submit withholds
    uint32_t stall_sink = 0;        // where that flag goes instead
    std::unique_ptr<PmCopyPool> copy_pool;  // host copies of the input chunks
    uint32_t io_chunks[pm::kMaxLogN + 1] = {};  // chunks per log2 N (0: default)
    bool pending = false;
    uint32_t pending_n = 0, pending_out_chunks = 0;
    std::string last_error;

    void* device_alloc(size_t bytes);
};

int pm_fail(pm_context* ctx, int status, const std::string& why);
int pm_hip(pm_context* ctx, hipError_t e, const char* what);
int pm_check_n(pm_context* ctx, uint32_t n, uint32_t* logn);
int pm_run_plain(pm_context* ctx, const pmd::Job& job, uint32_t outer, hipStream_t stream);
// the tables of the context's backend, and their residue rows per polynomial
const pmd::Tables& pm_tables(const pm_context* ctx);
uint32_t pm_rows(const pm_context* ctx);
// order a plain-path product on `stream` after the previous one (whose
// residues share the workspace), and mark the workspace in use until the
// work enqueued on `stream` so far has run
hipError_t pm_workspace_acquire(pm_context* ctx, hipStream_t stream);
hipError_t pm_workspace_release(pm_context* ctx, hipStream_t stream);
uint32_t pm_default_outer(uint32_t logn);
uint32_t pm_default_batch_outer(uint32_t logn, bool quartic = false);
// tiles per block that fill 1024 threads: 2^(15 - tlog)
uint32_t pm_default_batch_tiles(uint32_t tlog, uint32_t residues_per_thread = 32);
// whether `outer` passes leave a tile of size 2^lo..2^15 for products of 2^logn
bool pm_outer_valid(uint32_t logn, uint32_t outer, uint32_t lo);
uint32_t pm_io_chunks(const pm_context* ctx, uint32_t logn);
