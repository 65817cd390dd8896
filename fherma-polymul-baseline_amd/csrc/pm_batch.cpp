// Batched products of uneven N, one launch per phase, captured once in a
// hipGraph and replayed.
//
// The reference has no batch path (one product per fherma_run call,
// rns/solve.cu:721-853); this builds a work list over many products -- residue
// conversion blocks, optional outer NTT stages, fused convolution tiles
// grouped by tile size and ordered prime-major, CRT blocks -- in which every
// phase is one contiguous run of items, and launches one kernel per phase over
// its run.  PM_BATCH_PERSISTENT runs the same work list in one launch instead:
// one block per compute unit, grid barriers between the phases (no block waits
// on another block's item; an earlier design whose blocks claimed items and
// waited on per-product counters was removed after it hung the GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "pm_batch.h"
#include "pm_device.h"
#include "pm_kernels.h"
#include "pm_runtime.h"
#include "polymul.h"

using pmd::kLimbs;

namespace {

uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// A batch may run as 2-4 staggered parts on as many graph branches, each
// starting when the one before has converted its residues, so a part's
// HBM-bound conversion runs beside the previous part's tile blocks; later
// parts are smaller, since the last one's tiles and CRT run alone at the
// end.  Each part holds at least kPartMinProducts products.  Cumulative
// fractions of the products per part count:
constexpr uint32_t kMaxParts = 4;
constexpr uint32_t kPartMinProducts = 4;
constexpr double kPartCut[kMaxParts + 1][kMaxParts + 1] = {
    {0.0, 1.0}, {0.0, 1.0}, {0.0, 0.55, 1.0}, {0.0, 0.40, 0.75, 1.0}, {0.0, 0.30, 0.57, 0.80, 1.0}};

// Parts when the context leaves the choice open: 3 where some product runs
// 2^15 tiles (a 1024-thread block fills a CU, and only blocks of another
// part's conversion or CRT fill the CUs its tail leaves), else 2 (the second
// part's conversion runs beside the first part's tiles, its tiles beside the
// first part's CRT).  Measured in steady state (50 warm launches, 200 timed):
// 128 x N = 2^15 (2^13 tiles) 2 parts 62.6k against 61.8-62.0k products/s
// for 1, 3 alike, 4 slower; 32 x N = 2^17 (2^15 tiles) 3 parts 11.97-12.05k
// against 11.50-11.59k in one part.
uint32_t auto_parts(const std::vector<pmd::Job>& jobs) {
    for (const pmd::Job& j : jobs)
        if (j.logn - j.outer == 15) return 3;
    return 2;
}

}  // namespace

pm_batch::~pm_batch() {
    if (last_run) (void)hipEventDestroy(last_run);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : allocations) (void)hipFree(p);
}

void* pm_batch::alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    allocations.push_back(p);
    device_bytes += bytes;
    return p;
}

// Work list of jobs [first, first + count) of the batch (their device job
// table at d_jobs + first): items and units uploaded, arguments and phase
// ranges returned.
static int upload_worklist(pm_context* ctx, pm_batch* b, const pmd::Job* d_jobs, uint32_t first, uint32_t count,
                           pmd::BatchArgs& ba, pmd::PhaseRanges& ranges) {
    const uint32_t k = b->tables.k;  // residue rows of the batch's backend
    std::vector<pmd::Item> items;
    std::vector<pmd::Unit> units;
    pm_build_worklist(k, b->jobs.data() + first, count, true, items, units, ctx->batch_tiles, ctx->batch_rpt);
    if (!pm_phase_ranges(items, ranges)) return pm_fail(ctx, PM_EINVAL, "work list phases are not contiguous");
    if (units.empty()) units.push_back({0, 0});
    auto* d_items = static_cast<pmd::Item*>(b->alloc(items.size() * sizeof(pmd::Item)));
    auto* d_units = static_cast<pmd::Unit*>(b->alloc(units.size() * sizeof(pmd::Unit)));
    if (!d_items || !d_units) return pm_fail(ctx, PM_ENOMEM, "batch table allocation failed");
    hipError_t e = hipMemcpy(d_items, items.data(), items.size() * sizeof(pmd::Item), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_units, units.data(), units.size() * sizeof(pmd::Unit), hipMemcpyHostToDevice);
    if (e != hipSuccess) return pm_hip(ctx, e, "upload batch tables");
    ba = pmd::BatchArgs{d_jobs + first, d_items, d_units, uint32_t(items.size())};
    return PM_OK;
}

// The persistent launch (one graph node after the barrier reset): the whole
// work list, built for its 512-thread tile blocks.
static int build_persistent(pm_context* ctx, pm_batch* b, const pmd::Job* d_jobs) {
    const uint32_t k = b->tables.k;
    uint32_t tpb[16] = {};
    for (uint32_t t = 5; t <= uint32_t(pmd::kPersistentMaxTileLog); ++t) tpb[t] = (pmd::kPersistentThreads * 32u) >> t;
    std::vector<pmd::Item> items;
    std::vector<pmd::Unit> units;
    pm_build_worklist(k, b->jobs.data(), b->count, true, items, units, tpb);
    if (!pm_phase_ranges(items, b->ranges)) return pm_fail(ctx, PM_EINVAL, "work list phases are not contiguous");
    if (units.empty()) units.push_back({0, 0});
    auto* d_items = static_cast<pmd::Item*>(b->alloc(items.size() * sizeof(pmd::Item)));
    auto* d_units = static_cast<pmd::Unit*>(b->alloc(units.size() * sizeof(pmd::Unit)));
    b->sync = static_cast<uint32_t*>(b->alloc(4 * sizeof(uint32_t)));
    if (!d_items || !d_units || !b->sync) return pm_fail(ctx, PM_ENOMEM, "batch table allocation failed");
    hipError_t e = hipMemcpy(d_items, items.data(), items.size() * sizeof(pmd::Item), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_units, units.data(), units.size() * sizeof(pmd::Unit), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->sync, 0, 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = pmd::persistent_blocks(b->tables.cus, &b->persistent_blocks);
    if (e != hipSuccess) return pm_hip(ctx, e, "persistent batch set-up");
    if (!b->persistent_blocks) return pm_fail(ctx, PM_EUNSUPPORTED, "persistent batch: a block does not fit a CU");
    b->args = pmd::BatchArgs{d_jobs, d_items, d_units, uint32_t(items.size())};
    b->n_items = b->args.n_items;
    b->parts = 1;
    const pmd::GridSync gs{b->sync, b->sync + pmd::kPersistentBarriers, ctx->wait_ticks};
    hipStream_t stream = nullptr;
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const hipError_t le = pmd::launch_persistent(b->tables, b->args, b->ranges, gs, b->persistent_blocks, stream);
        const hipError_t ce = hipStreamEndCapture(stream, &b->graph);
        e = le != hipSuccess ? le : ce;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->last_run, hipEventDisableTiming);
    if (stream) (void)hipStreamDestroy(stream);
    if (e != hipSuccess) return pm_hip(ctx, e, "capture persistent batch graph");
    return PM_OK;
}

int pm_batch_build(pm_context* ctx, pm_batch* b, const uint32_t* ns, const uint32_t* const* d_a,
                   const uint32_t* const* d_b, uint32_t* const* d_c) {
    b->tables = pm_tables(ctx);  // the backend at creation stays the batch's
    b->crt_waves = ctx->batch_crt_waves;
    if (b->persistent && b->tables.quartic)
        return pm_fail(ctx, PM_EUNSUPPORTED, "persistent batch: general backend only");
    const uint32_t k = b->tables.k, count = b->count;
    b->logn.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t n = ns[i];
        if (n < 32 || (n & (n - 1))) return pm_fail(ctx, PM_EUNSUPPORTED, "batch sizes must be powers of two >= 32");
        b->logn[i] = uint32_t(__builtin_ctz(n));
        if (b->logn[i] > ctx->max_logn) return pm_fail(ctx, PM_EUNSUPPORTED, "batch N exceeds the context's max_n");
    }
    // device memory: inputs/outputs (unless the caller owns them) and residues
    std::vector<pmd::Job> jobs(count);
    for (uint32_t i = 0; i < count; ++i) {
        const size_t n = size_t(1) << b->logn[i];
        pmd::Job& j = jobs[i];
        j.logn = b->logn[i];
        j.outer = ctx->batch_outer[b->tables.quartic ? PM_BACKEND_QUARTIC : PM_BACKEND_GENERAL][b->logn[i]];
        if (b->persistent) {
            // tiles of at most 2^14 (a 512-thread block), outer stages fused
            const uint32_t least = j.logn > uint32_t(pmd::kPersistentMaxTileLog) ? j.logn - pmd::kPersistentMaxTileLog : 0u;
            if (j.outer < least) j.outer = least;
            if (j.outer && !pmd::outer_fusable(j.logn, j.outer))
                return pm_fail(ctx, PM_EUNSUPPORTED, "persistent batch: N up to 2^16");
        }
        j.fused = !b->separate_outer && !b->tables.quartic && pmd::outer_fusable(j.logn, j.outer) ? 1u : 0u;
        const bool caller_io = d_a && d_b && d_c;
        if (caller_io) {
            j.a = d_a[i];
            j.b = d_b[i];
            j.c = d_c[i];
        } else {
            auto* io = static_cast<uint32_t*>(b->alloc(3 * n * kLimbs * sizeof(uint32_t)));
            if (!io) return pm_fail(ctx, PM_ENOMEM, "batch input/output allocation failed");
            j.a = io;
            j.b = io + n * kLimbs;
            j.c = io + 2 * n * kLimbs;
        }
        auto* res = static_cast<uint32_t*>(b->alloc(2 * size_t(k) * n * sizeof(uint32_t)));
        if (!res) return pm_fail(ctx, PM_ENOMEM, "batch residue allocation failed");
        j.ra = res;
        j.rb = res + size_t(k) * n;
    }
    auto* d_jobs = static_cast<pmd::Job*>(b->alloc(jobs.size() * sizeof(pmd::Job)));
    if (!d_jobs) return pm_fail(ctx, PM_ENOMEM, "batch table allocation failed");
    hipError_t e = hipMemcpy(d_jobs, jobs.data(), jobs.size() * sizeof(pmd::Job), hipMemcpyHostToDevice);
    if (e != hipSuccess) return pm_hip(ctx, e, "upload batch tables");
    b->jobs = jobs;
    if (b->persistent) return build_persistent(ctx, b, d_jobs);

    // the whole batch as one work list (per-phase timing; graphs of one part)
    int st = upload_worklist(ctx, b, d_jobs, 0, count, b->args, b->ranges);
    if (st != PM_OK) return st;
    b->n_items = b->args.n_items;

    // capture the phase launches once; runs replay the graph, in one part or
    // in staggered parts on as many branches (kPartCut)
    uint32_t parts = ctx->batch_parts ? ctx->batch_parts : auto_parts(b->jobs);
    while (parts > 1 && count < parts * kPartMinProducts) --parts;
    b->parts = parts;
    std::vector<pmd::BatchArgs> part_args(parts);
    std::vector<pmd::PhaseRanges> part_ranges(parts);
    if (parts > 1) {
        for (uint32_t q = 0; q < parts; ++q) {
            const uint32_t lo = uint32_t(count * kPartCut[parts][q]);
            const uint32_t hi = q + 1 == parts ? count : uint32_t(count * kPartCut[parts][q + 1]);
            st = upload_worklist(ctx, b, d_jobs, lo, hi - lo, part_args[q], part_ranges[q]);
            if (st != PM_OK) return st;
        }
    } else {
        part_args[0] = b->args;
        part_ranges[0] = b->ranges;
    }
    std::vector<hipStream_t> streams(parts, nullptr);
    std::vector<hipEvent_t> first_done(parts, nullptr), part_done(parts, nullptr);
    for (uint32_t q = 0; q < parts && e == hipSuccess; ++q) {
        e = hipStreamCreateWithFlags(&streams[q], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&first_done[q], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&part_done[q], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipStreamBeginCapture(streams[0], hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = hipSuccess;
        for (uint32_t q = 0; q < parts && le == hipSuccess; ++q) {
            if (q > 0) le = hipStreamWaitEvent(streams[q], first_done[q - 1], 0);  // fork
            const auto phases = pmd::phase_list(part_ranges[q], b->scalar_conversion, b->scalar_crt, b->crt_waves);
            for (size_t i = 0; i < phases.size() && le == hipSuccess; ++i) {
                le = pmd::launch_phase(b->tables, part_args[q], phases[i].kind, phases[i].first, phases[i].count,
                                       streams[q]);
                if (i == 0 && le == hipSuccess) le = hipEventRecord(first_done[q], streams[q]);
            }
            if (q > 0 && le == hipSuccess) le = hipEventRecord(part_done[q], streams[q]);
        }
        for (uint32_t q = 1; q < parts && le == hipSuccess; ++q) le = hipStreamWaitEvent(streams[0], part_done[q], 0);
        hipError_t ce = hipStreamEndCapture(streams[0], &b->graph);
        e = le != hipSuccess ? le : ce;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->last_run, hipEventDisableTiming);
    for (uint32_t q = 0; q < parts; ++q) {
        if (first_done[q]) (void)hipEventDestroy(first_done[q]);
        if (part_done[q]) (void)hipEventDestroy(part_done[q]);
        if (streams[q]) (void)hipStreamDestroy(streams[q]);
    }
    if (e != hipSuccess) return pm_hip(ctx, e, "capture batch graph");
    return PM_OK;
}

bool pm_phase_ranges(const std::vector<pmd::Item>& items, pmd::PhaseRanges& r) {
    r = pmd::PhaseRanges{};
    // each phase (and each tile class) must be one contiguous run of items
    auto add = [](uint32_t i, uint32_t& first, uint32_t& count) {
        if (count == 0) first = i;
        else if (first + count != i) return false;
        ++count;
        return true;
    };
    for (uint32_t i = 0; i < items.size(); ++i) {
        const pmd::Item& it = items[i];
        bool ok = true;
        switch (it.type) {
            case pmd::kItemConv: ok = add(i, r.conv_first, r.conv_count); break;
            case pmd::kItemOuterFwd: ok = add(i, r.fwd_first, r.fwd_count); break;
            case pmd::kItemTile: {
                const uint32_t tlog = it.a0 & 0xFu, units = it.a0 >> 8;
                const uint32_t rpt16 = (it.a0 & pmd::kItemRpt16) ? 1u : 0u;
                const bool first_of_class = r.tile_count[tlog] == 0;
                // a block holds at most 1024 threads: 2^15 residues at 32 per thread, 2^14 at 16
                ok = tlog >= 5 && tlog <= 15 && units >= 1 && (units << tlog) <= (rpt16 ? 16384u : 32768u) &&
                     (!rpt16 || (tlog >= uint32_t(pmd::kRpt16MinLog) && tlog <= uint32_t(pmd::kRpt16MaxLog))) &&
                     (first_of_class || r.tile_rpt16[tlog] == rpt16) && add(i, r.tile_first[tlog], r.tile_count[tlog]);
                if (ok) r.tile_rpt16[tlog] = rpt16;
                if (ok && units > r.tile_units[tlog]) r.tile_units[tlog] = units;
                break;
            }
            case pmd::kItemOuterInv: ok = add(i, r.inv_first, r.inv_count); break;
            case pmd::kItemCrt: ok = add(i, r.crt_first, r.crt_count); break;
            default: ok = false;
        }
        if (!ok) return false;
    }
    return true;
}

void pm_build_worklist(uint32_t k, const pmd::Job* jobs, uint32_t count, bool largest_first,
                       std::vector<pmd::Item>& items, std::vector<pmd::Unit>& units,
                       const uint32_t* tiles_per_block, const uint32_t* residues_per_thread) {
    // Order: conversions, outer forward stages, fused tiles by size class,
    // outer inverse stages, CRT blocks.  Every dependency of an item points to
    // an earlier item, so claiming items in order cannot deadlock.
    std::vector<uint32_t> order(count);
    std::iota(order.begin(), order.end(), 0u);
    if (largest_first)
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t x, uint32_t y) { return jobs[x].logn > jobs[y].logn; });
    items.clear();
    units.clear();
    for (uint32_t j : order) {
        const uint32_t blocks = ceil_div(1u << jobs[j].logn, pmd::kConvItemCoefs);
        for (uint32_t blk = 0; blk < blocks; ++blk)
            for (uint32_t poly = 0; poly < 2; ++poly) items.push_back({pmd::kItemConv, j, poly, blk});
    }
    for (uint32_t j : order)
        if (jobs[j].outer && !jobs[j].fused)
            for (uint32_t x = 0; x < 2 * k; ++x) items.push_back({pmd::kItemOuterFwd, j, x, 0});
    for (uint32_t tlog = 15; tlog >= 5; --tlog) {
        // tiles per block: by default 1024 threads x 32 residues / tile size
        const uint32_t per_item = tiles_per_block ? tiles_per_block[tlog] : pm_default_batch_tiles(tlog);
        // prime-major: blocks that run at the same time (and the blocks one
        // XCD receives round-robin) share a prime, so its twiddle tables stay
        // in that XCD's L2 instead of every block streaming its own
        std::vector<pmd::Unit> cls;
        for (uint32_t prime = 0; prime < k; ++prime)
            for (uint32_t j : order) {
                if (jobs[j].logn - jobs[j].outer != tlog) continue;
                for (uint32_t s = 0; s < (1u << jobs[j].outer); ++s) cls.push_back({j, (prime << 2) | s});
            }
        const uint32_t rpt16 = residues_per_thread && residues_per_thread[tlog] == 16 ? pmd::kItemRpt16 : 0u;
        for (size_t first = 0; first < cls.size(); first += per_item) {
            const uint32_t n_units = uint32_t(std::min<size_t>(per_item, cls.size() - first));
            items.push_back({pmd::kItemTile, 0, tlog | rpt16 | (n_units << 8), uint32_t(units.size())});
            units.insert(units.end(), cls.begin() + first, cls.begin() + first + n_units);
        }
    }
    for (uint32_t j : order)
        if (jobs[j].outer && !jobs[j].fused)
            for (uint32_t prime = 0; prime < k; ++prime) items.push_back({pmd::kItemOuterInv, j, prime, 0});
    for (uint32_t j : order) {
        const uint32_t blocks = ceil_div(1u << jobs[j].logn, pmd::kBlock);
        for (uint32_t blk = 0; blk < blocks; ++blk) items.push_back({pmd::kItemCrt, j, blk, 0});
    }
}

extern "C" {

int pm_batch_create(pm_context* ctx, uint32_t count, const uint32_t* ns, const uint32_t* const* d_a,
                    const uint32_t* const* d_b, uint32_t* const* d_c, uint32_t flags, pm_batch** out) {
    if (!ctx || !ns || !out || count == 0) return PM_EINVAL;
    constexpr uint32_t kPhasedFlags = PM_BATCH_PHASED | PM_BATCH_SCALAR_CONVERSION | PM_BATCH_SCALAR_CRT |
                                      PM_BATCH_SEPARATE_OUTER;
    if (flags & ~uint32_t(PM_BATCH_PERSISTENT | kPhasedFlags)) return PM_EINVAL;
    if ((flags & PM_BATCH_PERSISTENT) && (flags & kPhasedFlags)) return PM_EINVAL;
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return pm_hip(ctx, e, "set device");
    pm_batch* b = new (std::nothrow) pm_batch();
    if (!b) return pm_fail(ctx, PM_ENOMEM, "host allocation failed");
    b->ctx = ctx;
    b->count = count;
    b->persistent = (flags & PM_BATCH_PERSISTENT) != 0;
    b->scalar_conversion = (flags & PM_BATCH_SCALAR_CONVERSION) != 0;
    b->scalar_crt = (flags & PM_BATCH_SCALAR_CRT) != 0;
    // the scalar conversion and CRT keep consecutive coefficients: outer stages as passes of their own
    b->separate_outer = (flags & (PM_BATCH_SEPARATE_OUTER | PM_BATCH_SCALAR_CONVERSION | PM_BATCH_SCALAR_CRT)) != 0;
    const int st = pm_batch_build(ctx, b, ns, d_a, d_b, d_c);
    if (st != PM_OK) {
        delete b;
        return st;
    }
    *out = b;
    return PM_OK;
}

void pm_batch_destroy(pm_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    // the last replay may be on a caller's stream: wait for it by its event
    if (b->last_run) (void)hipEventSynchronize(b->last_run);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

int pm_batch_io(pm_batch* b, uint32_t i, uint32_t** d_a, uint32_t** d_b, uint32_t** d_c) {
    if (!b || i >= b->count) return PM_EINVAL;
    if (d_a) *d_a = const_cast<uint32_t*>(b->jobs[i].a);
    if (d_b) *d_b = const_cast<uint32_t*>(b->jobs[i].b);
    if (d_c) *d_c = b->jobs[i].c;
    return PM_OK;
}

int pm_batch_upload(pm_batch* b, uint32_t i, const uint32_t* a, const uint32_t* bb) {
    if (!b || i >= b->count || !a || !bb) return PM_EINVAL;
    const size_t bytes = (size_t(1) << b->logn[i]) * kLimbs * sizeof(uint32_t);
    // a replay runs on a non-blocking stream, which these copies do not order
    // behind: wait for the last one so its inputs are not overwritten mid-run
    hipError_t e = hipEventSynchronize(b->last_run);
    if (e == hipSuccess) e = hipMemcpy(const_cast<uint32_t*>(b->jobs[i].a), a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(const_cast<uint32_t*>(b->jobs[i].b), bb, bytes, hipMemcpyHostToDevice);
    return pm_hip(b->ctx, e, "batch upload");
}

int pm_batch_download(pm_batch* b, uint32_t i, uint32_t* c) {
    if (!b || i >= b->count || !c) return PM_EINVAL;
    const size_t bytes = (size_t(1) << b->logn[i]) * kLimbs * sizeof(uint32_t);
    hipError_t e = hipEventSynchronize(b->last_run);  // the result of the last replay is whole
    if (e == hipSuccess) e = hipMemcpy(c, b->jobs[i].c, bytes, hipMemcpyDeviceToHost);
    return pm_hip(b->ctx, e, "batch download");
}

int pm_batch_run(pm_batch* b, void* hip_stream) {
    if (!b) return PM_EINVAL;
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : b->ctx->stream;
    hipError_t e = hipGraphLaunch(b->exec, stream);
    if (e == hipSuccess) e = hipEventRecord(b->last_run, stream);
    return pm_hip(b->ctx, e, "launch batch graph");
}

int pm_batch_wait(pm_batch* b, void* hip_stream) {
    if (!b) return PM_EINVAL;
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : b->ctx->stream;
    // phased launches have no in-kernel waits: a device error is all there is
    // to report; the persistent launch also reports a barrier wait that gave up
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return pm_hip(b->ctx, e, "batch wait");
    if (b->persistent) {
        uint32_t err = 0;
        e = hipMemcpy(&err, b->sync + pmd::kPersistentBarriers, sizeof(err), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return pm_hip(b->ctx, e, "batch wait");
        if (err) return pm_fail(b->ctx, PM_ETIMEOUT, "persistent batch: a grid barrier wait ran out (blocks not resident)");
    }
    return PM_OK;
}

int pm_plan_worklist(uint32_t k, uint32_t count, const uint32_t* ns, uint32_t* n_items, uint32_t* items,
                     uint32_t* n_units, uint32_t* units) {
    if (!ns || !n_items || !n_units || k == 0 || k > 64) return PM_EINVAL;
    std::vector<pmd::Job> jobs(count);
    for (uint32_t i = 0; i < count; ++i) {
        if (ns[i] < 32 || (ns[i] & (ns[i] - 1)) || ns[i] > (1u << 17)) return PM_EUNSUPPORTED;
        jobs[i].logn = uint32_t(__builtin_ctz(ns[i]));
        jobs[i].outer = pm_default_batch_outer(jobs[i].logn);
        jobs[i].fused = pmd::outer_fusable(jobs[i].logn, jobs[i].outer) ? 1u : 0u;
    }
    std::vector<pmd::Item> it;
    std::vector<pmd::Unit> un;
    pm_build_worklist(k, jobs.data(), count, true, it, un);
    pmd::PhaseRanges r;
    if (!pm_phase_ranges(it, r)) return PM_EINVAL;
    const uint32_t cap_items = *n_items, cap_units = *n_units;
    *n_items = uint32_t(it.size());
    *n_units = uint32_t(un.size());
    if (items && cap_items >= it.size()) std::memcpy(items, it.data(), it.size() * sizeof(pmd::Item));
    if (units && cap_units >= un.size()) std::memcpy(units, un.data(), un.size() * sizeof(pmd::Unit));
    return PM_OK;
}

int pm_batch_phase_times(pm_batch* b, void* hip_stream, uint32_t max, uint32_t* kinds, double* ms,
                         uint32_t* count) {
    if (!b || !count) return PM_EINVAL;
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : b->ctx->stream;
    const std::vector<pmd::Phase> phases =
        pmd::phase_list(b->ranges, b->scalar_conversion, b->scalar_crt, b->crt_waves);
    *count = uint32_t(phases.size());
    if (phases.size() > max || !kinds || !ms) return PM_OK;
    std::vector<hipEvent_t> ev(phases.size() + 1, nullptr);
    hipError_t e = hipSuccess;
    for (auto& x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    for (size_t i = 0; i < phases.size() && e == hipSuccess; ++i) {
        e = hipEventRecord(ev[i], stream);
        if (e == hipSuccess)
            e = pmd::launch_phase(b->tables, b->args, phases[i].kind, phases[i].first, phases[i].count, stream);
    }
    if (e == hipSuccess) e = hipEventRecord(ev[phases.size()], stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[phases.size()]);
    for (size_t i = 0; i < phases.size() && e == hipSuccess; ++i) {
        float t = 0;
        e = hipEventElapsedTime(&t, ev[i], ev[i + 1]);
        kinds[i] = phases[i].kind;
        ms[i] = t;
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    return pm_hip(b->ctx, e, "time batch phases");
}

int pm_batch_info(const pm_batch* b, uint32_t* items, uint32_t* parts, uint64_t* device_bytes) {
    if (!b) return PM_EINVAL;
    if (items) *items = b->n_items;
    if (parts) *parts = b->parts;
    if (device_bytes) *device_bytes = b->device_bytes;
    return PM_OK;
}

}  // extern "C"
