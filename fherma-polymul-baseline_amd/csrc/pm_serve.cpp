// Continuous-batching request loop (no counterpart in the reference, which
// serves one product per fherma_run call, rns/solve.cu:721-853).
//
// Requests of different N under one q are submitted at any time.  A native
// scheduler thread admits them in arrival order while their device memory
// fits an HBM budget (first-fit over one arena), groups whatever has been
// admitted into the next batch launch (the phased kernels of pm_batch.cpp
// fed with a freshly written work list), and keeps up to `max_inflight`
// launches running at once on separate streams, so requests that arrive
// while others are on the GPU start as soon as memory and a launch slot free
// up, and their work overlaps on the device.  Tables stay resident in the
// context.  Each result depends only on its own inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pm_batch.h"
#include "pm_device.h"
#include "pm_kernels.h"
#include "pm_runtime.h"
#include "polymul.h"

using Clock = std::chrono::steady_clock;
using pmd::kLimbs;

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size) with coalescing frees.
class Arena {
public:
    void reset(size_t size) {
        free_.clear();
        free_[0] = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, start = it->first + bytes;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t peak() const { return peak_; }
    size_t used() const { return used_; }

private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0, peak_ = 0;
};

struct Request {
    uint64_t id = 0;
    uint32_t logn = 0;
    const uint32_t* a = nullptr;
    const uint32_t* b = nullptr;
    uint32_t* c = nullptr;
    size_t offset = 0, bytes = 0;  // arena slice: a, b, c, ra, rb
    Clock::time_point submitted, finished;
    // device addresses of a and b when both are page-locked host memory (the
    // launch's gather kernel reads them), of c when it is (the scatter kernel
    // writes it); null otherwise
    const uint32_t* a_dev = nullptr;
    const uint32_t* b_dev = nullptr;
    uint32_t* c_dev = nullptr;
    bool ab_pinned = false;
    bool c_pinned = false;
    bool done = false;
    int status = PM_OK;
};

struct Launch {
    hipStream_t stream = nullptr;
    hipEvent_t finished = nullptr;
    pmd::Job* d_jobs = nullptr;
    pmd::Item* d_items = nullptr;
    pmd::Unit* d_units = nullptr;
    // pinned staging of the work list (async upload)
    pmd::Job* h_jobs = nullptr;
    pmd::Item* h_items = nullptr;
    pmd::Unit* h_units = nullptr;
    // span tables of the gather (inputs and work list in) and scatter (page-
    // locked results out) kernels, read by those kernels from host memory
    pmd::CopySeg* h_in_segs = nullptr;
    pmd::CopySeg* h_out_segs = nullptr;
    std::vector<Request*> requests;
    bool busy = false;
};

}  // namespace

struct pm_server {
    pm_context* ctx = nullptr;
    pmd::Tables tables{};  // the context's backend when the server was created
    uint32_t crt_waves = 0;  // waves per block of the MFMA CRT (the context's at creation)
    uint32_t max_batch = 0, max_inflight = 0;
    size_t max_items = 0, max_units = 0;
    uint64_t budget = 0;
    uint8_t* arena_base = nullptr;
    Arena arena;
    std::vector<Launch> launches;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;                   // submitted, not admitted
    std::deque<Request*> admitted;                  // memory assigned, not launched
    std::unordered_map<uint64_t, Request*> all;     // until waited for
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    uint64_t completed = 0, launches_done = 0, max_concurrent = 0, failed = 0;
    std::string error;  // reason of the last failed launch (guarded by mu)

    size_t request_bytes(uint32_t logn) const {
        const size_t n = size_t(1) << logn;
        return align_up(3 * n * kLimbs * 4) + align_up(2 * size_t(tables.k) * n * 4);
    }
    // Under a backlog the page-locked inputs of requests from 2^13
    // coefficients (1.8 MB a polynomial) go in by DMA on the copy engines
    // instead of the gather kernel: a kernel reading host memory and one
    // writing it share about 63 GB/s of PCIe between them on MI355X, a
    // copy-engine read beside the scatter kernel's writes about 85 GB/s.  A
    // DMA call can now and then hold this thread in the runtime for some
    // milliseconds, so a server that keeps up uses the kernel alone.
    static constexpr uint32_t kDmaMinLogN = 13;
    void run();
    bool start_launch(Launch& L, std::vector<Request*>& reqs, bool dma_inputs, std::string& why);
    void finish_launch(Launch& L, hipError_t status);
};

bool pm_server::start_launch(Launch& L, std::vector<Request*>& reqs, bool dma_inputs, std::string& why) {
    const uint32_t k = tables.k;
    std::vector<pmd::Item> items;
    std::vector<pmd::Unit> units;
    const uint32_t count = uint32_t(reqs.size());
    for (uint32_t j = 0; j < count; ++j) {
        Request* r = reqs[j];
        const size_t n = size_t(1) << r->logn;
        uint8_t* base = arena_base + r->offset;
        pmd::Job& job = L.h_jobs[j];
        job.logn = r->logn;
        job.outer = ctx->batch_outer[tables.quartic ? PM_BACKEND_QUARTIC : PM_BACKEND_GENERAL][r->logn];
        job.fused = !tables.quartic && pmd::outer_fusable(job.logn, job.outer) ? 1u : 0u;
        job.a = reinterpret_cast<const uint32_t*>(base);
        job.b = job.a + n * kLimbs;
        job.c = reinterpret_cast<uint32_t*>(base) + 2 * n * kLimbs;
        job.ra = reinterpret_cast<uint32_t*>(base + align_up(3 * n * kLimbs * 4));
        job.rb = job.ra + size_t(k) * n;
    }
    pm_build_worklist(k, L.h_jobs, count, false, items, units, ctx->batch_tiles, ctx->batch_rpt);
    if (items.size() > max_items || units.size() > max_units) {
        why = "work list exceeds the launch slot";
        return false;
    }
    pmd::PhaseRanges ranges;
    if (!pm_phase_ranges(items, ranges)) {  // checked before anything is enqueued
        why = "work list phases are not contiguous";
        return false;
    }
    std::memcpy(L.h_items, items.data(), items.size() * sizeof(pmd::Item));
    std::memcpy(L.h_units, units.data(), std::max<size_t>(units.size(), 1) * sizeof(pmd::Unit));
    // Inputs and the work list go in through one gather kernel, page-locked
    // results come back through one scatter kernel: with the conversion, tile
    // and CRT launches the whole batch is kernels on one queue, and this thread
    // makes no call that waits on the device (a copy-engine transfer behind a
    // kernel can hold the enqueueing thread in the runtime for ~10 ms).  Only
    // pageable inputs take a runtime copy.
    hipError_t e = hipSuccess;
    uint32_t n_in = 0;
    uint64_t max_words = 0;
    auto span = [&max_words](pmd::CopySeg* segs, uint32_t& n, const void* src, void* dst, size_t bytes) {
        segs[n++] = pmd::CopySeg{static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), bytes / 4};
        max_words = std::max<uint64_t>(max_words, bytes / 4);
    };
    for (uint32_t j = 0; j < count && e == hipSuccess; ++j) {
        const size_t bytes = (size_t(1) << reqs[j]->logn) * kLimbs * 4;
        uint32_t* da = const_cast<uint32_t*>(L.h_jobs[j].a);
        uint32_t* db = const_cast<uint32_t*>(L.h_jobs[j].b);
        if (reqs[j]->ab_pinned && !(dma_inputs && reqs[j]->logn >= kDmaMinLogN)) {
            span(L.h_in_segs, n_in, reqs[j]->a_dev, da, bytes);
            span(L.h_in_segs, n_in, reqs[j]->b_dev, db, bytes);
        } else {  // pageable inputs, or page-locked ones by DMA under a backlog
            e = hipMemcpyAsync(da, reqs[j]->a, bytes, hipMemcpyHostToDevice, L.stream);
            if (e == hipSuccess) e = hipMemcpyAsync(db, reqs[j]->b, bytes, hipMemcpyHostToDevice, L.stream);
        }
    }
    span(L.h_in_segs, n_in, L.h_jobs, L.d_jobs, count * sizeof(pmd::Job));
    span(L.h_in_segs, n_in, L.h_items, L.d_items, items.size() * sizeof(pmd::Item));
    span(L.h_in_segs, n_in, L.h_units, L.d_units, std::max<size_t>(units.size(), 1) * sizeof(pmd::Unit));
    if (e == hipSuccess) e = pmd::launch_copy_segments(L.h_in_segs, n_in, max_words, false, L.stream);
    pmd::BatchArgs ba{};
    ba.jobs = L.d_jobs;
    ba.items = L.d_items;
    ba.units = L.d_units;
    ba.n_items = uint32_t(items.size());
    if (e == hipSuccess) e = pmd::launch_phased(tables, ba, ranges, false, false, crt_waves, L.stream);
    // results into page-locked caller buffers go back behind the kernels on
    // the same stream; other buffers are copied on completion
    uint32_t n_out = 0;
    max_words = 0;
    for (uint32_t j = 0; j < count; ++j)
        if (reqs[j]->c_pinned)
            span(L.h_out_segs, n_out, L.h_jobs[j].c, reqs[j]->c_dev, (size_t(1) << reqs[j]->logn) * kLimbs * 4);
    if (e == hipSuccess) e = pmd::launch_copy_segments(L.h_out_segs, n_out, max_words, true, L.stream);
    if (e == hipSuccess) e = hipEventRecord(L.finished, L.stream);
    if (e != hipSuccess) {
        // what was enqueued still uses the requests' memory: let it drain
        // before the caller releases that memory
        (void)hipStreamSynchronize(L.stream);
        why = std::string("launch: ") + hipGetErrorString(e);
        return false;
    }
    L.requests = reqs;
    L.busy = true;
    return true;
}

void pm_server::finish_launch(Launch& L, hipError_t status) {
    // called once L.finished is no longer pending: with status hipSuccess
    // every kernel and copy of the launch is done; otherwise the device
    // reported an error and the launch's requests fail with it
    hipError_t e = status;
    // results to pageable callers' buffers
    for (size_t j = 0; j < L.requests.size() && e == hipSuccess; ++j) {
        if (L.requests[j]->c_pinned) continue;
        const size_t bytes = (size_t(1) << L.requests[j]->logn) * kLimbs * 4;
        e = hipMemcpy(L.requests[j]->c, L.h_jobs[j].c, bytes, hipMemcpyDeviceToHost);
    }
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lock(mu);
    if (e != hipSuccess) error = std::string("device: ") + hipGetErrorString(e);
    for (Request* r : L.requests) {
        r->finished = now;
        r->status = e != hipSuccess ? PM_EDEVICE : PM_OK;
        if (r->status != PM_OK) ++failed;
        r->done = true;
        arena.release(r->offset, r->bytes);
        ++completed;
    }
    L.requests.clear();
    L.busy = false;
    ++launches_done;
    cv_done.notify_all();
    cv_work.notify_all();
}

void pm_server::run() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::vector<Request*> batch;
        Launch* slot = nullptr;
        bool backlog = false;  // requests wait beyond this launch: the server is behind
        {
            std::unique_lock<std::mutex> lock(mu);
            // admission: arrival order, while the head fits the budget
            while (!pending.empty()) {
                Request* r = pending.front();
                size_t off = 0;
                if (!arena.alloc(r->bytes, &off)) break;
                r->offset = off;
                admitted.push_back(r);
                pending.pop_front();
            }
            bool any_busy = false;
            for (Launch& L : launches) {
                if (L.busy) any_busy = true;
                else if (!slot) slot = &L;
            }
            if (slot && !admitted.empty()) {
                while (!admitted.empty() && batch.size() < max_batch) {
                    batch.push_back(admitted.front());
                    admitted.pop_front();
                }
                backlog = !admitted.empty() || !pending.empty();
            } else if (!any_busy) {
                if (stop && pending.empty() && admitted.empty()) return;
                cv_work.wait_for(lock, std::chrono::milliseconds(2));
                continue;
            }
        }
        if (!batch.empty()) {
            std::string why;
            if (!start_launch(*slot, batch, backlog, why)) {
                std::lock_guard<std::mutex> lock(mu);
                error = why;
                for (Request* r : batch) {
                    r->status = PM_EDEVICE;
                    r->done = true;
                    r->finished = Clock::now();
                    arena.release(r->offset, r->bytes);
                    ++failed;
                }
                cv_done.notify_all();
            } else {
                uint64_t busy = 0;
                for (Launch& L : launches) busy += L.busy;
                std::lock_guard<std::mutex> lock(mu);
                max_concurrent = std::max(max_concurrent, busy);
            }
            continue;  // admit / launch more before polling completions
        }
        // poll completions; yield briefly when nothing finished
        bool progressed = false;
        for (Launch& L : launches) {
            if (!L.busy) continue;
            const hipError_t q = hipEventQuery(L.finished);
            if (q != hipErrorNotReady) {
                finish_launch(L, q);
                progressed = true;
            }
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

extern "C" {

int pm_server_create(pm_context* ctx, uint64_t hbm_budget_bytes, uint32_t max_batch, uint32_t max_inflight,
                     pm_server** out) {
    if (!ctx || !out || max_batch == 0 || max_inflight == 0) return PM_EINVAL;
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return pm_hip(ctx, e, "set device");
    pm_server* s = new (std::nothrow) pm_server();
    if (!s) return pm_fail(ctx, PM_ENOMEM, "host allocation failed");
    s->ctx = ctx;
    s->tables = pm_tables(ctx);
    s->crt_waves = ctx->batch_crt_waves;
    s->budget = hbm_budget_bytes;
    s->max_batch = max_batch;
    s->max_inflight = max_inflight;
    if (s->request_bytes(ctx->max_logn) > hbm_budget_bytes) {
        delete s;
        return pm_fail(ctx, PM_EBUDGET, "HBM budget below one product of max_n");
    }
    e = hipMalloc(reinterpret_cast<void**>(&s->arena_base), hbm_budget_bytes);
    if (e != hipSuccess) {
        delete s;
        return pm_hip(ctx, e, "allocate serving arena");
    }
    s->arena.reset(hbm_budget_bytes);
    const uint32_t k = s->tables.k, n = 1u << ctx->max_logn;
    const size_t per_req_items = 3 * ((n + pmd::kBlock - 1) / pmd::kBlock) + 3 * k + size_t(k) * 4 + 1;
    s->max_items = per_req_items * max_batch + 64;
    s->max_units = size_t(k) * 4 * max_batch + 1;
    s->launches.resize(max_inflight);
    for (Launch& L : s->launches) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&L.finished, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&L.d_jobs), max_batch * sizeof(pmd::Job));
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&L.d_items), s->max_items * sizeof(pmd::Item));
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&L.d_units), s->max_units * sizeof(pmd::Unit));
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&L.h_jobs), max_batch * sizeof(pmd::Job));
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&L.h_items), s->max_items * sizeof(pmd::Item));
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&L.h_units), s->max_units * sizeof(pmd::Unit));
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&L.h_in_segs), (2 * size_t(max_batch) + 3) * sizeof(pmd::CopySeg));
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&L.h_out_segs), size_t(max_batch) * sizeof(pmd::CopySeg));
    }
    if (e != hipSuccess) {
        const int st = pm_hip(ctx, e, "serving launch slots");
        pm_server_destroy(s);
        return st;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return PM_OK;
}

int pm_server_submit(pm_server* s, uint32_t n, const uint32_t* a, const uint32_t* b, uint32_t* c, uint64_t* id) {
    if (!s || !a || !b || !c || !id) return PM_EINVAL;
    if (n < 32 || (n & (n - 1))) return pm_fail(s->ctx, PM_EUNSUPPORTED, "served sizes are powers of two >= 32");
    const uint32_t logn = uint32_t(__builtin_ctz(n));
    if (logn > s->ctx->max_logn) return pm_fail(s->ctx, PM_EUNSUPPORTED, "N exceeds the context's max_n");
    auto* r = new (std::nothrow) Request();
    if (!r) return PM_ENOMEM;
    r->logn = logn;
    r->a = a;
    r->b = b;
    r->c = c;
    r->bytes = s->request_bytes(logn);
    // the device address of page-locked host memory, null for pageable memory
    auto page_locked = [](const void* p) -> void* {
        hipPointerAttribute_t attr{};
        if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
            (void)hipGetLastError();  // pageable memory: not an error
            return nullptr;
        }
        return attr.type == hipMemoryTypeHost ? attr.devicePointer : nullptr;
    };
    r->a_dev = static_cast<const uint32_t*>(page_locked(a));
    r->b_dev = static_cast<const uint32_t*>(page_locked(b));
    r->c_dev = static_cast<uint32_t*>(page_locked(c));
    r->ab_pinned = r->a_dev && r->b_dev;
    r->c_pinned = r->c_dev != nullptr;
    if (r->bytes > s->budget) {
        delete r;
        return pm_fail(s->ctx, PM_EBUDGET, "request exceeds the HBM budget");
    }
    {
        std::lock_guard<std::mutex> lock(s->mu);
        r->id = s->next_id++;
        r->submitted = Clock::now();
        s->pending.push_back(r);
        s->all[r->id] = r;
    }
    s->cv_work.notify_all();
    *id = r->id;
    return PM_OK;
}

int pm_server_wait(pm_server* s, uint64_t id, double* latency_seconds) {
    if (!s) return PM_EINVAL;
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return PM_EINVAL;
    Request* r = it->second;
    s->cv_done.wait(lock, [r] { return r->done; });
    if (latency_seconds) *latency_seconds = std::chrono::duration<double>(r->finished - r->submitted).count();
    const int st = r->status;
    s->all.erase(it);
    delete r;
    if (st != PM_OK)
        s->ctx->last_error = s->error.empty() ? std::string("served product failed")
                                              : "served product failed: " + s->error;
    return st;
}

int pm_server_stats(pm_server* s, uint64_t* completed, uint64_t* launches, uint64_t* max_concurrent,
                    uint64_t* peak_bytes) {
    if (!s) return PM_EINVAL;
    std::lock_guard<std::mutex> lock(s->mu);
    if (completed) *completed = s->completed;
    if (launches) *launches = s->launches_done;
    if (max_concurrent) *max_concurrent = s->max_concurrent;
    if (peak_bytes) *peak_bytes = s->arena.peak();
    return PM_OK;
}

void pm_server_destroy(pm_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Launch& L : s->launches) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        if (L.stream) (void)hipStreamDestroy(L.stream);
        if (L.finished) (void)hipEventDestroy(L.finished);
        (void)hipFree(L.d_jobs);
        (void)hipFree(L.d_items);
        (void)hipFree(L.d_units);
        (void)hipHostFree(L.h_jobs);
        (void)hipHostFree(L.h_items);
        (void)hipHostFree(L.h_units);
        (void)hipHostFree(L.h_in_segs);
        (void)hipHostFree(L.h_out_segs);
    }
    (void)hipFree(s->arena_base);
    for (auto& kv : s->all) delete kv.second;
    delete s;
}

}  // extern "C"
