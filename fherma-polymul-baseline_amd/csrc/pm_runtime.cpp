// Context, resident tables and the plain product path behind include/polymul.h.
#include <hip/hip_runtime.h>

#include <immintrin.h>
#include <sched.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <string>
#include <vector>

#include "pm_device.h"
#include "pm_kernels.h"
#include "pm_params.h"
#include "pm_placement.h"
#include "pm_runtime.h"
#include "polymul.h"

using pmd::kLimbs;

namespace {

std::once_flag g_configure_once;
hipError_t g_configure_status = hipSuccess;
thread_local std::string g_create_error;

int create_failed(int status, const std::string& why) {
    g_create_error = why;
    return status;
}

}  // namespace

int pm_fail(pm_context* ctx, int status, const std::string& why) {
    if (ctx) ctx->last_error = why;
    return status;
}

int pm_hip(pm_context* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return PM_OK;
    return pm_fail(ctx, e == hipErrorOutOfMemory ? PM_ENOMEM : PM_EDEVICE,
                   std::string(what) + ": " + hipGetErrorString(e));
}

void* pm_context::device_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    allocations.push_back(p);
    return p;
}

template <class T>
static const T* upload(pm_context* ctx, const std::vector<T>& v, int* status) {
    if (*status != PM_OK) return nullptr;
    void* p = ctx->device_alloc(v.size() * sizeof(T));
    if (!p) {
        *status = pm_fail(ctx, PM_ENOMEM, "device allocation of a parameter table failed");
        return nullptr;
    }
    ctx->table_bytes += v.size() * sizeof(T);
    hipError_t e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) *status = pm_hip(ctx, e, "upload parameter table");
    return static_cast<const T*>(p);
}

// outer passes of the plain path: a single product has only k tiles of 2^15,
// so from N = 2^14 on it is split into 2-4 smaller tiles per prime (more
// blocks) at the cost of the outer passes' HBM traffic
uint32_t pm_default_outer(uint32_t logn) { return logn <= 13 ? 0u : (logn == 14 ? 1u : 2u); }

// batches: outer stages (run inside the conversion and the CRT) down to a
// tile of 2^13 where two suffice; one 2^13 tile per 256-thread block lets
// four blocks share a compute unit, so one block's loads and barriers overlap
// the others' arithmetic.  Measured at 128 x N = 2^15 (tile phase, same
// residues): 2^15 tiles in 1024-thread blocks 1.68 ms, 2^14 tiles one per
// block 1.45 ms, 2^13 tiles one per block 1.24 ms.
// The quartic backend runs outer stages as passes of their own (two HBM
// round trips), so its batches keep the largest tile: measured at 128 x N =
// 2^15 and 64 rows, 2^13 tiles plus two outer passes 2.52 ms, one 2^15 tile
// about 1.8 ms.
uint32_t pm_default_batch_outer(uint32_t logn, bool quartic) {
    if (quartic) return logn > 15 ? logn - 15 : 0u;
    return logn <= 13 ? 0u : (logn - 13 < 2 ? logn - 13 : 2u);
}

// one tile per block from 2^13 up (2^15: the whole 1024-thread block);
// smaller tiles fill a 1024-thread block
uint32_t pm_default_batch_tiles(uint32_t tlog, uint32_t residues_per_thread) {
    // 32 per thread: one tile per block from 2^13 up (256 threads or more),
    // 1024-thread blocks below; 16 per thread: the same from 2^12 up
    const uint32_t one_from = residues_per_thread == 16 ? 12u : 13u;
    const uint32_t block_log = residues_per_thread == 16 ? 14u : 15u;  // residues of a 1024-thread block
    if (tlog >= one_from && tlog <= 15) return 1u;
    return tlog >= 5 && tlog < one_from ? 1u << (block_log - tlog) : 1u;
}

bool pm_outer_valid(uint32_t logn, uint32_t outer, uint32_t lo) {
    return outer <= 2 && outer <= logn && logn - outer <= 15 && logn - outer >= lo;
}

// input chunks of the host-buffer path (the output goes in half as many):
// the set value, else pieces of at most 2^12 coefficients (~0.46 MB), 2 to
// 16 of them
uint32_t pm_io_chunks(const pm_context* ctx, uint32_t logn) {
    if (ctx->io_chunks[logn]) return ctx->io_chunks[logn];
    uint32_t chunks = 2;
    while (chunks < pm_context::kMaxIoChunks && (2u << logn) / chunks > 4096u) chunks *= 2;
    return chunks;
}

constexpr int kCollectTimeoutSeconds = 60;
constexpr uint64_t kDefaultIoWaitMs = 5000;  // bound of one device wait on a host flag

extern "C" {

int pm_create(uint32_t max_n, uint32_t w_bits, uint32_t limbs, const uint32_t* q, int device,
              pm_context** out) {
    g_create_error.clear();
    if (!out) return create_failed(PM_EINVAL, "null output pointer");
    *out = nullptr;
    if (!q) return create_failed(PM_EINVAL, "null modulus");
    if (max_n < 2 || (max_n & (max_n - 1)) || max_n > (1u << pm::kMaxLogN))
        return create_failed(PM_EUNSUPPORTED, "coverage: power-of-two 2 <= N <= 131072");
    if (w_bits != pm::kWidth || limbs != pm::kLimbs)
        return create_failed(PM_EUNSUPPORTED, "coverage: W=868, L=28");
    pm_context* ctx = new (std::nothrow) pm_context();
    if (!ctx) return create_failed(PM_ENOMEM, "host allocation failed");
    ctx->device = device;
    ctx->max_logn = uint32_t(__builtin_ctz(max_n));
    const std::string why = pm::build_params(ctx->max_logn, q, limbs, ctx->params);
    if (!why.empty()) {
        delete ctx;
        return create_failed(PM_EUNSUPPORTED, "coverage: " + why);
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete ctx;
        return create_failed(PM_EDEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    }
    std::call_once(g_configure_once, [] { g_configure_status = pmd::configure_kernels(); });
    if (g_configure_status != hipSuccess) {
        delete ctx;
        return create_failed(PM_EDEVICE, std::string("kernel configuration: ") + hipGetErrorString(g_configure_status));
    }
    e = pmd::scratch_address(device, &ctx->scratch);
    if (e != hipSuccess) {
        delete ctx;
        return create_failed(PM_EDEVICE, std::string("static scratch address: ") + hipGetErrorString(e));
    }
    int status = PM_OK;
    const pm::Params& s = ctx->params;
    const uint32_t k = s.k, quads = (k + 3) / 4;

    // conversion operands: [quad][step][lane] x 16 bytes
    std::vector<uint8_t> conv_w(size_t(quads) * 2 * 64 * 16, 0);
    for (uint32_t quad = 0; quad < quads; ++quad)
        for (uint32_t step = 0; step < 2; ++step)
            for (uint32_t lane = 0; lane < 64; ++lane)
                for (uint32_t e8 = 0; e8 < 16; ++e8) {
                    const uint32_t row = lane % 16, prime = quad * 4 + row / 4, digit = row % 4;
                    const uint32_t byte = 64 * step + 16 * (lane / 16) + e8;
                    int8_t v = 0;
                    if (prime < k && byte < pm::kConvBytes)
                        v = s.conv_digits[(size_t(prime) * pm::kConvBytes + byte) * 4 + digit];
                    conv_w[((size_t(quad) * 2 + step) * 64 + lane) * 16 + e8] = uint8_t(v);
                }
    std::vector<int32_t> conv_corr(size_t(quads) * 4 * 4, 0);
    std::copy(s.conv_correction.begin(), s.conv_correction.end(), conv_corr.begin());
    std::vector<pmd::Tw> fwd(s.forward.size()), inv(s.inverse.size()), scale(s.scale.size()),
        limb_pow(s.limb_pow.size());
    for (size_t i = 0; i < limb_pow.size(); ++i) limb_pow[i] = {s.limb_pow[i].value, s.limb_pow[i].shoup};
    // forward twiddles carry -w mod 2^32: the butterfly forms q*p - Y*w as q*p + Y*(-w)
    for (size_t i = 0; i < fwd.size(); ++i) fwd[i] = {0u - s.forward[i].value, s.forward[i].shoup};
    for (size_t i = 0; i < inv.size(); ++i) inv[i] = {s.inverse[i].value, s.inverse[i].shoup};
    // complement twiddles p - w of the inverse (its complemented outputs)
    std::vector<pmd::Tw> invc(s.inverse.size());
    for (size_t i = 0; i < invc.size(); ++i) {
        const pm::Twiddle c = pm::complement_twiddle(s.inverse[i], s.primes[i >> s.max_logn]);
        invc[i] = {c.value, c.shoup};
    }
    // the tile's scale and block-product constants, per prime
    const uint32_t zrow = 1u << (ctx->max_logn - 1);
    std::vector<pmd::Tw> zs(size_t(s.k) * zrow);
    {
        std::vector<pm::Twiddle> s4(ctx->max_logn + 1), z(zrow);
        for (uint32_t i = 0; i < s.k; ++i) {
            pm::block_product_constants(s.primes[i], ctx->max_logn, s.forward.data() + (size_t(i) << s.max_logn),
                                        s.scale.data() + i, s.k, s4.data(), z.data());
            for (uint32_t logn = 0; logn <= ctx->max_logn; ++logn)
                scale[size_t(logn) * s.k + i] = {s4[logn].value, s4[logn].shoup};
            for (uint32_t x = 0; x < zrow; ++x) zs[size_t(i) * zrow + x] = {z[x].value, z[x].shoup};
        }
    }

    pmd::Tables& tb = ctx->tables;
    tb.primes = upload(ctx, s.primes, &status);
    tb.mont = upload(ctx, s.mont_neg_inv, &status);
    tb.fwd = upload(ctx, fwd, &status);
    tb.inv = upload(ctx, inv, &status);
    tb.invc = upload(ctx, invc, &status);
    tb.scale = upload(ctx, scale, &status);
    tb.zs = upload(ctx, zs, &status);
    tb.conv_w = reinterpret_cast<const uint4*>(upload(ctx, conv_w, &status));
    tb.conv_corr = upload(ctx, conv_corr, &status);
    tb.conv_sh16 = upload(ctx, s.conv_shoup16, &status);
    tb.limb_pow = upload(ctx, limb_pow, &status);
    tb.crt_m30 = upload(ctx, s.crt_bases30, &status);
    tb.crt_round = upload(ctx, s.crt_round, &status);
    tb.crt_negkp = upload(ctx, s.crt_neg_kp, &status);
    tb.crt_a = reinterpret_cast<const uint4*>(upload(ctx, s.crt_mfma_a, &status));
    {
        std::vector<int32_t> corr(s.crt_mfma_corr);  // [row] -> int4 per 4 rows
        tb.crt_corr = reinterpret_cast<const int4*>(upload(ctx, corr, &status));
    }
    tb.q = upload(ctx, s.q, &status);
    tb.mu0 = s.barrett_mu[0];
    tb.mu1 = s.barrett_mu[1];
    tb.k = k;
    tb.quads = quads;
    tb.max_logn = ctx->max_logn;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    tb.cus = uint32_t(cus);
    if (status != PM_OK) {
        const std::string reason = ctx->last_error;
        pm_destroy(ctx);
        return create_failed(status, reason);
    }
    // plain-path workspace for N up to max_n
    const size_t coef_bytes = size_t(max_n) * kLimbs * sizeof(uint32_t);
    const size_t res_bytes = size_t(k) * max_n * sizeof(uint32_t);
    ctx->d_a = static_cast<uint32_t*>(ctx->device_alloc(2 * coef_bytes));  // a then b: one transfer target
    ctx->d_b = ctx->d_a ? ctx->d_a + size_t(max_n) * kLimbs : nullptr;
    ctx->d_c = static_cast<uint32_t*>(ctx->device_alloc(coef_bytes));
    ctx->d_ra = static_cast<uint32_t*>(ctx->device_alloc(res_bytes));
    ctx->d_rb = static_cast<uint32_t*>(ctx->device_alloc(res_bytes));
    ctx->workspace_bytes = 3 * coef_bytes + 2 * res_bytes;
    ctx->workspace_rows = k;
    if (!ctx->d_a || !ctx->d_b || !ctx->d_c || !ctx->d_ra || !ctx->d_rb) {
        pm_destroy(ctx);
        return create_failed(PM_ENOMEM, "device workspace allocation failed");
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->workspace_free, hipEventDisableTiming) != hipSuccess) {
        pm_destroy(ctx);
        return create_failed(PM_EDEVICE, "stream creation failed");
    }
    // page-locked staging of the host-buffer path, mapped into the device
    // (the device reads the inputs and writes the result there while the host
    // copies other chunks), and a fine-grained (coherent) page of flags
    // through which each side tells the other that a chunk has landed
    const unsigned mapped = hipHostMallocMapped;
    if (hipHostMalloc(reinterpret_cast<void**>(&ctx->h_in), 2 * coef_bytes, mapped) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&ctx->h_out), coef_bytes, mapped) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&ctx->h_flags), 4096, mapped | hipHostMallocCoherent) != hipSuccess) {
        pm_destroy(ctx);
        return create_failed(PM_ENOMEM, "page-locked staging allocation failed");
    }
    std::memset(ctx->h_flags, 0, 4096);
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->hd_in), ctx->h_in, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->hd_out), ctx->h_out, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->hd_flags), ctx->h_flags, 0) != hipSuccess) {
        pm_destroy(ctx);
        return create_failed(PM_EDEVICE, "mapping page-locked staging into the device failed");
    }
    // egress chunk counters, then the ingest's mirror of the ready flags and its error word
    const size_t io_words = pmd::kMaxIoChunks + (pmd::kMaxIoChunks + 1) * pmd::kMirrorStride;
    ctx->d_io_counters = static_cast<uint32_t*>(ctx->device_alloc(io_words * sizeof(uint32_t)));
    if (ctx->d_io_counters && hipMemset(ctx->d_io_counters, 0, io_words * sizeof(uint32_t)) != hipSuccess) {
        pm_destroy(ctx);
        return create_failed(PM_EDEVICE, "clearing the transfer counters failed");
    }
    if (!ctx->d_io_counters) {
        pm_destroy(ctx);
        return create_failed(PM_ENOMEM, "device workspace allocation failed");
    }
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    ctx->clock_khz = uint64_t(rate_khz);
    ctx->wait_ticks = ctx->clock_khz * kDefaultIoWaitMs;
    {
        // copy workers on the device's NUMA node, one per physical core, off
        // the creating thread's core (the thread that will call submit)
        std::vector<int> cpus;
        int node = -1;
        if (pm_placement_enabled()) {
            char bus[64] = {};
            if (hipDeviceGetPCIBusId(bus, sizeof(bus), device) == hipSuccess) node = pm_pci_numa_node(bus);
            cpus = pm_worker_order(pm_allowed_cpus(node), sched_getcpu());
        }
        ctx->numa_node = node;
        const unsigned cores = cpus.empty() ? std::thread::hardware_concurrency() : unsigned(cpus.size());
        const unsigned workers = cores > 4 ? 3u : (cores > 1 ? cores - 1 : 0u);
        ctx->copy_pool.reset(new (std::nothrow) PmCopyPool(workers, cpus));
        if (!ctx->copy_pool) {
            pm_destroy(ctx);
            return create_failed(PM_ENOMEM, "host allocation failed");
        }
    }
    for (uint32_t logn = 0; logn <= pm::kMaxLogN; ++logn) {
        ctx->outer[logn] = pm_default_outer(logn);
        ctx->batch_outer[PM_BACKEND_GENERAL][logn] = pm_default_batch_outer(logn, false);
        ctx->batch_outer[PM_BACKEND_QUARTIC][logn] = pm_default_batch_outer(logn, true);
    }
    for (uint32_t tlog = 0; tlog < 16; ++tlog) {
        ctx->batch_tiles[tlog] = pm_default_batch_tiles(tlog);
        ctx->batch_rpt[tlog] = 32;
    }
    // one product of size max_n on zero inputs: the runtime loads each kernel
    // at its first launch, which would otherwise land in the first timed call
    {
        std::vector<uint32_t> zeros(size_t(max_n) * kLimbs, 0u), sink(zeros.size());
        const int st = pm_multiply(ctx, max_n, zeros.data(), zeros.data(), sink.data());
        if (st != PM_OK) {
            const std::string reason = ctx->last_error;
            pm_destroy(ctx);
            return create_failed(st, "warm-up product: " + reason);
        }
    }
    *out = ctx;
    return PM_OK;
}

int pm_plan_rns(uint32_t max_n, const uint32_t* q, uint32_t* k, uint32_t* primes, uint32_t* psi_max,
                uint32_t* product_bits) {
    g_create_error.clear();
    if (!q || !k) return create_failed(PM_EINVAL, "null argument");
    if (max_n < 2 || (max_n & (max_n - 1)) || max_n > (1u << pm::kMaxLogN))
        return create_failed(PM_EUNSUPPORTED, "coverage: power-of-two 2 <= N <= 131072");
    pm::Params s;
    const std::string why = pm::build_params(uint32_t(__builtin_ctz(max_n)), q, pm::kLimbs, s);
    if (!why.empty()) return create_failed(PM_EUNSUPPORTED, "coverage: " + why);
    *k = s.k;
    for (uint32_t i = 0; i < s.k; ++i) {
        if (primes) primes[i] = s.primes[i];
        if (psi_max) psi_max[i] = s.psi_max[i];
    }
    if (product_bits) *product_bits = s.product_bits;
    return PM_OK;
}

int pm_plan_quartic(uint32_t max_n, const uint32_t* q, uint32_t* k, uint32_t* primes, uint32_t* roots,
                    uint32_t* iotas, uint32_t* product_bits) {
    g_create_error.clear();
    if (!q || !k) return create_failed(PM_EINVAL, "null argument");
    if (max_n < 4 || (max_n & (max_n - 1)) || max_n > (1u << pm::kMaxLogN))
        return create_failed(PM_EUNSUPPORTED, "coverage: power-of-two 4 <= N <= 131072");
    pm::QuarticParams s;
    const std::string why = pm::build_quartic_params(uint32_t(__builtin_ctz(max_n)), q, s);
    if (!why.empty()) return create_failed(PM_EUNSUPPORTED, "coverage: " + why);
    *k = s.k;
    for (uint32_t i = 0; i < s.k; ++i) {
        if (primes) primes[i] = s.primes[i];
        if (roots) roots[i] = s.root4[i];
        if (iotas) iotas[i] = s.iota[i].value;
    }
    if (product_bits) *product_bits = s.product_bits;
    return PM_OK;
}

void pm_destroy(pm_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->workspace_free) {
        // the last plain product may have run on a caller's stream
        if (ctx->workspace_used) (void)hipEventSynchronize(ctx->workspace_free);
        (void)hipEventDestroy(ctx->workspace_free);
    }
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }

    for (void* p : ctx->allocations) (void)hipFree(p);
    ctx->copy_pool.reset();  // joins the copy threads
    if (ctx->h_in) (void)hipHostFree(ctx->h_in);
    if (ctx->h_out) (void)hipHostFree(ctx->h_out);
    if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
    delete ctx;
}

const char* pm_last_error(const pm_context* ctx) { return ctx ? ctx->last_error.c_str() : g_create_error.c_str(); }

int pm_get_info(const pm_context* ctx, pm_info_t* info) {
    if (!ctx || !info) return PM_EINVAL;
    info->max_n = 1u << ctx->max_logn;
    info->k = ctx->params.k;
    info->product_bits = ctx->params.product_bits;
    info->device = uint32_t(ctx->device);
    info->table_bytes = ctx->table_bytes;
    info->workspace_bytes = ctx->workspace_bytes;
    return PM_OK;
}

int pm_get_prime(const pm_context* ctx, uint32_t i, uint32_t* p, uint32_t* psi_max) {
    if (!ctx || i >= ctx->params.k) return PM_EINVAL;
    if (p) *p = ctx->params.primes[i];
    if (psi_max) *psi_max = ctx->params.psi_max[i];
    return PM_OK;
}

int pm_synchronize(pm_context* ctx) {
    if (!ctx) return PM_EINVAL;
    return pm_hip(ctx, hipStreamSynchronize(ctx->stream), "synchronize");
}

}  // extern "C"

int pm_check_n(pm_context* ctx, uint32_t n, uint32_t* logn) {
    if (n < 2 || (n & (n - 1))) return pm_fail(ctx, PM_EINVAL, "N must be a power of two >= 2");
    *logn = uint32_t(__builtin_ctz(n));
    if (*logn > ctx->max_logn) return pm_fail(ctx, PM_EUNSUPPORTED, "N exceeds the context's max_n");
    return PM_OK;
}

hipError_t pm_workspace_acquire(pm_context* ctx, hipStream_t stream) {
    return ctx->workspace_used ? hipStreamWaitEvent(stream, ctx->workspace_free, 0) : hipSuccess;
}

hipError_t pm_workspace_release(pm_context* ctx, hipStream_t stream) {
    const hipError_t e = hipEventRecord(ctx->workspace_free, stream);
    if (e == hipSuccess) ctx->workspace_used = true;
    return e;
}

const pmd::Tables& pm_tables(const pm_context* ctx) { return ctx->backend ? ctx->qtables : ctx->tables; }

uint32_t pm_rows(const pm_context* ctx) { return pm_tables(ctx).k; }

int pm_run_plain(pm_context* ctx, const pmd::Job& job_in, uint32_t outer, hipStream_t stream) {
    // the outer stages run inside the conversion and the CRT where they can
    // (the MFMA kernels of the general backend)
    const pmd::Tables& tb = pm_tables(ctx);
    pmd::Job job = job_in;
    job.outer = outer;
    job.fused = !tb.quartic && pmd::outer_fusable(job.logn, outer) ? 1u : 0u;
    job.conv_slices = ctx->conv_slices[job.logn];
    hipError_t e = pm_workspace_acquire(ctx, stream);
    if (e == hipSuccess) e = pmd::launch_convert(tb, job, stream);
    if (e == hipSuccess) e = pmd::launch_ntt(tb, job, outer, stream);
    if (e == hipSuccess) e = pmd::launch_crt(tb, job, stream);
    if (e == hipSuccess) e = pm_workspace_release(ctx, stream);
    return pm_hip(ctx, e, "launch plain product");
}

extern "C" {

int pm_multiply_device(pm_context* ctx, uint32_t n, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                       void* hip_stream) {
    if (!ctx || !d_a || !d_b || !d_c) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
    pmd::Job job{logn, 0, d_a, d_b, d_c, ctx->d_ra, ctx->d_rb};
    return pm_run_plain(ctx, job, ctx->outer[logn], stream);
}

int pm_multiply_device_noalloc(pm_context* ctx, uint32_t n, const uint32_t* d_a, const uint32_t* d_b,
                               uint32_t* d_c, void* hip_stream) {
    if (!ctx || !d_a || !d_b || !d_c) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    const pmd::Tables& tb = pm_tables(ctx);
    const uint64_t words = 2ull * tb.k * n;
    if (words > pmd::kScratchWords)
        return pm_fail(ctx, PM_EUNSUPPORTED,
                       "2 x k x N = " + std::to_string(words) + " residues exceed the static scratch of " +
                           std::to_string(pmd::kScratchWords) + " (16 MiB: N <= 32768 for k <= 64)");
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
    // every residue the kernels read was written by this call's conversion:
    // nothing in the scratch needs clearing
    const uint32_t outer = ctx->outer[logn];
    pmd::Job job{logn, outer, d_a, d_b, d_c, ctx->scratch, ctx->scratch + size_t(tb.k) * n,
                 !tb.quartic && pmd::outer_fusable(logn, outer) ? 1u : 0u, ctx->conv_slices[logn]};
    hipError_t e = pmd::launch_convert(tb, job, stream);
    if (e == hipSuccess) e = pmd::launch_ntt(tb, job, outer, stream);
    if (e == hipSuccess) e = pmd::launch_crt(tb, job, stream);
    return pm_hip(ctx, e, "launch allocation-free product");
}

int pm_multiply(pm_context* ctx, uint32_t n, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    if (!ctx || !a || !b || !c) return PM_EINVAL;
    if (c == a || c == b) return pm_fail(ctx, PM_EINVAL, "c may not alias a or b");
    const int st = pm_multiply_submit(ctx, n, a, b);
    if (st != PM_OK) return st;
    return pm_multiply_collect(ctx, c);
}

// The host-buffer path.  Every kernel of the product is enqueued first; the
// first one converts a|b (one array of 2n coefficients) straight from
// page-locked staging, each 64-coefficient group as soon as the host flags
// its chunk ready, so host copies, PCIe transfers and conversion overlap.
// The NTTs and CRT run from device memory, and an egress kernel returns the
// result into page-locked staging chunk by chunk, flagging each chunk as it
// lands; collect copies each chunk out as soon as its flag is up.
namespace {

volatile uint32_t* ready_flags(pm_context* ctx) { return ctx->h_flags; }
volatile uint32_t* done_flags(pm_context* ctx) { return ctx->h_flags + pmd::kMaxIoChunks; }
volatile uint32_t* error_flag(pm_context* ctx) { return ctx->h_flags + 2 * pmd::kMaxIoChunks; }

// One result chunk out of staging into the caller's array, split over the copy
// workers and this thread: the next chunk is still crossing PCIe meanwhile, so
// only the last chunk's copy is exposed, and that at several cores' rate.
void copy_out(pm_context* ctx, uint32_t* dst, const uint32_t* src, size_t bytes) {
    constexpr size_t kMinPiece = size_t(64) << 10;
    const size_t parts_max = size_t(ctx->copy_pool->workers()) + 1;
    size_t parts = bytes / kMinPiece;
    parts = parts < 1 ? 1 : (parts > parts_max ? parts_max : parts);
    if (parts == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    PmCopyPool::Chunk pieces[PmCopyPool::kMaxChunks];
    const size_t piece = (bytes / parts + 63) & ~size_t(63);
    uint32_t count = 0;
    for (size_t off = 0; off < bytes && count < PmCopyPool::kMaxChunks; off += piece, ++count) {
        const size_t len = bytes - off < piece ? bytes - off : piece;
        pieces[count] = {reinterpret_cast<char*>(dst) + off, reinterpret_cast<const char*>(src) + off, len, nullptr,
                         0};
    }
    ctx->copy_pool->run(pieces, count);
}

}  // namespace

int pm_multiply_submit(pm_context* ctx, uint32_t n, const uint32_t* a, const uint32_t* b) {
    if (!ctx || !a || !b) return PM_EINVAL;
    if (ctx->pending) return pm_fail(ctx, PM_EINVAL, "a submitted product has not been collected");
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return pm_hip(ctx, e, "set device");
    const uint32_t chunks = pm_io_chunks(ctx, logn);
    const uint32_t in_chunks = chunks < 2 ? 2 : chunks;  // a and b are separate arrays
    const uint32_t out_chunks = chunks < 2 ? 1 : chunks / 2;
    const uint32_t groups_per_poly = (n + pmd::kConvPerWave - 1) / pmd::kConvPerWave;
    const uint32_t groups_per_chunk = 2 * groups_per_poly / in_chunks;
    const size_t words = size_t(n) * kLimbs;
    if (++ctx->io_seq == 0) ctx->io_seq = 1;
    *error_flag(ctx) = 0;
    pmd::IoArgs io{ctx->hd_flags, ctx->hd_flags + pmd::kMaxIoChunks, ctx->hd_flags + 2 * pmd::kMaxIoChunks,
                   ctx->d_io_counters, ctx->d_io_counters + pmd::kMaxIoChunks, ctx->io_seq, in_chunks,
                   groups_per_chunk, 0, ctx->wait_ticks};
    const pmd::Tables& tb = pm_tables(ctx);
    pmd::Job job{logn, 0, ctx->hd_in, ctx->hd_in + words, ctx->d_c, ctx->d_ra, ctx->d_rb};
    e = pm_workspace_acquire(ctx, ctx->stream);
    // the general backend converts straight from page-locked staging as chunks
    // land; the quartic one converts from device memory after one transfer
    // of the whole staging (enqueued behind the host copies below)
    if (e == hipSuccess && !tb.quartic) e = pmd::launch_ingest_convert(tb, job, io, ctx->stream);
    // input chunk j: coefficients [j, j+1) * span of the concatenated a|b.  The
    // other launches come after the copies: a launch costs several us of host
    // time, and the copies are what the device is waiting for
    const size_t span_words = words * 2 / in_chunks;
    if (e == hipSuccess) {
        PmCopyPool::Chunk chunks_of_input[pmd::kMaxIoChunks];
        for (uint32_t j = 0; j < in_chunks; ++j) {
            const size_t off = j * span_words;
            volatile uint32_t* flag = int32_t(j) == ctx->stall_chunk ? &ctx->stall_sink : &ready_flags(ctx)[j];
            chunks_of_input[j] = {ctx->h_in + off, off < words ? a + off : b + (off - words),
                                  span_words * sizeof(uint32_t), flag, ctx->io_seq};
        }
        ctx->stall_chunk = -1;
        ctx->copy_pool->run(chunks_of_input, in_chunks);  // returns once a and b are no longer needed
    }
    if (e == hipSuccess && tb.quartic) {
        // the egress chunk counters, which the general path's ingest kernel clears
        e = hipMemsetAsync(ctx->d_io_counters, 0, pmd::kMaxIoChunks * sizeof(uint32_t), ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ctx->d_a, ctx->h_in, 2 * words * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream);
        job.a = ctx->d_a;
        job.b = ctx->d_a + words;  // the staging's layout: b right behind this product's a
        if (e == hipSuccess) e = pmd::launch_convert(tb, job, ctx->stream);
    }
    if (e == hipSuccess) e = pmd::launch_ntt(tb, job, ctx->outer[logn], ctx->stream);
    if (e == hipSuccess) e = pmd::launch_crt(tb, job, ctx->stream);
    if (e == hipSuccess) e = pmd::launch_egress(ctx->d_c, ctx->hd_out, words / out_chunks, out_chunks, io, ctx->stream);
    if (e == hipSuccess) e = pm_workspace_release(ctx, ctx->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(ctx->stream);
        return pm_hip(ctx, e, "launch product");
    }
    ctx->pending = true;
    ctx->pending_n = n;
    ctx->pending_out_chunks = out_chunks;
    return PM_OK;
}

int pm_multiply_collect(pm_context* ctx, uint32_t* c) {
    if (!ctx || !c) return PM_EINVAL;
    if (!ctx->pending) return pm_fail(ctx, PM_EINVAL, "no submitted product to collect");
    ctx->pending = false;
    const uint32_t chunks = ctx->pending_out_chunks, seq = ctx->io_seq;
    const size_t chunk_words = size_t(ctx->pending_n) * kLimbs / chunks;
    const auto start = std::chrono::steady_clock::now();
    for (uint32_t j = 0; j < chunks; ++j) {
        // spin on the device-written flag; now and then look at the stream so
        // a device failure ends the wait instead of the time limit
        for (uint32_t spins = 0; __atomic_load_n(&done_flags(ctx)[j], __ATOMIC_ACQUIRE) != seq; ++spins) {
            _mm_pause();
            if ((spins & 4095u) != 4095u) continue;
            const hipError_t q = hipStreamQuery(ctx->stream);
            if (q != hipSuccess && q != hipErrorNotReady) {
                (void)hipStreamSynchronize(ctx->stream);
                return pm_hip(ctx, q, "product");
            }
            if (q == hipSuccess && __atomic_load_n(&done_flags(ctx)[j], __ATOMIC_ACQUIRE) != seq) {
                return pm_fail(ctx, PM_EDEVICE, "product finished without flagging its result");
            }
            if (std::chrono::steady_clock::now() - start > std::chrono::seconds(kCollectTimeoutSeconds)) {
                (void)hipStreamSynchronize(ctx->stream);
                return pm_fail(ctx, PM_ETIMEOUT, "product did not complete within the time limit");
            }
        }
        copy_out(ctx, c + j * chunk_words, ctx->h_out + j * chunk_words, chunk_words * sizeof(uint32_t));
    }
    if (*error_flag(ctx) != 0) {
        (void)hipStreamSynchronize(ctx->stream);
        return pm_fail(ctx, PM_ETIMEOUT, "a wait for an input chunk gave up");
    }
    return PM_OK;
}

int pm_set_io_wait_limit(pm_context* ctx, uint32_t microseconds) {
    if (!ctx) return PM_EINVAL;
    ctx->wait_ticks = microseconds ? ctx->clock_khz * microseconds / 1000u : ctx->clock_khz * kDefaultIoWaitMs;
    return PM_OK;
}

int pm_place_near_device(int device, int* node, int* cpus_on_node) {
    if (node) *node = -1;
    if (cpus_on_node) *cpus_on_node = 0;
    char bus[64] = {};
    if (hipDeviceGetPCIBusId(bus, sizeof(bus), device) != hipSuccess) return PM_EDEVICE;
    const int numa = pm_pci_numa_node(bus);
    if (node) *node = numa;
    if (numa < 0) return PM_EUNSUPPORTED;
    std::vector<int> cpus;
    for (const PmCpu& c : pm_allowed_cpus(numa)) cpus.push_back(c.cpu);
    if (cpus_on_node) *cpus_on_node = int(cpus.size());
    return pm_pin_thread(cpus) ? PM_OK : PM_EUNSUPPORTED;
}

int pm_inject_io_stall(pm_context* ctx, int32_t chunk) {
    if (!ctx || chunk < -1 || chunk >= int32_t(pmd::kMaxIoChunks)) return PM_EINVAL;
    ctx->stall_chunk = chunk;
    return PM_OK;
}

int pm_set_io_chunks(pm_context* ctx, uint32_t n, uint32_t chunks) {
    if (!ctx) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    if (chunks != 0 && (chunks > pm_context::kMaxIoChunks || (chunks & (chunks - 1)) || n / chunks < 64 ||
                        n % chunks))
        return pm_fail(ctx, PM_EINVAL, "chunks must be 0 or a power of two <= 16 leaving >= 64 coefficients each");
    ctx->io_chunks[logn] = chunks;
    return PM_OK;
}

int pm_stage_convert(pm_context* ctx, uint32_t n, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_ra,
                     uint32_t* d_rb, void* hip_stream) {
    if (!ctx || !d_a || !d_b || !d_ra || !d_rb) return PM_EINVAL;
    if ((reinterpret_cast<uintptr_t>(d_ra) | reinterpret_cast<uintptr_t>(d_rb)) & 15u)
        return pm_fail(ctx, PM_EINVAL, "residue buffers must be 16-byte aligned");
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
    pmd::Job job{logn, 0, d_a, d_b, nullptr, d_ra, d_rb};
    return pm_hip(ctx, pmd::launch_convert(pm_tables(ctx), job, stream), "launch conversion");
}

int pm_stage_convert_scalar(pm_context* ctx, uint32_t n, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_ra,
                            uint32_t* d_rb, void* hip_stream) {
    if (!ctx || !d_a || !d_b || !d_ra || !d_rb) return PM_EINVAL;
    if ((reinterpret_cast<uintptr_t>(d_ra) | reinterpret_cast<uintptr_t>(d_rb)) & 15u)
        return pm_fail(ctx, PM_EINVAL, "residue buffers must be 16-byte aligned");
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
    pmd::Job job{logn, 0, d_a, d_b, nullptr, d_ra, d_rb};
    return pm_hip(ctx, pmd::launch_convert_scalar(pm_tables(ctx), job, stream), "launch scalar conversion");
}

int pm_stage_ntt(pm_context* ctx, uint32_t n, uint32_t outer, uint32_t* d_ra, uint32_t* d_rb, void* hip_stream) {
    if (!ctx || !d_ra || !d_rb) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    if (outer > 2 || outer > logn || logn - outer > 15)
        return pm_fail(ctx, PM_EUNSUPPORTED, "outer stages must leave a tile of at most 2^15 (outer <= 2)");
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
    pmd::Job job{logn, 0, nullptr, nullptr, nullptr, d_ra, d_rb};
    return pm_hip(ctx, pmd::launch_ntt(pm_tables(ctx), job, outer, stream), "launch NTT convolution");
}

// The quartic backend's device tables: per residue row (channel j of prime i
// at row 4i + j) the prime, Montgomery constant, twiddles (the prime's,
// repeated for its four channels so the tile kernels index them by row) and
// post-inverse scale; per prime the conversion and CRT constants.
static int build_quartic(pm_context* ctx) {
    const std::string why = pm::build_quartic_params(ctx->max_logn, ctx->params.q.data(), ctx->qparams);
    if (!why.empty()) return pm_fail(ctx, PM_EUNSUPPORTED, "quartic backend: " + why);
    const pm::QuarticParams& s = ctx->qparams;
    const uint32_t k = s.k, rows = 4 * k, max_n = 1u << ctx->max_logn;
    std::vector<uint32_t> rp(rows), rm(rows);
    const uint32_t zrow = 1u << (ctx->max_logn - 1);
    std::vector<pmd::Tw> fwd(size_t(rows) * max_n), inv(size_t(rows) * max_n), invc(size_t(rows) * max_n),
        scale(size_t(ctx->max_logn + 1) * rows), zs(size_t(rows) * zrow);
    std::vector<pm::Twiddle> s4(ctx->max_logn + 1), z(zrow);
    for (uint32_t row = 0; row < rows; ++row) {
        const uint32_t i = row / 4;
        rp[row] = s.primes[i];
        rm[row] = s.mont_neg_inv[i];
        for (uint32_t x = 0; x < max_n; ++x) {
            const pm::Twiddle f = s.forward[size_t(i) * max_n + x], g = s.inverse[size_t(i) * max_n + x];
            fwd[size_t(row) * max_n + x] = {0u - f.value, f.shoup};  // the butterfly's negated forward twiddle
            inv[size_t(row) * max_n + x] = {g.value, g.shoup};
            const pm::Twiddle c = pm::complement_twiddle(g, s.primes[i]);
            invc[size_t(row) * max_n + x] = {c.value, c.shoup};
        }
        // the tile's scale and block-product constants (the prime's, for each of its channels)
        if (row % 4 == 0)
            pm::block_product_constants(s.primes[i], ctx->max_logn, s.forward.data() + size_t(i) * max_n,
                                        s.scale.data() + i, k, s4.data(), z.data());
        for (uint32_t logn = 0; logn <= ctx->max_logn; ++logn)
            scale[size_t(logn) * rows + row] = {s4[logn].value, s4[logn].shoup};
        for (uint32_t x = 0; x < zrow; ++x) zs[size_t(row) * zrow + x] = {z[x].value, z[x].shoup};
    }
    auto tw = [](const std::vector<pm::Twiddle>& v) {
        std::vector<pmd::Tw> out(v.size());
        for (size_t i = 0; i < v.size(); ++i) out[i] = {v[i].value, v[i].shoup};
        return out;
    };
    int status = PM_OK;
    pmd::Tables tb = ctx->tables;  // device facts (CUs, max_logn) as the general tables have them
    tb.primes = upload(ctx, rp, &status);
    tb.mont = upload(ctx, rm, &status);
    tb.fwd = upload(ctx, fwd, &status);
    tb.inv = upload(ctx, inv, &status);
    tb.invc = upload(ctx, invc, &status);
    tb.scale = upload(ctx, scale, &status);
    tb.zs = upload(ctx, zs, &status);
    tb.k = rows;
    tb.quads = 0;
    tb.conv_w = nullptr;
    tb.conv_corr = nullptr;
    tb.conv_sh16 = nullptr;
    tb.limb_pow = nullptr;
    tb.crt_m30 = nullptr;
    tb.crt_round = nullptr;
    tb.crt_negkp = nullptr;
    tb.crt_a = nullptr;
    tb.crt_corr = nullptr;
    tb.quartic = 1;
    tb.qk = k;
    tb.qdelta = s.delta;
    tb.qprimes = upload(ctx, s.primes, &status);
    tb.q_limb_pow = upload(ctx, tw(s.limb_pow), &status);
    tb.q_rpow = upload(ctx, tw(s.rpow), &status);
    tb.q_rinv = upload(ctx, tw(s.rinv), &status);
    tb.q_iota = upload(ctx, tw(s.iota), &status);
    tb.q_m30 = upload(ctx, s.m30, &status);
    tb.q_round = upload(ctx, s.round, &status);
    tb.q_kp30 = upload(ctx, s.kp30, &status);
    tb.q_negc = upload(ctx, s.negc, &status);
    if (status != PM_OK) return status;
    ctx->qtables = tb;
    ctx->quartic_built = true;
    return PM_OK;
}

int pm_set_backend(pm_context* ctx, uint32_t backend) {
    if (!ctx || backend > PM_BACKEND_QUARTIC) return PM_EINVAL;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return pm_hip(ctx, e, "set device");
    if (backend == PM_BACKEND_QUARTIC && !ctx->quartic_built) {
        const int st = build_quartic(ctx);
        if (st != PM_OK) return st;
    }
    // products already enqueued keep the residue workspace until they finish
    if (ctx->workspace_used) e = hipEventSynchronize(ctx->workspace_free);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return pm_hip(ctx, e, "wait for the plain path");
    const uint32_t rows = backend == PM_BACKEND_QUARTIC ? ctx->qtables.k : ctx->tables.k;
    if (rows > ctx->workspace_rows) {
        // a larger plain-path residue workspace (the old one is released with the context)
        const size_t res_bytes = size_t(rows) << ctx->max_logn << 2;
        auto* ra = static_cast<uint32_t*>(ctx->device_alloc(res_bytes));
        auto* rb = static_cast<uint32_t*>(ctx->device_alloc(res_bytes));
        if (!ra || !rb) return pm_fail(ctx, PM_ENOMEM, "device workspace allocation failed");
        ctx->d_ra = ra;
        ctx->d_rb = rb;
        ctx->workspace_rows = rows;
        ctx->workspace_bytes += 2 * res_bytes;
    }
    ctx->backend = backend;
    return PM_OK;
}

int pm_get_backend(const pm_context* ctx, uint32_t* backend, uint32_t* rows) {
    if (!ctx) return PM_EINVAL;
    if (backend) *backend = ctx->backend;
    if (rows) *rows = pm_rows(ctx);
    return PM_OK;
}

int pm_set_outer(pm_context* ctx, uint32_t n, uint32_t outer) {
    if (!ctx) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    if (outer > 2 || outer > logn || logn - outer > 15 || logn - outer < 1)
        return pm_fail(ctx, PM_EUNSUPPORTED, "outer stages must leave a tile of 2..2^15 (outer <= 2)");
    ctx->outer[logn] = outer;
    return PM_OK;
}

int pm_set_batch_outer(pm_context* ctx, uint32_t n, uint32_t outer) {
    if (!ctx) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    if (!pm_outer_valid(logn, outer, 5))
        return pm_fail(ctx, PM_EUNSUPPORTED, "batch outer stages must leave a tile of 2^5..2^15 (outer <= 2)");
    ctx->batch_outer[ctx->backend][logn] = outer;
    return PM_OK;
}

int pm_set_batch_tiles_per_block(pm_context* ctx, uint32_t tile_n, uint32_t tiles) {
    if (!ctx) return PM_EINVAL;
    if (tile_n < 32 || tile_n > 32768 || (tile_n & (tile_n - 1)))
        return pm_fail(ctx, PM_EINVAL, "tile size must be a power of two in [32, 32768]");
    const uint32_t tlog = uint32_t(__builtin_ctz(tile_n)), rpt = ctx->batch_rpt[tlog];
    if (tiles == 0) tiles = pm_default_batch_tiles(tlog, rpt);
    // a block holds rpt residues per thread, at most 1024 threads, whole waves
    if ((tiles & (tiles - 1)) || uint64_t(tiles) * tile_n > 1024u * rpt || uint64_t(tiles) * tile_n < 64u * rpt)
        return pm_fail(ctx, PM_EINVAL, "tiles per block: a power of two with 64 x r <= tiles x tile size <= 1024 x r "
                                       "(r residues per thread)");
    ctx->batch_tiles[tlog] = tiles;
    return PM_OK;
}

int pm_set_batch_residues_per_thread(pm_context* ctx, uint32_t tile_n, uint32_t residues) {
    if (!ctx) return PM_EINVAL;
    if (tile_n < 32 || tile_n > 32768 || (tile_n & (tile_n - 1)))
        return pm_fail(ctx, PM_EINVAL, "tile size must be a power of two in [32, 32768]");
    const uint32_t tlog = uint32_t(__builtin_ctz(tile_n));
    if (residues == 0) residues = 32;
    if (residues != 32 && residues != 16)
        return pm_fail(ctx, PM_EINVAL, "residues per thread: 16 or 32");
    if (residues == 16 && (tlog < uint32_t(pmd::kRpt16MinLog) || tlog > uint32_t(pmd::kRpt16MaxLog)))
        return pm_fail(ctx, PM_EUNSUPPORTED, "16 residues per thread: tiles of 2^6 .. 2^14");
    ctx->batch_rpt[tlog] = residues;
    ctx->batch_tiles[tlog] = pm_default_batch_tiles(tlog, residues);  // keep the pair valid
    return PM_OK;
}

int pm_get_batch_residues_per_thread(pm_context* ctx, uint32_t tile_n, uint32_t* residues) {
    if (!ctx || !residues || tile_n < 32 || tile_n > 32768 || (tile_n & (tile_n - 1))) return PM_EINVAL;
    *residues = ctx->batch_rpt[__builtin_ctz(tile_n)];
    return PM_OK;
}

int pm_set_batch_crt_waves(pm_context* ctx, uint32_t waves) {
    if (!ctx) return PM_EINVAL;
    if (waves == 0) waves = pmd::kCrtMfmaWaves;
    if (waves != 2 && waves != 4 && waves != 8) return pm_fail(ctx, PM_EINVAL, "CRT waves per block: 2, 4 or 8");
    ctx->batch_crt_waves = waves;
    return PM_OK;
}

int pm_set_batch_parts(pm_context* ctx, uint32_t parts) {
    if (!ctx) return PM_EINVAL;
    if (parts > 4) return pm_fail(ctx, PM_EINVAL, "batch parts: 1 to 4, or 0 to choose per batch");
    ctx->batch_parts = parts;
    return PM_OK;
}

int pm_get_batch_parts(pm_context* ctx, uint32_t* parts) {
    if (!ctx || !parts) return PM_EINVAL;
    *parts = ctx->batch_parts;
    return PM_OK;
}

int pm_get_batch_crt_waves(pm_context* ctx, uint32_t* waves) {
    if (!ctx || !waves) return PM_EINVAL;
    *waves = ctx->batch_crt_waves;
    return PM_OK;
}

int pm_set_conversion_slices(pm_context* ctx, uint32_t n, uint32_t slices) {
    if (!ctx) return PM_EINVAL;
    uint32_t logn = 0;
    const int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    if (slices != 0 && slices != 1 && slices != 2 && slices != 4)
        return pm_fail(ctx, PM_EINVAL, "conversion slices: 1, 2 or 4 (0: chosen at launch)");
    ctx->conv_slices[logn] = slices;
    return PM_OK;
}

int pm_get_conversion_slices(pm_context* ctx, uint32_t n, uint32_t* slices) {
    if (!ctx || !slices) return PM_EINVAL;
    uint32_t logn = 0;
    const int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    *slices = ctx->conv_slices[logn];
    return PM_OK;
}

int pm_get_batch_tiles_per_block(pm_context* ctx, uint32_t tile_n, uint32_t* tiles) {
    if (!ctx || !tiles || tile_n < 32 || tile_n > 32768 || (tile_n & (tile_n - 1))) return PM_EINVAL;
    *tiles = ctx->batch_tiles[__builtin_ctz(tile_n)];
    return PM_OK;
}

int pm_get_batch_outer(pm_context* ctx, uint32_t n, uint32_t* outer) {
    if (!ctx || !outer) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    *outer = ctx->batch_outer[ctx->backend][logn];
    return PM_OK;
}

int pm_get_outer(pm_context* ctx, uint32_t n, uint32_t* outer) {
    if (!ctx || !outer) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    *outer = ctx->outer[logn];
    return PM_OK;
}

int pm_stage_crt_scalar(pm_context* ctx, uint32_t n, const uint32_t* d_t, uint32_t* d_c, void* hip_stream) {
    if (!ctx || !d_t || !d_c) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
    pmd::Job job{logn, 0, nullptr, nullptr, d_c, const_cast<uint32_t*>(d_t), nullptr};
    return pm_hip(ctx, pmd::launch_crt_scalar(pm_tables(ctx), job, stream), "launch scalar CRT");
}

int pm_stage_crt(pm_context* ctx, uint32_t n, const uint32_t* d_t, uint32_t* d_c, void* hip_stream) {
    if (!ctx || !d_t || !d_c) return PM_EINVAL;
    uint32_t logn = 0;
    int st = pm_check_n(ctx, n, &logn);
    if (st != PM_OK) return st;
    hipStream_t stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->stream;
    pmd::Job job{logn, 0, nullptr, nullptr, d_c, const_cast<uint32_t*>(d_t), nullptr};
    return pm_hip(ctx, pmd::launch_crt(pm_tables(ctx), job, stream), "launch CRT");
}

}  // extern "C"
