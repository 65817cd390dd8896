/*
 * polymul.h -- C-ABI of the MI355X (gfx950) exact negacyclic product
 *     c = a * b  mod (X^N + 1, q),   q odd, 2^864 < q < 2^868,
 * with coefficients as 28 little-endian uint32 limbs (fherma.h:31-40).
 *
 * Plain pointers and sizes only; every call returns a pm_status and the
 * context keeps a text reason (pm_last_error).  This is the layer the C++
 * entry points of fherma.h (fherma_init / fherma_run / fherma_free,
 * fherma.h:46-48) are implemented on, and what an FFI binds.
 */
#ifndef POLYMUL_H
#define POLYMUL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_context pm_context;

typedef enum {
    PM_OK = 0,
    PM_EINVAL = 1,        /* bad argument (null pointer, N not a power of two, ...) */
    PM_EUNSUPPORTED = 2,  /* outside coverage: N > max_n, q even or not in (2^864, 2^868) */
    PM_EDEVICE = 3,       /* HIP runtime error (text in pm_last_error)              */
    PM_ENOMEM = 4,        /* device or host allocation failed                       */
    PM_EBUDGET = 5,       /* request exceeds the admission budget                   */
    PM_ETIMEOUT = 6       /* a bounded wait expired                                 */
} pm_status;

#define PM_LIMBS 28
#define PM_WIDTH 868

typedef struct {
    uint32_t max_n;        /* largest supported N (power of two, <= 2^17) */
    uint32_t k;            /* RNS primes */
    uint32_t product_bits; /* bits of the prime product P */
    uint32_t device;       /* HIP device ordinal */
    uint64_t table_bytes;  /* device bytes held by resident tables */
    uint64_t workspace_bytes; /* device bytes held by the plain-path workspace */
} pm_info_t;

/* Host-only parameter plan (no device needed): the RNS of pm_create for this
 * q and max_n.  primes/psi_max may be NULL or hold 64 entries.  Same
 * coverage checks and prime rule as pm_create (rns/host_setup.h:61-78). */
int pm_plan_rns(uint32_t max_n, const uint32_t* q, uint32_t* k, uint32_t* primes, uint32_t* psi_max,
                uint32_t* product_bits);

/* Host-only plan of the quartic backend for q = 2^868 - delta, 0 < delta <
 * 2^28 (role of quartic/host_setup.h:41-110; no device needed): the primes
 * (c 2^18 + 1 < 2^30 with delta a fourth power mod p, as many as make P >
 * 4 max_n (2^217-1)^2 (1 + 3 delta)), each prime's fourth root r of delta,
 * the fourth root of unity i (i^2 = -1) its channels use, and P's bit
 * length.  primes/roots/iotas may be NULL or hold 32 entries.
 * PM_EUNSUPPORTED when q is not such a modulus. */
int pm_plan_quartic(uint32_t max_n, const uint32_t* q, uint32_t* k, uint32_t* primes, uint32_t* roots,
                    uint32_t* iotas, uint32_t* product_bits);
/* Replaces the input-independent part of fherma_init (rns/solve.cu:581-718):
 * coverage checks, RNS parameters, resident device tables, workspace for N up
 * to max_n.  w_bits must be 868 and limbs 28 (fherma.h:28-31). */
int pm_create(uint32_t max_n, uint32_t w_bits, uint32_t limbs, const uint32_t* q, int device,
              pm_context** out);
/* Replaces fherma_free (fherma.h:48). */
void pm_destroy(pm_context* ctx);
/* reason of the last failure on ctx; with ctx == NULL, of this thread's last
 * failed pm_create */
const char* pm_last_error(const pm_context* ctx);
int pm_get_info(const pm_context* ctx, pm_info_t* info);
/* prime i of the RNS and its primitive 2^(log2(max_n)+1)-th root of unity */
int pm_get_prime(const pm_context* ctx, uint32_t i, uint32_t* p, uint32_t* psi_max);

/* Replaces the timed body of fherma_run (rns/solve.cu:721-853): host buffers
 * a, b, c of n*28 uint32 each (c may not alias a or b).  Equal to
 * pm_multiply_submit followed by pm_multiply_collect. */
int pm_multiply(pm_context* ctx, uint32_t n, const uint32_t* a, const uint32_t* b, uint32_t* c);
/* The two halves of pm_multiply, so a caller can do host work (such as
 * allocating the output) while the device computes -- the role of the
 * reference's chunked input pipeline and output completion
 * (rns/solve.cu:492-503,721-853; input_pipeline.h:12-54; output_completion.h:63-91).
 * submit copies a and b chunk by chunk into page-locked staging, each chunk's
 * transfer starting while the next is copied, enqueues the product and
 * returns once a and b are no longer needed; collect waits for the result's
 * chunks and copies each into c as it lands.  One product may be pending per
 * context; PM_EINVAL otherwise. */
int pm_multiply_submit(pm_context* ctx, uint32_t n, const uint32_t* a, const uint32_t* b);
int pm_multiply_collect(pm_context* ctx, uint32_t* c);
/* Transfer chunks of the host-buffer path for products of size n (no
 * counterpart in the reference): the inputs a|b go in `chunks` pieces and
 * the result in chunks/2.  0 restores the default (pieces of at most 2^12
 * coefficients, 2..16 of them); otherwise a power of two <= 16 with
 * n / chunks >= 64. */
int pm_set_io_chunks(pm_context* ctx, uint32_t n, uint32_t chunks);
/* Bound of one device-side wait for a host-buffer input chunk (no
 * counterpart in the reference): 0 restores the default of 5 s.  A wait that
 * runs out latches a timeout; the product still drains, and collect returns
 * PM_ETIMEOUT. */
int pm_set_io_wait_limit(pm_context* ctx, uint32_t microseconds);
/* Host placement (reference host_affinity.h:14-46, pin_near_gpu): restricts
 * the calling thread to the CPUs of `device`'s NUMA node that the process may
 * use, so that page-locked staging allocated afterwards (pm_create) and the
 * caller's copies stay on the GPU's node.  *node and *cpus_on_node report the
 * node and how many CPUs the thread may now use (either may be NULL).
 * PM_EUNSUPPORTED, placement unchanged, when the topology is unknown.
 * pm_create pins its copy workers on that node by itself;
 * PM_HOST_PLACEMENT=0 in the environment turns both off. */
int pm_place_near_device(int device, int* node, int* cpus_on_node);
/* Failure injection for tests (no counterpart in the reference): the next
 * pm_multiply_submit copies input chunk `chunk` but never raises its ready
 * flag, as a host that stalls would; -1 cancels.  The device wait for it runs
 * out after the wait limit and collect returns PM_ETIMEOUT.  One-shot. */
int pm_inject_io_stall(pm_context* ctx, int32_t chunk);
/* Same product with device-resident a, b, c on a HIP stream (0 = the
 * context's stream); returns once the work is enqueued. */
int pm_multiply_device(pm_context* ctx, uint32_t n, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                       void* hip_stream);
/* Allocation-free variant of pm_multiply_device (no counterpart in the
 * reference), for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops.  Same product and arguments; its
 * residues live in a static 16 MiB array of the library (one per device,
 * held until the process ends) instead of the context's workspace, so the
 * call makes no allocation.  Limit: 2 x k x n <= 4,194,304 residues (n <=
 * 32768 for every q; PM_EUNSUPPORTED beyond).  Every call on one GPU shares
 * that array: calls that overlap on a device (other streams, other
 * contexts, other threads) race and must be serialized by the caller. */
int pm_multiply_device_noalloc(pm_context* ctx, uint32_t n, const uint32_t* d_a, const uint32_t* d_b,
                               uint32_t* d_c, void* hip_stream);
/* Backend of the context (no counterpart as an API in the reference, which
 * picks its backend at build time, CMakeLists.txt:29-38): PM_BACKEND_GENERAL,
 * the RNS of any odd q (the default), or PM_BACKEND_QUARTIC for q = 2^868 -
 * delta, 0 < delta < 2^28 -- four 217-bit components and four channels per
 * prime (quartic/host_setup.h:41-110, quartic/solve.cu:289-316,906-1040),
 * whose tables are built by its first selection (PM_EUNSUPPORTED for any
 * other q).  Every later product of the context, and batches and servers
 * created afterwards, use it; both give the same words.  get: the backend and
 * its residue rows per polynomial (k primes, or 4 per quartic prime). */
#define PM_BACKEND_GENERAL 0u
#define PM_BACKEND_QUARTIC 1u
int pm_set_backend(pm_context* ctx, uint32_t backend);
int pm_get_backend(const pm_context* ctx, uint32_t* backend, uint32_t* rows);
/* Waits for the context's stream. */
int pm_synchronize(pm_context* ctx);
/* Plain-path schedule (no counterpart in the reference): the number of outer
 * radix-2 stages (0..2) run as separate passes before/after the fused tiles
 * for products of size n; the default is 0 up to N = 2^13, 1 at 2^14 and 2
 * above.  More outer stages give more, smaller tiles -- more blocks for a
 * single product.  Set by the autotuner in the Python mirror
 * (`autotune_plain`). */
int pm_set_outer(pm_context* ctx, uint32_t n, uint32_t outer);
int pm_get_outer(pm_context* ctx, uint32_t n, uint32_t* outer);
/* Batch schedule (no counterpart in the reference): outer radix-2 stages
 * (0..2, leaving a tile of 2^5..2^15) for products of size n in batches and
 * served launches created afterwards, kept per backend (set/get act on the
 * context's current one).  The general backend runs one or two outer stages
 * inside its conversion and CRT; its default leaves a 2^13 tile (0 stages up
 * to 2^13, 1 at 2^14, 2 above).  The quartic backend runs them as passes of
 * their own; its default is max(0, log2 n - 15), the largest tile.  Set by
 * `autotune_batch` in the Python mirror. */
int pm_set_batch_outer(pm_context* ctx, uint32_t n, uint32_t outer);
int pm_get_batch_outer(pm_context* ctx, uint32_t n, uint32_t* outer);
/* Batch schedule (no counterpart in the reference, whose CRT launches fixed
 * 128-thread blocks, quartic/solve.cu:1280,1299): waves per block of the
 * MFMA CRT in batches and served launches created afterwards -- 8 (the
 * default; 0 selects it) fill a compute unit with one block, 4 and 2 spread
 * a small batch over more compute units.  PM_EINVAL for other values.  Set by
 * `autotune_batch` in the Python mirror. */
int pm_set_batch_crt_waves(pm_context* ctx, uint32_t waves);
int pm_get_batch_crt_waves(pm_context* ctx, uint32_t* waves);
/* Batch schedule (no counterpart in the reference): the staggered parts
 * (1-4) a batch created afterwards splits into, each part on a graph branch
 * of its own that starts when the part before has converted its residues;
 * a part holds at least 4 products.  0 (the default) chooses per batch: 3
 * where some product runs tiles of 2^15 residues, else 1.  PM_EINVAL above
 * 4.  pm_batch_info reports a batch's count.  Set by `autotune_batch` in
 * the Python mirror. */
int pm_set_batch_parts(pm_context* ctx, uint32_t parts);
int pm_get_batch_parts(pm_context* ctx, uint32_t* parts);
/* Plain-path schedule (no counterpart in the reference; the role of its
 * primes-per-conversion-block constant, quartic/solve.cu:325): the slices of
 * the primes that the conversion of a product of size n spreads over
 * separate blocks (1, 2 or 4; 0, the default, chooses at launch: up to 4
 * while that leaves fewer than about 4 waves per compute unit).  The
 * host-buffer path (pm_multiply, pm_submit) converts as it ingests and does
 * not use it.  Set by `autotune_plain` in the Python mirror. */
int pm_set_conversion_slices(pm_context* ctx, uint32_t n, uint32_t slices);
int pm_get_conversion_slices(pm_context* ctx, uint32_t n, uint32_t* slices);
/* Batch schedule (no counterpart in the reference; the role of the
 * reference's compile-time tile and primes-per-block knobs,
 * quartic/host_setup.h:11-12, quartic/solve.cu:325): tiles of size tile_n
 * (a power of two in [32, 32768]) that one block holds, in batches and
 * served launches created afterwards.  A block has tiles x tile_n / r
 * threads (r residues each, pm_set_batch_residues_per_thread): a power of two
 * with 64 r <= tiles x tile_n <= 1024 r; 0 restores the default (one tile
 * per block from 2^13 up, 1024-thread blocks below, at r = 32).  Smaller
 * blocks let several run on one compute unit.  Set by `autotune_batch`. */
int pm_set_batch_tiles_per_block(pm_context* ctx, uint32_t tile_n, uint32_t tiles);
int pm_get_batch_tiles_per_block(pm_context* ctx, uint32_t tile_n, uint32_t* tiles);
/* Batch schedule (no counterpart in the reference; the role of its
 * compile-time registers-per-thread of the fused tile, quartic/solve.cu:576 and
 * quartic/host_setup.h:11-12): residues each thread of a tile of size tile_n
 * holds in registers, 32 (the default; 0 selects it) or 16 for tiles of
 * 2^6 .. 2^14 (twice the threads, half the registers, one more LDS exchange
 * per transform).  Setting it resets the size's tiles per block to that
 * choice's default (pm_set_batch_tiles_per_block then counts threads as tiles
 * x tile_n / residues).  Batches and served launches created afterwards use
 * it; the persistent mode keeps 32.  Set by `autotune_batch`. */
int pm_set_batch_residues_per_thread(pm_context* ctx, uint32_t tile_n, uint32_t residues);
int pm_get_batch_residues_per_thread(pm_context* ctx, uint32_t tile_n, uint32_t* residues);

/* ---- stage entry points (device buffers, context stream) ------------------
 * The three kernels of the plain path, exposed so each can be checked alone.
 * Residue buffers are [k][n] uint32, 16-byte aligned (PM_EINVAL otherwise). */
/* residues of a and b: ra[i*n+j] = a_j mod p_i (a representative below 4*p_i, the
 * NTT's lazy input range) -- rns/solve.cu:138-157 */
int pm_stage_convert(pm_context* ctx, uint32_t n, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_ra,
                     uint32_t* d_rb, void* hip_stream);
/* same contract by scalar limb reduction (sum of 28 Shoup products with
 * 2^(32j) mod p_i, rns/solve.cu:138-157; values below 2*p_i) -- the comparison
 * path of the MFMA conversion */
int pm_stage_convert_scalar(pm_context* ctx, uint32_t n, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_ra,
                            uint32_t* d_rb, void* hip_stream);
/* forward NTTs, pointwise, inverse, scale: ra <- N^-1 (M_i mod p_i)^-1 * (a*b mod p_i) --
 * rns/solve.cu:225-333 (here incomplete transforms down to blocks of 4, multiplied
 * modulo X^4 - zeta); outer = number of radix-2 stages run as separate passes */
int pm_stage_ntt(pm_context* ctx, uint32_t n, uint32_t outer, uint32_t* d_ra, uint32_t* d_rb, void* hip_stream);
/* CRT + mod q from scaled residues -- rns/solve.cu:336-420 (int8 MFMA sums) */
int pm_stage_crt(pm_context* ctx, uint32_t n, const uint32_t* d_t, uint32_t* d_c, void* hip_stream);
/* the same with the sums on the vector ALUs (comparison path, identical output) */
int pm_stage_crt_scalar(pm_context* ctx, uint32_t n, const uint32_t* d_t, uint32_t* d_c, void* hip_stream);

/* ---- batched products (no counterpart in the reference) --------------------
 * Many products of uneven N (powers of two, 32 <= N <= max_n) under the
 * context's q, run from a dependency-ordered work list whose launches are
 * captured once in a hipGraph and replayed by pm_batch_run.  With
 * d_a/d_b/d_c all non-NULL the caller owns the n_i*28-word device buffers of
 * product i; otherwise the batch allocates them (pm_batch_io returns them). */
typedef struct pm_batch pm_batch;
/* flags: PM_BATCH_PHASED (default, 0 also selects it) runs one launch per
 * phase -- conversion, outer stages, one per tile size class, CRT -- inside
 * the captured graph, each kernel register-allocated for its own phase.
 * PM_BATCH_PERSISTENT runs the same work list in one persistent launch: one
 * 512-thread block per compute unit loops over the conversion items, then the
 * tiles, then the CRT groups, with a grid barrier between phases (bounded
 * waits: a barrier that cannot complete makes pm_batch_wait return
 * PM_ETIMEOUT instead of hanging).  General backend, N up to 2^16 (tiles of at
 * most 2^14, outer stages inside conversion and CRT); PM_EUNSUPPORTED
 * otherwise.  Same words as the phased graph. */
#define PM_BATCH_PERSISTENT 1u
#define PM_BATCH_PHASED 2u
#define PM_BATCH_SCALAR_CONVERSION 4u /* phased: scalar limb reduction instead of int8 MFMA */
#define PM_BATCH_SCALAR_CRT 8u        /* phased: CRT sums on the vector ALUs instead of int8 MFMA */
/* phased: the outer NTT stages (N > tile size) as launches of their own over
 * the residues in HBM, instead of inside the conversion (forward) and the
 * MFMA CRT (inverse); implied by either scalar flag (comparison path) */
#define PM_BATCH_SEPARATE_OUTER 16u
int pm_batch_create(pm_context* ctx, uint32_t count, const uint32_t* ns, const uint32_t* const* d_a,
                    const uint32_t* const* d_b, uint32_t* const* d_c, uint32_t flags, pm_batch** out);
void pm_batch_destroy(pm_batch* batch);
int pm_batch_io(pm_batch* batch, uint32_t i, uint32_t** d_a, uint32_t** d_b, uint32_t** d_c);
/* synchronous host <-> device copies of product i's operands / result; each
 * first waits for the batch's last replay (on whichever stream it ran) */
int pm_batch_upload(pm_batch* batch, uint32_t i, const uint32_t* a, const uint32_t* b);
int pm_batch_download(pm_batch* batch, uint32_t i, uint32_t* c);
/* enqueue one replay (0 = context stream); pm_batch_wait synchronizes that
 * stream and reports a device error.  pm_batch_destroy waits for the last
 * replay, on whichever stream it was enqueued. */
int pm_batch_run(pm_batch* batch, void* hip_stream);
int pm_batch_wait(pm_batch* batch, void* hip_stream);
/* The batch's work-list items, the staggered parts of its graph and the
 * device memory it holds (any pointer may be null). */
int pm_batch_info(const pm_batch* batch, uint32_t* items, uint32_t* parts, uint64_t* device_bytes);
/* Runs the batch's phases once as direct launches on a stream with HIP events
 * between them and returns each phase's kind (0 conversion, 1 outer forward,
 * 2 | tlog<<8 tile class, 3 outer inverse, 4 CRT) and milliseconds.  On entry
 * max is the capacity of kinds/ms; *count returns the number of phases. */
int pm_batch_phase_times(pm_batch* batch, void* hip_stream, uint32_t max, uint32_t* kinds, double* ms,
                         uint32_t* count);
/* Host-only: the work list a batch of these sizes runs (k primes).  On entry
 * *n_items / *n_units hold the capacities of items (4 words each: type, job,
 * a0, a1) and units (2 words each: job, prime<<2|sub-block); on return the
 * sizes needed.  Arrays are filled only when large enough. */
int pm_plan_worklist(uint32_t k, uint32_t count, const uint32_t* ns, uint32_t* n_items, uint32_t* items,
                     uint32_t* n_units, uint32_t* units);

/* ---- serving (no counterpart in the reference) ------------------------------
 * Continuous batching: requests of any N (powers of two, 32 <= N <= max_n)
 * are submitted at any time; a native scheduler thread admits them in
 * arrival order while their device memory (3*N*112 + 2*k*N*4 bytes) fits
 * hbm_budget_bytes, groups what has been admitted into the next batch
 * launch (at most max_batch products), and keeps up to max_inflight launches
 * running concurrently on separate streams (each launch runs the phased
 * kernels of the batch path on the launch's own stream).  a, b, c are host buffers of
 * n*28 words that must stay valid until pm_server_wait returns for the id.
 * Page-locked a and b (both) are read, and a page-locked c written, by
 * kernels of the launch, so a launch of such requests is kernels only; other
 * buffers take runtime copies (pageable c after the launch completes). */
typedef struct pm_server pm_server;
int pm_server_create(pm_context* ctx, uint64_t hbm_budget_bytes, uint32_t max_batch, uint32_t max_inflight,
                     pm_server** out);
/* non-blocking; PM_EBUDGET if the request alone exceeds the budget */
int pm_server_submit(pm_server* server, uint32_t n, const uint32_t* a, const uint32_t* b, uint32_t* c,
                     uint64_t* id);
/* blocks until request id is done; latency = completion - submission */
int pm_server_wait(pm_server* server, uint64_t id, double* latency_seconds);
int pm_server_stats(pm_server* server, uint64_t* completed, uint64_t* launches, uint64_t* max_concurrent,
                    uint64_t* peak_bytes);
/* finishes queued work, then releases the server */
void pm_server_destroy(pm_server* server);

#ifdef __cplusplus
}
#endif
#endif
