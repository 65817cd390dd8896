// Device-side data shared by the host runtime and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace pmd {

constexpr uint32_t kLimbs = 28;
constexpr uint32_t kCrtLimbs = 29;
constexpr uint32_t kWaveSize = 64;
constexpr uint32_t kBlock = 1024;          // threads of an NTT tile block
constexpr uint32_t kTileElems = 32768;     // residues held by one NTT block (32 per thread)
constexpr uint32_t kLdsWords = kTileElems + kTileElems / 32;  // padded exchange buffer
constexpr uint32_t kConvPerWave = 64;      // coefficients per wave in the conversion
constexpr uint32_t kConvPerBlock = kConvPerWave * (kBlock / kWaveSize);  // 1024
// batch conversion: an item is 2048 coefficients of one polynomial, 16 slots
// of 128 that a wave converts as two halves of 64
constexpr uint32_t kConvSlotCoefs = 128;
constexpr uint32_t kConvItemCoefs = 2048;
constexpr uint32_t kConvSmallBlock = 256;  // threads of the plain-path conversion blocks
constexpr uint32_t kCrtPerBlock = 256;
// MFMA CRT: groups of 64 coefficients per wave, 8 waves per block; per wave
// the LDS holds 31 words (low 32 bits + signed high 16 bits) per coefficient
constexpr uint32_t kCrtWaveCoefs = 64;
constexpr uint32_t kCrtMfmaWaves = 8;   // default waves per block of the batched CRT (2, 4 or 8)
// waves per block of the single-product CRT: one product has few groups (512
// at N = 2^15), and blocks of 4 spread them over twice as many CUs
constexpr uint32_t kCrtPlainWaves = 4;
constexpr uint32_t kCrtWords = 31;
constexpr size_t kLdsPerCu = 160u * 1024u;  // gfx950
// words per row: a partial-sum store writes words 4r..4r+3 of 16 coefficients,
// lanes (g, col) at g*stride + col; stride = 16 mod 64 puts the 64 lanes on
// 64 distinct banks (and each 32-lane half on distinct banks of 32)
constexpr uint32_t kCrtLdsStride = 80;
constexpr uint32_t kCrtWaveLds = kCrtWords * kCrtLdsStride * 6;  // bytes: u32 low + i16 high

// Static residue scratch of the allocation-free plain product: ra and rb of
// up to 2 x 64 primes x 2^15 residues (16 MiB per device).
constexpr uint32_t kScratchWords = 2u * 64u * 32768u;

struct Tw {
    uint32_t value, shoup;
};

// Resident per-modulus tables (device pointers), passed by value to kernels.
struct Tables {
    const uint32_t* primes;       // [k]
    const uint32_t* mont;         // [k] -p^-1 mod 2^32
    const Tw* fwd;                // [k][max_n] psi^brv(i), value stored negated mod 2^32
    const Tw* inv;                // [k][max_n] psi^-brv(i)
    const Tw* invc;               // [k][max_n] p - psi^-brv(i): the inverse's complemented outputs
    const Tw* scale;              // [max_logn+1][k] the tile's scale (N >= 4: 4 times the post-inverse scale)
    const Tw* zs;                 // [k][2^(max_logn-1)] zeta_G * scale of block G at 2^(logn-2) + G
    const uint4* conv_w;          // [quads][2][64] int8x16 weight operands
    const int32_t* conv_corr;     // [quads*4][4]
    const uint32_t* conv_sh16;    // [k] floor(2^48/p)
    const Tw* limb_pow;           // [k][28] 2^(32j) mod p (scalar conversion)
    const uint32_t* crt_m30;      // [k][29]
    const uint32_t* crt_round;    // [k] floor(2^42/p)
    const uint32_t* crt_negkp;    // [2k+1][28]
    const uint4* crt_a;           // [8 row tiles][4 k steps][64 lanes] int8x16 MFMA operands
    const int4* crt_corr;         // [32]: row corrections of rows 4r..4r+3
    const uint32_t* q;            // [28]
    uint32_t mu0, mu1;            // floor(2^906/q)
    uint32_t k, quads, max_logn, cus;  // cus: compute units of the device
    // Quartic backend (q = 2^868 - qdelta): the rows above (primes, mont,
    // fwd, inv, invc, scale, zs, k) are then the 4 qk channels, row 4i + j channel j of
    // prime i, and these hold the per-prime constants of the conversion and
    // the CRT; quartic = 0 and null pointers for the general RNS.
    uint32_t quartic, qk, qdelta, pad_;
    const uint32_t* qprimes;      // [qk]
    const Tw* q_limb_pow;         // [qk][7] 2^(32 l) mod p
    const Tw* q_rpow;             // [qk][4] r^j
    const Tw* q_rinv;             // [qk][4] r^-j
    const Tw* q_iota;             // [qk] i, i^2 = -1
    const uint32_t* q_m30;        // [qk][18] M_i in 30-bit limbs
    const uint32_t* q_round;      // [qk] floor(2^42 / p_i)
    const uint32_t* q_kp30;       // [qk+1][18] K P in 30-bit limbs
    const uint32_t* q_negc;       // [28] (-qk P (1 + B + B^2 + B^3)) mod q
};
constexpr uint32_t kQuarticMaxPrimes = 17;
constexpr uint32_t kQuarticLimbs30 = 18;

// One product of the batch / the plain path.
struct Job {
    uint32_t logn;
    uint32_t outer;      // radix-2 stages run as separate passes (tile = 2^(logn-outer))
    const uint32_t* a;   // [n][28] coefficients (device)
    const uint32_t* b;
    uint32_t* c;         // [n][28] result (device)
    uint32_t* ra;        // [k][n] residues of a, reused for CRT inputs
    uint32_t* rb;        // [k][n] residues of b
    // 1: the outer stages (outer = 1 or 2, n >= 256) run inside the
    // conversion (forward) and the CRT (inverse) instead of as passes of
    // their own; both then take a group's four column tiles of 16
    // coefficients a quarter of n apart, so each lane holds a whole radix-4
    // column of its prime
    uint32_t fused;
    // plain-path conversion: slices of the primes over blockIdx.z (1, 2 or
    // 4), 0 for the launch's own choice
    uint32_t conv_slices;
};

// whether a product of 2^logn with `outer` outer stages can fuse them
constexpr bool outer_fusable(uint32_t logn, uint32_t outer) { return outer >= 1 && outer <= 2 && logn >= 8; }

// Batch work list: the items of each phase (conversion groups, outer-stage
// columns, tiles, CRT groups); a tile item covers one or more units.
enum ItemType : uint32_t { kItemConv = 0, kItemOuterFwd = 1, kItemTile = 2, kItemOuterInv = 3, kItemCrt = 4 };

struct Item {
    uint32_t type;
    uint32_t job;   // product (tile items: unused)
    uint32_t a0;    // conv: poly; outer fwd: poly*k+prime; tile: tlog | kItemRpt16? | count<<8; outer inv: prime; crt: block
    uint32_t a1;    // conv: coefficient block; tile: first unit
};

struct Unit {
    uint32_t job;
    uint32_t prime_s;  // prime << 2 | sub-block
};

struct BatchArgs {
    const Job* jobs;
    const Item* items;
    const Unit* units;
    uint32_t n_items;
};

constexpr uint32_t kErrTimeout = 2;

// Batch tiles at 16 residues per thread (a schedule choice beside the default
// 32) are compiled for tile sizes 2^6 .. 2^14; a block holds at most 1024
// threads, so 2^14 residues.  A tile item's a0 carries the choice in bit 4.
constexpr int kRpt16MinLog = 6;
constexpr int kRpt16MaxLog = 14;
constexpr uint32_t kItemRpt16 = 1u << 4;

// Persistent batch (one launch for the whole work list): one block of 512
// threads per compute unit; tiles up to 2^14 residues (the block's 16384
// residues); two grid barriers, their arrival counters and a shared error
// word in device memory, cleared by a kernel ahead of each launch.
constexpr uint32_t kPersistentThreads = 512;
constexpr int kPersistentMaxTileLog = 14;
constexpr uint32_t kPersistentBarriers = 2;
struct GridSync {
    uint32_t* counters;   // [kPersistentBarriers]
    uint32_t* error;      // kErrTimeout when a barrier wait gave up
    uint64_t wait_ticks;  // bound of one barrier wait (wall-clock ticks)
};

// Host-buffer path: flags in page-locked host memory (fine-grained) that the
// host sets as input chunks land in staging (ready) and the device sets as
// result chunks land in staging (done); counters (device memory) count the
// finished pieces of each result chunk.
constexpr uint32_t kMaxIoChunks = 16;
constexpr uint32_t kMirrorStride = 64;  // words between the device copies of the ready flags
// One span of a gather copy: words 32-bit words from src to dst (either may
// be page-locked host memory).
struct CopySeg {
    const uint32_t* src;
    uint32_t* dst;
    uint64_t words;
};

struct IoArgs {
    const uint32_t* ready;      // [kMaxIoChunks] host-written
    uint32_t* done;             // [kMaxIoChunks] device-written
    uint32_t* error;            // host-mapped: kErrTimeout when a wait gave up
    uint32_t* counters;         // [kMaxIoChunks] device memory
    uint32_t* mirror;           // [kMaxIoChunks + 1][kMirrorStride] device memory: ready flags as seen, then an error word
    uint32_t seq;               // this product's sequence number (never 0)
    uint32_t chunks;            // input chunks
    uint32_t groups_per_chunk;  // ingest: 64-coefficient groups per input chunk
    uint32_t pad;
    uint64_t wait_ticks;        // bound of one wait (wall-clock ticks)
};

}  // namespace pmd
