// gfx950 kernels for the exact 868-bit negacyclic product over a 30-bit RNS.
//
//  * residue conversion: exact int8 x int8 -> int32 MFMA contraction of the
//    coefficient bytes against balanced base-256 digits of 256^j mod p_i
//    (role of rns/tensor_prepare.h:20-71 and rns/solve.cu:138-157);
//  * fused per-prime convolution tile: Harvey lazy odd-root Cooley-Tukey
//    forward NTTs of a and b stopped at blocks of 4 residues, the scaled
//    block product modulo X^4 - zeta, the Gentleman-Sande inverse (on
//    complemented operands where the stage's mode is known at compile time)
//    -- spectra stay in registers/LDS (role of quartic/solve.cu:641-770 and
//    rns/solve.cu:225-333);
//  * outer radix-2/radix-4 stages of N > 2^13 tiles, fused into the
//    conversion and the CRT or as passes of their own (role of
//    quartic/solve.cu:771-817);
//  * exact CRT on int8 MFMA (or 30-bit limb accumulation), nearest-integer
//    quotient and a Barrett reduction modulo any q > 2^864 (role of
//    rns/solve.cu:336-420).
// All arithmetic is exact 32/64-bit integer arithmetic.
#include <hip/hip_runtime.h>

#include "pm_device.h"
#include "pm_kernels.h"

namespace pmd {

typedef int v4i __attribute__((ext_vector_type(4)));

// A pointer the caller knows is the same in every lane of the wave, moved to
// scalar registers (a buffer descriptor must be wave-uniform).
template <class T>
__device__ __forceinline__ T* wave_uniform(T* ptr) {
    const uint64_t v = reinterpret_cast<uint64_t>(ptr);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v)), hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<T*>(uint64_t(lo) | (uint64_t(hi) << 32));
}

// Buffer resources (V#) for raw buffer loads and stores: 32-bit data format,
// no swizzle, no stride (gfx9 V# word 3); loads past the range read zero.
constexpr uint32_t kRawBufferWord3 = 0x00020000u;
constexpr uint32_t kOutOfRange = 0x80000000u;  // an offset past any buffer range used here
// Cache policy of the fused tile's residue loads and result stores: nt
// (streaming).  Each residue is read once and each result is read once more,
// by the CRT; streaming them keeps the L2 for the twiddle tables that the
// blocks of one prime share.  At 128 x N = 2^15 the tile phase drops from
// 1.72 to 1.66 ms; nt on the conversion's 64-byte store runs or on the CRT's
// loads and 16-byte output stores is slower.
constexpr int kTileStreamPolicy = 2;

// ---------------------------------------------------------------------------
// small-modulus primitives (p < 2^30)
// ---------------------------------------------------------------------------
// a*w mod p, result in [0,2p), for any 32-bit a (Shoup).
__device__ __forceinline__ uint32_t shoup(uint32_t a, Tw w, uint32_t p) {
    const uint32_t quotient = __umulhi(a, w.shoup);
    return a * w.value - quotient * p;
}
// The same with np = 2^32 - p, which the caller keeps opaque to the optimiser
// (an empty asm): a*w + q*np is one multiply and one multiply-add (the low
// half of v_mad_u64_u32) instead of two multiplies and a subtract.
__device__ __forceinline__ uint32_t shoup_np(uint32_t a, Tw w, uint32_t np) {
    return a * w.value + __umulhi(a, w.shoup) * np;
}
// x < 2m -> x mod m   (x - m wraps above x when x < m)
__device__ __forceinline__ uint32_t csub(uint32_t x, uint32_t m) {
    const uint32_t t = x - m;
    return t < x ? t : x;
}
// Cooley-Tukey: X in [0,4p), Y any -> both in [0,4p).  The Shoup product
// t = Y*w mod p in [0,2p) is carried negated, nt = q*p - Y*w (mod 2^32), so
// the outputs are x - nt and x + 2p + nt: one three-input add instead of an
// add and a subtract.  Forward twiddles are stored as w.value = -w mod 2^32,
// so nt = q*p + Y*w.value is one multiply and one multiply-add (the low half
// of v_mad_u64_u32): 7 VALU instructions per butterfly.  The three-input add
// is written out because the compiler, left to itself, reassociates it into
// code that needs more registers (and spills at 128 VGPRs).
__device__ __forceinline__ void ct_bfly(uint32_t& X, uint32_t& Y, Tw w, uint32_t p, uint32_t p2) {
    const uint32_t x = csub(X, p2);
    const uint32_t quotient = __umulhi(Y, w.shoup);
    const uint32_t nt = quotient * p + Y * w.value;
    uint32_t y;
    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(y) : "v"(x), "v"(p2), "v"(nt));
    X = x - nt;
    Y = y;
}
// Gentleman-Sande: X, Y in [0,2p) -> both in [0,2p).  The Shoup product is
// formed as d*w + q*(2^32 - p): one multiply and one multiply-add (the low
// half of v_mad_u64_u32), 8 VALU instructions per butterfly.  2^32 - p is
// passed through an empty asm so that the optimiser cannot see it is -p and
// fold the sum back into two multiplies and a subtract.
__device__ __forceinline__ void gs_bfly(uint32_t& X, uint32_t& Y, Tw w, uint32_t p, uint32_t p2) {
    const uint32_t s = X + Y;
    const uint32_t d = X - Y + p2;
    X = csub(s, p2);
    uint32_t np = 0u - p;
    asm("" : "+v"(np));
    const uint32_t quotient = __umulhi(d, w.shoup);
    Y = d * w.value + quotient * np;
}
// Gentleman-Sande on complemented operands.  In the tile's inverse a value
// whose position is the multiplied (Y) operand of the next stage can be held
// complemented, Yc = 2p - Y in [0,2p] (congruent to -Y), so that the stage
// forms X - Y + 2p = X + Yc with one add and X + Y = X - Yc (+2p) with one
// subtract and a min (7 instructions instead of 8).  The outputs are normal
// (X', Y' in [0,2p)) when the next stage takes them as they are, or
// complemented when they are the next stage's Y operands (both outputs of a
// butterfly share that bit of their position).  wc is the complement twiddle
// (p - w, its Shoup quotient): Shoup(d, p - w) is congruent to -d*w.
// Yc in [0,2p], X in [0,2p) -> X', Y' in [0,2p)
__device__ __forceinline__ void gs_bfly_cn(uint32_t& X, uint32_t& Y, Tw w, uint32_t p, uint32_t p2) {
    const uint32_t e = X - Y;  // X - Yc in [-2p, 2p), congruent to X + Y
    const uint32_t d = X + Y;  // X - Y + 2p
    uint32_t np = 0u - p;
    asm("" : "+v"(np));
    X = min(e, e + p2);
    Y = d * w.value + __umulhi(d, w.shoup) * np;
}
// Yc in [0,2p], X in [0,2p) -> Xc', Yc' in [0,2p]
__device__ __forceinline__ void gs_bfly_cc(uint32_t& X, uint32_t& Y, Tw wc, uint32_t p, uint32_t p2) {
    const uint32_t f = Y - X;  // Yc - X in (-2p, 2p], congruent to -(X + Y)
    const uint32_t d = X + Y;
    uint32_t np = 0u - p;
    asm("" : "+v"(np));
    X = min(f, f + p2);
    Y = d * wc.value + __umulhi(d, wc.shoup) * np;
}
// X, Y in [0,2p) -> Xc', Yc' in [0,2p]
__device__ __forceinline__ void gs_bfly_nc(uint32_t& X, uint32_t& Y, Tw wc, uint32_t p, uint32_t p2) {
    const uint32_t u = 2u * p2 - (X + Y);  // 4p - (X + Y) in (0, 4p]
    const uint32_t d = X - Y + p2;
    uint32_t np = 0u - p;
    asm("" : "+v"(np));
    X = csub(u, p2);
    Y = d * wc.value + __umulhi(d, wc.shoup) * np;
}
// T*2^-32 mod p for a 64-bit T below 8p^2 (Montgomery REDC): m = T*(-p^-1)
// mod 2^32, (T + m*p) / 2^32 with T + m*p < 8p^2 + 2^32 p < 2^64 (p < 2^30),
// one multiply and one v_mad_u64_u32; the result is below 8p^2/2^32 + p < 3p.
__device__ __forceinline__ uint32_t redc64(uint64_t T, uint32_t p, uint32_t neg_inv) {
    const uint32_t m = uint32_t(T) * neg_inv;
    return uint32_t((uint64_t(m) * p + T) >> 32);
}
// a*b*2^-32 mod p in [0,2p) for a, b in [0,2p) (Montgomery REDC).
// (a*b + m*p) / 2^32 with m = lo * (-p^-1): the high half of m*p + lo is
// umulhi(m, p) plus the carry of lo + (m*p mod 2^32), one v_mad_u64_u32.
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t p, uint32_t neg_inv) {
    const uint32_t lo = a * b, hi = __umulhi(a, b);
    const uint32_t m = lo * neg_inv;
    return hi + uint32_t((uint64_t(m) * p + lo) >> 32);
}

// ---------------------------------------------------------------------------
// NTT tile geometry: a tile of T = 2^TLOG residues is held by 2^(TLOG-ELOG)
// threads with E = 2^ELOG registers each.  The TLOG index bits are split into
// NPH phases of (nearly) equal width; in phase P a thread's registers span
// that phase's bits (plus filler bits when the width is below ELOG), so every
// butterfly of the phase is register-local.  Phases are joined by an exchange
// through padded LDS (lds_pad).
// ---------------------------------------------------------------------------
// log2 of the residues per thread of a tile of 2^tlog: 32 per thread, in
// 1024-thread blocks of 128 VGPRs.  (64 per thread in 512-thread blocks of
// 256 VGPRs measured 26 % slower on the 2^15 tile: half the waves per CU.)
__host__ __device__ constexpr int tile_elog(int tlog) { return tlog < 5 ? tlog : 5; }
// threads of a block that holds one 2^15-residue tile (or several smaller ones)
__host__ __device__ constexpr uint32_t tile_block_threads(int tlog) {
    return tlog < 5 ? kBlock : (kTileElems >> tile_elog(tlog));
}

// LDS address of tile index x: x + x/32 spreads every phase's half-wave
// accesses over the 32 banks except in the 2^13 tile, whose 13 index bits
// split 5 + 4 + 4: there phase 1's lanes hold index bits {1,2,3,8,9} and
// x/32 puts bits 3 and 8 on the same bank.  x + 2(x/32) + x/512 gives the
// bits of each phase weights of distinct 2-adic valuation mod 32 (phase 0:
// bits 0-4; phase 1: 2, 4, 8, 272, 545; phase 2: 34, 68, 136, 272, 545), so
// no exchange of that tile conflicts (checked on the CPU over every exchange
// and register).  Both forms are additive over disjoint bit sets, so a
// thread's base address plus a register's constant offset still gives the
// slot.  (For 2^11 and 2^12 no form with two such terms is conflict-free.)
template <int TLOG, int EL>
__host__ __device__ constexpr uint32_t lds_pad(uint32_t x) {
    if constexpr (TLOG == 13 && EL == 5) return x + ((x >> 5) << 1) + (x >> 9);
    else return x + (x >> 5);
}

template <int TLOG, int EL = tile_elog(TLOG)>
struct Geo {
    static_assert(EL >= 1 && EL <= 5 && EL <= TLOG, "residues per thread: 2^1 .. 2^5, at most the tile");
    static constexpr int ELOG = EL;
    static constexpr int E = 1 << ELOG;
    static constexpr int TPTLOG = TLOG - ELOG;
    static constexpr int TPT = 1 << TPTLOG;
    static constexpr int NPH = (TLOG + ELOG - 1) / ELOG;
    static constexpr int T = 1 << TLOG;
    static constexpr int STRIDE = int(lds_pad<TLOG, EL>(uint32_t(T - 1))) + 1;  // LDS words of one tile
    static constexpr int width(int p) { return TLOG / NPH + (p < TLOG % NPH ? 1 : 0); }
    static constexpr int lo(int p) {
        int h = TLOG;
        for (int i = 0; i <= p; ++i) h -= width(i);
        return h;
    }
    static constexpr bool in_range(int p, int pos) { return pos >= lo(p) && pos < lo(p) + width(p); }
    static constexpr int reg_pos(int p, int j) {
        if (j < width(p)) return lo(p) + j;
        int need = j - width(p);
        for (int pos = 0; pos < TLOG; ++pos) {
            if (in_range(p, pos)) continue;
            if (need == 0) return pos;
            --need;
        }
        return -1;
    }
    static constexpr bool is_reg_pos(int p, int pos) {
        for (int j = 0; j < ELOG; ++j)
            if (reg_pos(p, j) == pos) return true;
        return false;
    }
    static constexpr int thr_pos(int p, int j) {
        int need = j;
        for (int pos = 0; pos < TLOG; ++pos) {
            if (is_reg_pos(p, pos)) continue;
            if (need == 0) return pos;
            --need;
        }
        return -1;
    }
    // the register bit holding index bit pos in phase p, or -1 (a thread bit)
    static constexpr int reg_bit_of(int p, int pos) {
        for (int j = 0; j < ELOG; ++j)
            if (reg_pos(p, j) == pos) return j;
        return -1;
    }
    // the phase whose stages run on index bit pos
    static constexpr int phase_of(int pos) {
        for (int p = 0; p < NPH; ++p)
            if (in_range(p, pos)) return p;
        return -1;
    }
    static constexpr uint32_t reg_scatter(int p, int r) {
        uint32_t x = 0;
        for (int j = 0; j < ELOG; ++j)
            if ((r >> j) & 1) x |= 1u << reg_pos(p, j);
        return x;
    }
};

// a thread's registers of one operand of a tile
template <int TLOG, int EL>
using TileRegs = uint32_t[Geo<TLOG, EL>::E];

template <int TLOG, int EL, int P>
__device__ __forceinline__ uint32_t thr_scatter(uint32_t tau) {
    using G = Geo<TLOG, EL>;
    uint32_t x = 0;
#pragma unroll
    for (int j = 0; j < G::TPTLOG; ++j) x |= ((tau >> j) & 1u) << G::thr_pos(P, j);
    return x;
}

constexpr uint32_t kLdsHalf = 16384;  // words: 64 KB, the reach of an LDS immediate offset

// Butterflies between scheduling barriers, which bound the twiddles in flight
// (VGPR budget of 1024-thread blocks): measured best at 16 for the 2^15 tile,
// while the smaller tiles spill with more than 8.
template <int TLOG, int EL>
constexpr int twiddle_chunk() { return TLOG == 15 && EL == 5 ? 16 : 8; }
template <int TLOG, int EL>
constexpr int inverse_twiddle_chunk() { return 2 * twiddle_chunk<TLOG, EL>(); }

// An exchange keeps every element inside its wave when the wave bits of the
// thread index (thread bits 6 and up) sit on the same index bits in both
// phases; then no block barrier is needed: LDS instructions of one wave
// execute in order, so its reads see its own writes.
template <int TLOG, int EL>
__host__ __device__ constexpr bool wave_local(int pfrom, int pto) {
    using G = Geo<TLOG, EL>;
    for (int j = 6; j < G::TPTLOG; ++j)
        if (G::thr_pos(pfrom, j) != G::thr_pos(pto, j)) return false;
    return true;
}

// LEAD: a wave-local exchange came before this block-wide one, so other waves
// may still be reading their part of the buffer.
template <int TLOG, int EL, int PFROM, int PTO, bool LEAD>
__device__ __forceinline__ void exchange(TileRegs<TLOG, EL>& v, uint32_t* lds, uint32_t tau) {
    using G = Geo<TLOG, EL>;
    constexpr bool kWave = wave_local<TLOG, EL>(PFROM, PTO);
    // thread and register bits are disjoint, so lds_pad(x|c) = lds_pad(x) + lds_pad(c):
    // one base address per thread plus immediate offsets.  An LDS immediate
    // offset is below 64 KB, so slots above that use a second base 64 KB up,
    // formed in this exchange (opaque to the optimiser): left to itself the
    // compiler gives each such slot an address register of its own, shared
    // by every exchange of the same layout, and spills them across the tile.
    const uint32_t wx = lds_pad<TLOG, EL>(thr_scatter<TLOG, EL, PFROM>(tau)),
                   rx = lds_pad<TLOG, EL>(thr_scatter<TLOG, EL, PTO>(tau));
    uint32_t wx_hi = wx + kLdsHalf, rx_hi = rx + kLdsHalf;
    if constexpr (G::STRIDE > kLdsHalf) asm volatile("" : "+v"(wx_hi), "+v"(rx_hi));
    const auto slot = [&](uint32_t x, uint32_t x_hi, uint32_t c) -> uint32_t& {
        return c < kLdsHalf ? lds[x + c] : lds[x_hi + (c - kLdsHalf)];
    };
    if constexpr (LEAD && !kWave) __syncthreads();
#pragma unroll
    for (int r = 0; r < G::E; ++r) slot(wx, wx_hi, lds_pad<TLOG, EL>(G::reg_scatter(PFROM, r))) = v[r];
    if constexpr (kWave) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < G::E; ++r) v[r] = slot(rx, rx_hi, lds_pad<TLOG, EL>(G::reg_scatter(PTO, r)));
    if constexpr (kWave) {
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// Twiddle tables are read-only while any kernel runs: read through the
// constant address space, so that uniform indices become scalar loads (the
// compiler cannot otherwise rule out that a table aliases the tile's own
// stores) and hold no vector registers.
typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
__device__ __forceinline__ Tw const_twiddle(const Tw* tw, uint32_t i) {
    const ConstWords e = (ConstWords)(tw + i);
    return Tw{e[0], e[1]};
}

// A work-list entry (item, product) at a wave-uniform index, read through the
// constant address space so that it lands in scalar registers: the lists are
// read-only while any kernel runs.
template <class T>
__device__ __forceinline__ T const_entry(const T* list, uint32_t i) {
    static_assert(sizeof(T) % 4 == 0, "whole words");
    const ConstWords w = (ConstWords)(list + i);
    T out;
    uint32_t* o = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (uint32_t j = 0; j < sizeof(T) / 4; ++j) o[j] = w[j];
    return out;
}

// Group index (index bits above beta) of register r in phase P.  When every
// thread bit of the phase lies below beta+1 the group -- hence the twiddle --
// is the same for all threads, which lets the compiler use scalar loads.
template <int TLOG, int EL, int P>
__device__ __forceinline__ uint32_t group_of(uint32_t tau_x, int r, int beta) {
    using G = Geo<TLOG, EL>;
    bool thread_bits_above = false;
    for (int j = 0; j < G::TPTLOG; ++j)
        if (G::thr_pos(P, j) > beta) thread_bits_above = true;
    const uint32_t reg_part = G::reg_scatter(P, r) >> (beta + 1);
    // thread and register bits are disjoint: + is |, and folds into offsets
    return thread_bits_above ? ((tau_x >> (beta + 1)) + reg_part) : reg_part;
}

// The tile's transforms are incomplete: the forward stages stop at span 4,
// where each block of 4 positions holds its operand modulo X^4 - zeta, and
// the inverse starts there (block_product between them).  Stages on index
// bits below kBlockLog are not run; a tile below 4 residues runs them all.
constexpr int kBlockLog = 2;
template <int TLOG>
__host__ __device__ constexpr int first_stage_bit() { return TLOG >= kBlockLog ? kBlockLog : 0; }

// Forward CT stages of phase P on two polynomials that share twiddles.
// Twiddle of local stage (m groups, group g) is table[bm*m + g] with
// bm = 2^u + s for sub-block s of a transform with u outer stages.
template <int TLOG, int EL, int P>
__device__ __forceinline__ void fwd_phase2(TileRegs<TLOG, EL>& a, TileRegs<TLOG, EL>& b,
                                           const Tw* __restrict__ tw, uint32_t bm, uint32_t tau_x,
                                           uint32_t p, uint32_t p2) {
    using G = Geo<TLOG, EL>;
#pragma unroll
    for (int bit = G::width(P) - 1; bit >= 0; --bit) {
        const int beta = G::lo(P) + bit;
        if (beta < first_stage_bit<TLOG>()) continue;
        const uint32_t m = 1u << (TLOG - 1 - beta);
        int done = 0;
#pragma unroll
        for (int r = 0; r < G::E; ++r) {
            if (r & (1 << bit)) continue;
            const Tw w = const_twiddle(tw, bm * m + group_of<TLOG, EL, P>(tau_x, r, beta));
            ct_bfly(a[r], a[r | (1 << bit)], w, p, p2);
            ct_bfly(b[r], b[r | (1 << bit)], w, p, p2);
            if (++done % twiddle_chunk<TLOG, EL>() == 0) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Complemented operands in the inverse (gs_bfly_cn): stage beta leaves the
// positions with bit beta+1 set complemented when that bit is a register bit
// of its phase, so the butterfly knows its mode at compile time (both outputs
// share the bit); the last stage and a stage whose next bit is a thread bit
// leave normal values.  The block product complements the positions with bit
// 2 set for the first stage the same way.
template <int TLOG, int EL>
__host__ __device__ constexpr bool inv_comp_after(int beta) {
    using G = Geo<TLOG, EL>;
    if (TLOG <= kBlockLog || beta + 1 >= TLOG) return false;
    return G::reg_bit_of(G::phase_of(beta), beta + 1) >= 0;
}
template <int TLOG, int EL>
__host__ __device__ constexpr bool inv_comp_before(int beta) {
    using G = Geo<TLOG, EL>;
    if (TLOG <= kBlockLog) return false;
    if (beta == kBlockLog) return G::reg_bit_of(G::NPH - 1, kBlockLog) >= 0;
    return inv_comp_after<TLOG, EL>(beta - 1);
}

// The stages of inverse phase P from register bit BIT on, one template level
// per stage, so that each stage's operand modes are compile-time constants.
template <int TLOG, int EL, int P, int BIT>
__device__ __forceinline__ void inv_stages(TileRegs<TLOG, EL>& a, const Tw* __restrict__ tw,
                                           const Tw* __restrict__ twc, uint32_t bm, uint32_t tau_x, uint32_t p,
                                           uint32_t p2) {
    using G = Geo<TLOG, EL>;
    if constexpr (BIT < G::width(P)) {
        constexpr int beta = G::lo(P) + BIT;
        if constexpr (beta >= first_stage_bit<TLOG>()) {
            constexpr bool cin = inv_comp_before<TLOG, EL>(beta);
            constexpr int jn = inv_comp_after<TLOG, EL>(beta) ? G::reg_bit_of(P, beta + 1) : -1;
            const uint32_t m = 1u << (TLOG - 1 - beta);
            int done = 0;
#pragma unroll
            for (int r = 0; r < G::E; ++r) {
                if (r & (1 << BIT)) continue;
                const uint32_t idx = bm * m + group_of<TLOG, EL, P>(tau_x, r, beta);
                if (jn >= 0 && ((r >> jn) & 1)) {
                    const Tw wc = const_twiddle(twc, idx);
                    if constexpr (cin)
                        gs_bfly_cc(a[r], a[r | (1 << BIT)], wc, p, p2);
                    else
                        gs_bfly_nc(a[r], a[r | (1 << BIT)], wc, p, p2);
                } else {
                    const Tw w = const_twiddle(tw, idx);
                    if constexpr (cin)
                        gs_bfly_cn(a[r], a[r | (1 << BIT)], w, p, p2);
                    else
                        gs_bfly(a[r], a[r | (1 << BIT)], w, p, p2);
                }
                if (++done % inverse_twiddle_chunk<TLOG, EL>() == 0) __builtin_amdgcn_sched_barrier(0);
            }
        }
        inv_stages<TLOG, EL, P, BIT + 1>(a, tw, twc, bm, tau_x, p, p2);
    }
}

template <int TLOG, int EL, int P>
__device__ __forceinline__ void inv_phase(TileRegs<TLOG, EL>& a, const Tw* __restrict__ tw,
                                          const Tw* __restrict__ twc, uint32_t bm, uint32_t tau_x, uint32_t p,
                                          uint32_t p2) {
    using G = Geo<TLOG, EL>;
    // the inverse table is indexed like the forward one: without this the
    // compiler keeps the forward phase's twiddle offsets live across the tile
    // (and spills them) instead of recomputing them here
    if constexpr (G::TPTLOG >= 6)
        asm volatile("" : "+v"(tau_x));  // bm is scalar here (conv_tile)
    else
        asm volatile("" : "+v"(tau_x), "+v"(bm));
    inv_stages<TLOG, EL, P, 0>(a, tw, twc, bm, tau_x, p, p2);
}

template <int TLOG, int EL, int P>
__device__ __forceinline__ void forward_from(TileRegs<TLOG, EL>& a, TileRegs<TLOG, EL>& b,
                                             uint32_t* lds, uint32_t tau, const Tw* tw, uint32_t bm,
                                             uint32_t p, uint32_t p2) {
    using G = Geo<TLOG, EL>;
    if constexpr (P > 0) {
        constexpr bool lead = P >= 2 && wave_local<TLOG, EL>(P - 2, P - 1);
        exchange<TLOG, EL, P - 1, P, lead>(a, lds, tau);
        exchange<TLOG, EL, P - 1, P, false>(b, lds, tau);
    }
    fwd_phase2<TLOG, EL, P>(a, b, tw, bm, thr_scatter<TLOG, EL, P>(tau), p, p2);
    if constexpr (P + 1 < G::NPH) forward_from<TLOG, EL, P + 1>(a, b, lds, tau, tw, bm, p, p2);
}

template <int TLOG, int EL, int P>
__device__ __forceinline__ void inverse_from(TileRegs<TLOG, EL>& a, uint32_t* lds, uint32_t tau,
                                             const Tw* tw, const Tw* twc, uint32_t bm, uint32_t p, uint32_t p2) {
    using G = Geo<TLOG, EL>;
    if constexpr (P + 1 < G::NPH) {
        // the exchange before this one: the inverse step above, or the last
        // forward step when this is the first inverse exchange
        constexpr bool lead =
            P + 2 < G::NPH ? wave_local<TLOG, EL>(P + 2, P + 1) : wave_local<TLOG, EL>(P, P + 1);
        exchange<TLOG, EL, P + 1, P, lead>(a, lds, tau);
    }
    inv_phase<TLOG, EL, P>(a, tw, twc, bm, thr_scatter<TLOG, EL, P>(tau), p, p2);
    if constexpr (P > 0) inverse_from<TLOG, EL, P - 1>(a, lds, tau, tw, twc, bm, p, p2);
}

// Product of the blocks of the incomplete transforms, in the last forward
// phase's registers: r..r+3 (r = 0 mod 4) are positions x..x+3 of one block,
// a and b modulo X^4 - zeta.  With A_i = a_i reduced to [0,p), B_j = b_j S
// and Z_j = b_j zeta S (Shoup, [0,2p); S the scale, zrow[x/4] = zeta S),
// coefficient k of the product is REDC of
//   sum_{i<=k} A_i B_{k-i} + sum_{i>k} A_i Z_{k-i+4}  (below 4 p 2p = 8p^2),
// in [0,3p), and after one conditional subtraction in [0,2p), the inverse
// stages' input range.  16 v_mad_u64_u32 per block; the scale of the result
// is folded in, so the store multiplies by nothing.
template <int TLOG, int EL>
__device__ __forceinline__ void block_product(TileRegs<TLOG, EL>& a, const TileRegs<TLOG, EL>& b, const Tw* zrow,
                                              uint32_t x_thr, Tw S, uint32_t p, uint32_t p2, uint32_t neg_inv) {
    using G = Geo<TLOG, EL>;
    constexpr int P = G::NPH - 1;
    static_assert(G::reg_pos(P, 0) == 0 && G::reg_pos(P, 1) == 1,
                  "the last phase holds index bits 0 and 1 in register bits 0 and 1");
    uint32_t np = 0u - p;
    asm volatile("" : "+v"(np));
#pragma unroll
    for (int r = 0; r < G::E; r += 4) {
        const Tw z = const_twiddle(zrow, (x_thr + G::reg_scatter(P, r)) >> 2);
        uint32_t A[4], B[4], Z[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            A[i] = csub(csub(a[r + i], p2), p);
            B[i] = shoup_np(b[r + i], S, np);
            Z[i] = i ? shoup_np(b[r + i], z, np) : 0u;
        }
        // positions with bit 2 set are the first inverse stage's Y operands:
        // complemented, csub(3p - t, p) in [0,2p] for t in [0,3p)
        constexpr int j2 = G::reg_bit_of(P, kBlockLog);
        const bool comp = inv_comp_before<TLOG, EL>(kBlockLog) && j2 >= 0 && ((r >> j2) & 1);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint64_t T = uint64_t(A[0]) * B[k];
#pragma unroll
            for (int i = 1; i < 4; ++i) T += uint64_t(A[i]) * (i <= k ? B[k - i] : Z[k - i + 4]);
            const uint32_t t = redc64(T, p, neg_inv);
            a[r + k] = comp ? csub(3u * p - t, p) : csub(t, p);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// One tile: forward(a) and forward(b) down to blocks of 4, the scaled block
// product, the inverse, the store ([0,p) for a whole transform).  (A tile
// below 4 residues runs complete transforms, a Montgomery pointwise product
// and the scale at the store.)  `valid` false means this thread's tile is padding: it joins every barrier
// but touches no global memory.
template <int TLOG, int EL = tile_elog(TLOG)>
__device__ __forceinline__ void conv_tile(const Tables& tb, const uint32_t* __restrict__ a_src,
                                          const uint32_t* __restrict__ b_src, uint32_t* __restrict__ dst,
                                          uint32_t prime, uint32_t bm, uint32_t logn_full,
                                          uint32_t* lds, uint32_t tau, bool valid) {
    using G = Geo<TLOG, EL>;
    uint32_t pi = valid ? prime : 0u;
    if constexpr (G::TPTLOG >= 6) {
        // a tile that spans whole waves has one prime and sub-block per wave:
        // scalar registers for them and all that follows (moduli, table bases)
        pi = __builtin_amdgcn_readfirstlane(pi);
        bm = __builtin_amdgcn_readfirstlane(bm);
    }
    const uint32_t p = tb.primes[pi], p2 = 2u * p, neg_inv = tb.mont[pi];
    const size_t table = size_t(pi) << tb.max_logn;
    uint32_t a[G::E], b[G::E];
    const uint32_t x0 = thr_scatter<TLOG, EL, 0>(tau);
    // Loads issued in the order the first forward stage pairs them (register r
    // with r + E/2, a then b), so that stage starts as the first loads land.
    // Unconditional (a padding tile points at a real tile of the same size and
    // only its stores are skipped).  When a tile spans whole waves, buffer
    // loads: one descriptor per operand in scalar registers, the thread's
    // offset in one VGPR and each register's offset as a scalar, so no per-load
    // 64-bit address occupies VGPRs.
    constexpr bool kWaveTile = G::TPTLOG >= 6;  // else the descriptors below go unused
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(wave_uniform(a_src)), short(0), int(G::T * 4u), int(kRawBufferWord3));
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(wave_uniform(b_src)), short(0), int(G::T * 4u), int(kRawBufferWord3));
#pragma unroll
    for (int h = 0; h < G::E / 2; ++h) {
        const int r0 = h, r1 = h + G::E / 2;
        if constexpr (kWaveTile) {
            a[r0] = __builtin_amdgcn_raw_buffer_load_b32(ra, x0 * 4u, G::reg_scatter(0, r0) * 4u,
                                                         kTileStreamPolicy);
            a[r1] = __builtin_amdgcn_raw_buffer_load_b32(ra, x0 * 4u, G::reg_scatter(0, r1) * 4u,
                                                         kTileStreamPolicy);
            b[r0] = __builtin_amdgcn_raw_buffer_load_b32(rb, x0 * 4u, G::reg_scatter(0, r0) * 4u,
                                                         kTileStreamPolicy);
            b[r1] = __builtin_amdgcn_raw_buffer_load_b32(rb, x0 * 4u, G::reg_scatter(0, r1) * 4u,
                                                         kTileStreamPolicy);
        } else {
            a[r0] = a_src[x0 + G::reg_scatter(0, r0)];
            a[r1] = a_src[x0 + G::reg_scatter(0, r1)];
            b[r0] = b_src[x0 + G::reg_scatter(0, r0)];
            b[r1] = b_src[x0 + G::reg_scatter(0, r1)];
        }
    }
    forward_from<TLOG, EL, 0>(a, b, lds, tau, tb.fwd + table, bm, p, p2);
    const Tw s = tb.scale[logn_full * tb.k + pi];
    constexpr bool kBlocks = TLOG >= kBlockLog;
    if constexpr (kBlocks) {
        // block G of the whole transform (zeta table entry 2^(logn-2) + G) is
        // block x/4 of sub-block s = bm - 2^u: bm * 2^(TLOG-2) + x/4
        const Tw* zrow = tb.zs + (size_t(pi) << (tb.max_logn - 1)) + (size_t(bm) << (TLOG - kBlockLog));
        block_product<TLOG, EL>(a, b, zrow, thr_scatter<TLOG, EL, G::NPH - 1>(tau), s, p, p2, neg_inv);
    } else {
#pragma unroll
        for (int r = 0; r < G::E; ++r) a[r] = mont_mul(csub(a[r], p2), csub(b[r], p2), p, neg_inv);
    }
    inverse_from<TLOG, EL, G::NPH - 1>(a, lds, tau, tb.inv + table, tb.invc + table, bm, p, p2);
    const __amdgpu_buffer_rsrc_t rd =
        __builtin_amdgcn_make_buffer_rsrc(wave_uniform(dst), short(0), int(G::T * 4u), int(kRawBufferWord3));
    // A tile that is a sub-block of a larger transform (bm > 1: outer inverse
    // stages follow, in the CRT or a pass of their own) stores its scaled
    // values in [0,2p), as those Gentleman-Sande stages take them (the CRT's
    // leave [0,2p), the pass's [0,p)).  A whole transform stores canonical
    // residues.
    uint32_t np = 0u - p;
    asm volatile("" : "+v"(np));
    const auto store = [&](auto canonical) {
#pragma unroll
        for (int r = 0; r < G::E; ++r) {
            uint32_t t;  // [0,2p)
            if constexpr (kBlocks)
                t = a[r];  // scaled in the block product
            else
                t = shoup_np(a[r], s, np);
            if constexpr (decltype(canonical)::value) t = csub(t, p);
            if (!valid) continue;
            if constexpr (kWaveTile)
                __builtin_amdgcn_raw_buffer_store_b32(t, rd, x0 * 4u, G::reg_scatter(0, r) * 4u, kTileStreamPolicy);
            else
                dst[x0 + G::reg_scatter(0, r)] = t;
        }
    };
    if (bm > 1u)
        store(std::false_type{});
    else
        store(std::true_type{});
}

// ---------------------------------------------------------------------------
// residue conversion on MFMA
// ---------------------------------------------------------------------------
// Wave-level: 64 coefficients starting at c0 (4 column tiles of 16), all
// primes.  Lane l supplies bytes [16*(l/16), +16) and [64+16*(l/16), +16) of
// coefficient c0 + 16*tile + l%16; the weight operand of quad Q holds rows
// (prime 4Q + i/4, digit i%4).  Output lane l gets the 4 digit sums of
// (prime 4Q + l/16, coefficient c0 + 16*tile + l%16).
// Per-block conversion tables in LDS: the MFMA weight operands and, per
// (quad, lane group), the prime, its Shoup constant of 2^16 and the digit-sum
// corrections -- read with LDS instructions so that no global load is issued
// between a group's prefetch and its use (the vector memory counter drains in
// order).
// fold: the digit-sum corrections of the recombination's two halves with
// their offsets, {128 (C0 + 256 C1) + p, 128 (C2 + 256 C3) + 2p} mod 2^32.
struct ConvLds {
    uint4 w[16 * 2 * 64];  // [quad][step][lane]
    uint32_t p[64];
    uint32_t sh16[64];
    uint2 fold[64];
    Tw otw[64][3];  // forward twiddles 1..3 of each prime: the fused outer stages
};

// whole block; ends with a barrier.  Slice [quad_begin, quad_end) of the quads.
__device__ __forceinline__ void conv_tables_to_lds(const Tables& tb, ConvLds& L, uint32_t quad_begin = 0,
                                                   uint32_t quad_end = 0xFFFFFFFFu) {
    quad_end = quad_end < tb.quads ? quad_end : tb.quads;
    for (uint32_t i = threadIdx.x + quad_begin * 128u; i < quad_end * 128u; i += blockDim.x) L.w[i] = tb.conv_w[i];
    for (uint32_t i = threadIdx.x; i < 64u; i += blockDim.x) {
        const uint32_t pi = i < tb.k ? i : 0u;
        L.p[i] = tb.primes[pi];
        L.sh16[i] = tb.conv_sh16[pi];
        const int4 corr = i < tb.quads * 4u ? reinterpret_cast<const int4*>(tb.conv_corr)[i] : make_int4(0, 0, 0, 0);
        L.fold[i] = make_uint2(uint32_t(corr.x + 256 * corr.y) + L.p[i], uint32_t(corr.z + 256 * corr.w) + 2u * L.p[i]);
        for (uint32_t j = 0; j < 3; ++j)  // twiddle 1 + j of the 2^max_logn table (every N shares its prefix)
            L.otw[i][j] = (1u + j) < (1u << tb.max_logn) ? tb.fwd[(size_t(pi) << tb.max_logn) + 1u + j] : Tw{0u, 0u};
    }
    __syncthreads();
}

typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) uint32_t* global_store_u32;

// Operand bytes of one 64-coefficient group as loaded (raw, unconditional
// buffer loads, so a prefetch can stay in flight): lane l holds 16 bytes at
// 16*(l/16) and at 64 + 16*(l/16) of coefficient c0 + 16*tile + l%16.  The
// buffer's range is the c_end coefficients: bytes past the 112-byte
// coefficient (l >= 48, second load) and coefficients past c_end are read as
// zero by the hardware range check, without a memory access.

// Column tile ct holds coefficients c0 + ct * cstride + (0..15): cstride 16
// for 64 consecutive coefficients, n/4 for the fused outer stages' layout.
__device__ __forceinline__ void conv_load(const uint32_t* __restrict__ coef, uint32_t c_end, uint32_t c0,
                                          uint32_t lane, v4u (&v0)[4], v4u (&v1)[4], uint32_t cstride = 16u) {
    const uint32_t col = lane & 15u, kg = lane >> 4;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(coef), short(0), int(c_end * (kLimbs * 4u)), int(kRawBufferWord3));
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const uint32_t byte0 = (c0 + cstride * ct + col) * (kLimbs * 4u) + 16u * kg;
        v0[ct] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, byte0, 0, 0);
        v1[ct] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, kg < 3 ? byte0 + 64u : kOutOfRange, 0, 0);
    }
}

// The same operand bytes read from page-locked host memory: plain loads,
// issued only by the lanes that need them (an out-of-range buffer load would
// still take a slot in the PCIe read path).
__device__ __forceinline__ void conv_load_host(const uint32_t* __restrict__ coef, uint32_t c_end, uint32_t c0,
                                               uint32_t lane, v4u (&v0)[4], v4u (&v1)[4]) {
    const uint32_t col = lane & 15u, kg = lane >> 4;
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const uint32_t c = c0 + 16u * ct + col;
        v0[ct] = v4u{0, 0, 0, 0};
        v1[ct] = v4u{0, 0, 0, 0};
        if (c < c_end) {
            const v4u* src = reinterpret_cast<const v4u*>(coef + size_t(c) * kLimbs);
            v0[ct] = src[kg];
            if (kg < 3) v1[ct] = src[4 + kg];
        }
    }
}

// Residue of one column tile from its two MFMA steps, not yet reduced.  With
// the exact digit sums T_d = sum_j a_j * digit_d(256^j mod p) (accumulator
// plus its correction), u = T0 + 256 T1 and w = T2 + 256 T3, the value
//     v = u + p + ((w + 2p) * 2^16 mod p, one Shoup step in [0,2p))
// is congruent to the coefficient mod p, lies in [0, 2^32), and v - 2p < 4p
// when v >= 2p: the exact per-prime ranges of u and w over any 112 coefficient
// bytes are checked when the tables are built (pm_params.cpp).  So v is a
// valid multiplied (upper) operand of the forward NTT's first butterflies as
// it is, and one conditional subtraction of 2p makes it a lower operand
// (below 4p).  The corrections and offsets are folded into two constants
// (ConvLds::fold): 7 VALU instructions per residue.
__device__ __forceinline__ uint32_t conv_residue(const v4i& w0, const v4i& w1, const v4i& b0, const v4i& b1,
                                                 uint2 fold, uint32_t p, Tw c16) {
    v4i acc = v4i{0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(w0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(w1, b1, acc, 0, 0, 0);
    // u = T0' + 256 T1' alone is one shift-add; the accumulators are read by
    // compiled code only, which waits for the MFMA results (an inline asm
    // reading them gets no wait states: it read unfinished results)
    const uint32_t u = uint32_t(acc[0]) + (uint32_t(acc[1]) << 8);
    const uint32_t w = uint32_t(acc[2]) + (uint32_t(acc[3]) << 8) + fold.y;
    // Shoup product by 2^16 as w 2^16 + floor(w sh16 / 2^32) (2^32 - p): one
    // multiply-add (the low half of v_mad_u64_u32), 2^32 - p kept opaque as in
    // gs_bfly
    uint32_t np = 0u - p;
    asm("" : "+v"(np));
    const uint32_t x = (w << 16) + __umulhi(w, c16.shoup) * np;
    uint32_t v;
    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(v) : "v"(u), "v"(fold.x), "v"(x));
    return v;
}

// Residues of the group for quads [quad_begin, quad_end) of primes; n is the
// row stride of the residue array, coefficients at or past c_end are not
// stored.  With outer = 1 or 2 (cstride = n/4) the lane's four residues are
// one column of the outer radix-2 / radix-4 Cooley-Tukey stages of its prime
// (positions x, x + n/4, x + n/2, x + 3n/4), which are applied before the
// store -- the outer stages of the forward NTT without an HBM pass of their
// own.
__device__ __forceinline__ void conv_compute(const Tables& tb, const v4u (&v0)[4], const v4u (&v1)[4], uint32_t n,
                                             uint32_t c_end, uint32_t c0, uint32_t* __restrict__ res_any,
                                             const ConvLds& L, uint32_t lane, uint32_t quad_begin = 0,
                                             uint32_t quad_end = 0xFFFFFFFFu, uint32_t cstride = 16u,
                                             uint32_t outer = 0u) {
    const global_store_u32 res = (global_store_u32)res_any;
    const uint32_t col = lane & 15u, kg = lane >> 4;
    v4i b0[4], b1[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        // bytes a in [0,255] -> a - 128 in [-128,127] (the zero bytes past the
        // coefficient become -128 but meet zero weight rows)
        const v4u m = v4u{0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u};
        b0[ct] = __builtin_bit_cast(v4i, v0[ct] ^ m);
        b1[ct] = __builtin_bit_cast(v4i, v1[ct] ^ m);
    }
    quad_end = quad_end < tb.quads ? quad_end : tb.quads;
    for (uint32_t quad = quad_begin; quad < quad_end; ++quad) {
        const v4i w0 = __builtin_bit_cast(v4i, L.w[(quad * 2 + 0) * 64 + lane]);
        const v4i w1 = __builtin_bit_cast(v4i, L.w[(quad * 2 + 1) * 64 + lane]);
        const uint32_t prime = quad * 4u + kg;
        const bool live = prime < tb.k;
        const uint32_t p = L.p[prime], p2 = 2u * p;
        const Tw c16{65536u, L.sh16[prime]};
        uint32_t r[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) r[ct] = csub(conv_residue(w0, w1, b0[ct], b1[ct], L.fold[prime], p, c16), p2);
        if (outer >= 1u) {
            // bit log2(n)-1 pairs x with x + n/2 (twiddle 1); bit log2(n)-2
            // pairs x with x + n/4 in each half (twiddles 2 and 3): the
            // stages of the separate outer pass, in the same order; outputs
            // in [0,4p) as the tile's first stage takes them
            const Tw w1 = L.otw[prime][0];
            ct_bfly(r[0], r[2], w1, p, p2);
            ct_bfly(r[1], r[3], w1, p, p2);
            if (outer >= 2u) {
                ct_bfly(r[0], r[1], L.otw[prime][1], p, p2);
                ct_bfly(r[2], r[3], L.otw[prime][2], p, p2);
            }
        }
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const uint32_t c = c0 + cstride * ct + col;
            if (live && c < c_end) res[size_t(prime) * n + c] = r[ct];
        }
    }
}

// A slot of 128 coefficients as two halves A and B, each four column tiles of
// 16 (operand bytes va / vb as conv_load gives them).  Contiguous layout
// (fused = false, cstride 16): A = [cA, cA + 64), B = [cA + 64, cA + 128).
// Fused layout (cstride = n/4): half h holds columns x + 16h .. x + 16h + 15
// of each quarter, so every (prime, quarter) run of the slot is 32 residues;
// after the outer stages a row swap (v_permlane16_swap: rows 1, 3 of A with
// rows 0, 2 of B) gives each lane pair of rows one prime's 32 consecutive
// residues, and each store writes whole 128-byte lines (16-residue runs left
// half-written lines that the L2 wrote back as more HBM traffic).
__device__ __forceinline__ void conv_slot_compute(const Tables& tb, const v4u (&va0)[4], const v4u (&va1)[4],
                                                  const v4u (&vb0)[4], const v4u (&vb1)[4], uint32_t n, uint32_t cA,
                                                  uint32_t cstride, bool fused, uint32_t outer,
                                                  uint32_t* __restrict__ res_any, const ConvLds& L, uint32_t lane,
                                                  uint32_t quad_begin = 0, uint32_t quad_end = 0xFFFFFFFFu) {
    const global_store_u32 res = (global_store_u32)res_any;
    const __amdgpu_buffer_rsrc_t rres =
        __builtin_amdgcn_make_buffer_rsrc(wave_uniform(res_any), short(0), int(tb.k * n * 4u), int(kRawBufferWord3));
    const uint32_t col = lane & 15u, kg = lane >> 4;
    const v4u m = v4u{0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u};
    v4i a0[4], a1[4], b0[4], b1[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        a0[ct] = __builtin_bit_cast(v4i, va0[ct] ^ m);
        a1[ct] = __builtin_bit_cast(v4i, va1[ct] ^ m);
        b0[ct] = __builtin_bit_cast(v4i, vb0[ct] ^ m);
        b1[ct] = __builtin_bit_cast(v4i, vb1[ct] ^ m);
    }
    quad_end = quad_end < tb.quads ? quad_end : tb.quads;
    for (uint32_t quad = quad_begin; quad < quad_end; ++quad) {
        const v4i w0 = __builtin_bit_cast(v4i, L.w[(quad * 2 + 0) * 64 + lane]);
        const v4i w1 = __builtin_bit_cast(v4i, L.w[(quad * 2 + 1) * 64 + lane]);
        const uint32_t prime = quad * 4u + kg;
        const uint32_t p = L.p[prime], p2 = 2u * p;
        const uint2 fold = L.fold[prime];
        const Tw c16{65536u, L.sh16[prime]};
        uint32_t ra[4], rb[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            ra[ct] = conv_residue(w0, w1, a0[ct], a1[ct], fold, p, c16);
            rb[ct] = conv_residue(w0, w1, b0[ct], b1[ct], fold, p, c16);
        }
        if (!fused) {
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const uint32_t c = cA + 16u * ct + col;
                if (prime < tb.k && c < n) res[size_t(prime) * n + c] = csub(ra[ct], p2);
                if (prime < tb.k && c + 64u < n) res[size_t(prime) * n + c + 64u] = csub(rb[ct], p2);
            }
            continue;
        }
        // fused layout: column tiles 0, 1 sit in the lower half of the
        // polynomial (lower operands of the first outer stage: below 4p),
        // tiles 2, 3 in the upper half (multiplied operands: any 32-bit value)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            ra[ct] = csub(ra[ct], p2);
            rb[ct] = csub(rb[ct], p2);
        }
        if (outer >= 1u) {  // the outer stages, as conv_compute applies them
            const Tw t1 = L.otw[prime][0];
            ct_bfly(ra[0], ra[2], t1, p, p2);
            ct_bfly(ra[1], ra[3], t1, p, p2);
            ct_bfly(rb[0], rb[2], t1, p, p2);
            ct_bfly(rb[1], rb[3], t1, p, p2);
            if (outer >= 2u) {
                const Tw t2 = L.otw[prime][1], t3 = L.otw[prime][2];
                ct_bfly(ra[0], ra[1], t2, p, p2);
                ct_bfly(ra[2], ra[3], t3, p, p2);
                ct_bfly(rb[0], rb[1], t2, p, p2);
                ct_bfly(rb[2], rb[3], t3, p, p2);
            }
        }
        // rows (lane / 16) of A: primes 4q..4q+3 at columns x..x+15; of B the
        // same primes at x+16..x+31.  After the swap A holds rows {A0, B0, A2,
        // B2} (primes 4q, 4q, 4q+2, 4q+2) and B rows {A1, B1, A3, B3}.
        // Buffer stores: the lane's byte offset of (pa, column tile 0) in a
        // VGPR, column tile ct and prime pb = pa + 1 as scalar offsets
        // (k * n * 4 <= 64 * 2^17 * 4 bytes: 32-bit offsets).
        const uint32_t pa = quad * 4u + (kg & 2u), pb = pa + 1u;
        const uint32_t voff = ((pa * n + cA + 16u * (kg & 1u) + col) << 2);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const auto sw = __builtin_amdgcn_permlane16_swap(ra[ct], rb[ct], false, false);
            const uint32_t soff = __builtin_amdgcn_readfirstlane(cstride * uint32_t(ct) * 4u);
            if (pa < tb.k) __builtin_amdgcn_raw_buffer_store_b32(sw[0], rres, voff, soff, 0);
            if (pb < tb.k) __builtin_amdgcn_raw_buffer_store_b32(sw[1], rres, voff, soff + 4u * n, 0);
        }
    }
}

// One group, loaded then converted (no prefetch).
__device__ __forceinline__ void convert_wave(const Tables& tb, const uint32_t* __restrict__ coef, uint32_t n,
                                             uint32_t c_end, uint32_t c0, uint32_t* __restrict__ res,
                                             const ConvLds& L, uint32_t lane, uint32_t quad_begin = 0,
                                             uint32_t quad_end = 0xFFFFFFFFu, uint32_t cstride = 16u,
                                             uint32_t outer = 0u) {
    v4u v0[4], v1[4];
    conv_load(coef, c_end, c0, lane, v0, v1, cstride);
    conv_compute(tb, v0, v1, n, c_end, c0, res, L, lane, quad_begin, quad_end, cstride, outer);
}

// Scalar limb reduction, the comparison path of the MFMA conversion (role of
// rns/solve.cu:138-157): one thread per coefficient, every prime,
// r = sum_j w_j * (2^(32j) mod p) with Shoup products, kept below 2p.
__device__ __forceinline__ void convert_scalar_one(const Tables& tb, const uint32_t* __restrict__ coef,
                                                   uint32_t n, uint32_t c, uint32_t* __restrict__ res) {
    uint32_t w[kLimbs];
    const uint4* src = reinterpret_cast<const uint4*>(coef + size_t(c) * kLimbs);
#pragma unroll
    for (int v = 0; v < 7; ++v) {
        const uint4 x = src[v];
        w[4 * v] = x.x;
        w[4 * v + 1] = x.y;
        w[4 * v + 2] = x.z;
        w[4 * v + 3] = x.w;
    }
    for (uint32_t i = 0; i < tb.k; ++i) {
        const uint32_t p = tb.primes[i], p2 = 2u * p;
        const Tw* pw = tb.limb_pow + size_t(i) * kLimbs;
        uint32_t acc = 0;
#pragma unroll
        for (int j = 0; j < int(kLimbs); ++j) acc = csub(acc + shoup(w[j], pw[j], p), p2);
        res[size_t(i) * n + c] = acc;
    }
}

// ---------------------------------------------------------------------------
// CRT: one output coefficient per thread
// ---------------------------------------------------------------------------
__device__ __forceinline__ void crt_finish(const Tables& tb, uint32_t (&x)[29], uint64_t est,
                                           uint32_t* __restrict__ out);
__device__ __forceinline__ void crt_reduce(const Tables& tb, uint32_t (&x)[29], uint64_t est);

__device__ __forceinline__ void crt_one(const Tables& tb, const uint32_t* __restrict__ t_res, uint32_t n,
                                        uint32_t c, uint32_t* __restrict__ out) {
    uint64_t acc[31];
#pragma unroll
    for (int l = 0; l < 31; ++l) acc[l] = 0;
    uint64_t est = 0;
    const uint32_t k = tb.k;
    for (uint32_t i0 = 0; i0 < k; i0 += 15) {
        // issue the group's 15 residue loads together, then accumulate
        uint32_t tv[15];
#pragma unroll
        for (int j = 0; j < 15; ++j) tv[j] = i0 + j < k ? t_res[size_t(i0 + j) * n + c] : 0u;  // < p < 2^30
#pragma unroll
        for (int j = 0; j < 15; ++j) {
            if (i0 + j >= k) break;
            const uint32_t i = i0 + j, t = tv[j];
            est += uint64_t(t) * tb.crt_round[i];
            const uint32_t* m = tb.crt_m30 + size_t(i) * kCrtLimbs;
#pragma unroll
            for (int l = 0; l < 29; ++l) acc[l] += uint64_t(t) * m[l];
        }
        // keep every limb below 2^35 so 15 more products (< 2^60 each) fit
#pragma unroll
        for (int l = 0; l < 30; ++l) {
            acc[l + 1] += (acc[l] >> 32) << 2;
            acc[l] &= 0xFFFFFFFFull;
        }
    }
    // exact 30-bit normalisation
    uint32_t limb[31];
#pragma unroll
    for (int l = 0; l < 30; ++l) {
        acc[l + 1] += acc[l] >> 30;
        limb[l] = uint32_t(acc[l]) & 0x3FFFFFFFu;
    }
    limb[30] = uint32_t(acc[30]);
    // 29 words of 32 bits
    uint32_t x[29];
#pragma unroll
    for (int w = 0; w < 29; ++w) {
        const int L = (32 * w) / 30, s = (32 * w) % 30;
        x[w] = (limb[L] >> s) | (L + 1 < 31 ? (limb[L + 1] << (30 - s)) : 0u);
    }
    crt_finish(tb, x, est, out + size_t(c) * kLimbs);
}

// The 28-word result of crt_reduce stored for one coefficient.
__device__ __forceinline__ void crt_finish(const Tables& tb, uint32_t (&x)[29], uint64_t est,
                                           uint32_t* __restrict__ out) {
    crt_reduce(tb, x, est);
    uint4* dst = reinterpret_cast<uint4*>(out);
#pragma unroll
    for (int w = 0; w < 28; w += 4) dst[w / 4] = make_uint4(x[w], x[w + 1], x[w + 2], x[w + 3]);
}

// From the 29 words of A = sum_i t_i (M_i mod q) and the rounding estimate
// est = sum_i t_i floor(2^42/p_i) of residues t_i < 2 p_i (canonical or one
// p above): K = nearest integer of sum_i t_i/p_i (at most 2k), X = A +
// (-K*P mod q) < 2^905, then Barrett reduction modulo q into x[0..27]
// (role of rns/solve.cu:336-420).
__device__ __forceinline__ void crt_reduce(const Tables& tb, uint32_t (&x)[29], uint64_t est) {
    const uint32_t k = tb.k;
    // nearest integer quotient of the CRT sum by P; add (-K*P) mod q
    uint32_t K = uint32_t((est + (uint64_t(1) << 41)) >> 42);
    K = K <= 2u * k ? K : 2u * k;  // t_i < 2 p_i give K <= 2k; the clamp keeps the table read in bounds
    const uint32_t* nk = tb.crt_negkp + size_t(K) * kLimbs;
    uint32_t cin = 0;
#pragma unroll
    for (int w = 0; w < 28; ++w) x[w] = __builtin_addc(x[w], nk[w], cin, &cin);
    x[28] += cin;
    // Barrett: X < 2^905, Qh = floor(floor(X/2^863) * floor(2^906/q) / 2^43) < 2^42.
    // X/q - Qh < (X mod 2^863)/q + floor(X/2^863)/2^43 + 1 < 1/2 + 1/2 + 1 for
    // q > 2^864, so the result is in [0, 2q).  (The truncation at 2^864 with
    // floor(2^908/q) leaves up to 2^864/q + 1/16 + 1, which passes 2 for q
    // just above 2^864.)
    const uint64_t xh = (uint64_t(x[26]) >> 31) | (uint64_t(x[27]) << 1) | (uint64_t(x[28]) << 33);
    const uint64_t mu = uint64_t(tb.mu0) | (uint64_t(tb.mu1) << 32);
    const uint64_t plo = xh * mu, phi = __umul64hi(xh, mu);
    const uint64_t qh = (phi << 21) | (plo >> 43);
    const uint32_t q0 = uint32_t(qh), q1 = uint32_t(qh >> 32);
    // X -= q0*q and X -= q1*q*2^32: two multiply chains (each word product
    // plus the previous carry fits 64 bits) with their own borrow chains; the
    // true result is in [0, 2q), so the wrapped 29-word arithmetic is exact
    const uint32_t* qw = tb.q;
    uint64_t acc0 = 0, acc1 = 0;
    uint32_t b0 = 0, b1 = 0;
#pragma unroll
    for (int w = 0; w < 29; ++w) {
        acc0 = uint64_t(q0) * (w < 28 ? qw[w] : 0u) + (acc0 >> 32);
        x[w] = __builtin_subc(x[w], uint32_t(acc0), b0, &b0);
        acc1 = uint64_t(q1) * (w >= 1 ? qw[w - 1] : 0u) + (acc1 >> 32);
        x[w] = __builtin_subc(x[w], uint32_t(acc1), b1, &b1);
    }
    // result in [0, 2q): one conditional subtraction
    uint32_t y[28];
    uint32_t bq = 0;
#pragma unroll
    for (int w = 0; w < 28; ++w) y[w] = __builtin_subc(x[w], qw[w], bq, &bq);
    const bool ge = x[28] >= bq;
#pragma unroll
    for (int w = 0; w < 28; ++w) x[w] = ge ? y[w] : x[w];
}

// CRT on int8 MFMA, one wave per group of 64 coefficients [c0, c0+64).  The
// sums A = sum_i t_i (M_i mod q) and est = sum_i t_i floor(2^42/p_i) are one
// exact integer GEMM D[j][c] = sum_{i,u} A[j][4i+u] * byte_u(t_i(c)) over byte
// positions j (rows, 8 tiles of 16) and coefficients (4 column tiles of 16),
// K = 64 primes x 4 bytes in 4 steps; the unsigned bytes enter as byte - 128
// and the row sums of A (the accumulators' start values) restore them.  Lane l
// of a tile holds rows 4(l/16)..+3 of coefficient l%16: one 32-bit word
// position of the sum, kept as a 47-bit signed partial in LDS (low 32 bits,
// high 16 bits); then each lane takes one coefficient, carries through its 29
// words and finishes as the scalar CRT does, so the output is bit-identical.

// B operands of a group: lane holds t of primes 16s+4g+{0..3} for coefficient
// c0+16ct+col; indices are clamped (padding primes meet zero columns of A,
// coefficients past n are never stored).
typedef const __attribute__((address_space(1))) uint32_t* global_u32;

// Column tile ct holds coefficients c0 + ct * cstride + (0..15) (cstride 16:
// 64 consecutive coefficients; n/4: the fused outer stages' layout).
// v[s][ct][j]: residue of prime 16s + 4(lane/16) + j at column tile ct.
__device__ __forceinline__ void crt_load_group(const uint32_t* t_res_any, uint32_t n, uint32_t k,
                                               uint32_t c0, uint32_t lane, uint32_t (&v)[4][4][4],
                                               uint32_t cstride = 16u) {
    const global_u32 t_res = (global_u32)t_res_any;  // device memory: global (not flat) loads
    typedef const __attribute__((address_space(1))) char* global_bytes;
    const uint32_t g = lane >> 4, col = lane & 15u;
    if (c0 + 3u * cstride + 16u <= n) {
        // whole group in range: buffer loads, one offset register per prime
        // and the 4 column tiles 4 * cstride bytes apart as a scalar offset
        // (k * n * 4 <= 64 * 2^17 * 4 bytes: 32-bit offsets)
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint32_t*>(wave_uniform(t_res_any)), short(0), int(k * n * 4u), int(kRawBufferWord3));
        const uint32_t cbytes = __builtin_amdgcn_readfirstlane(4u * cstride);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint32_t i = 16u * s + 4u * g + j;
                i = i < k ? i : k - 1u;
                const uint32_t row = (i * n + c0 + col) << 2;
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
                    v[s][ct][j] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, row, cbytes * uint32_t(ct), 0);
            }
        return;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            uint32_t c = c0 + cstride * ct + col;
            c = c < n ? c : n - 1u;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint32_t i = 16u * s + 4u * g + j;
                i = i < k ? i : k - 1u;
                v[s][ct][j] =
                    *reinterpret_cast<global_u32>(reinterpret_cast<global_bytes>(t_res) + ((i * n + c) << 2));
            }
        }
}

// Slot (uint4) of prime i's outer-stage constants in LDS: one uint4 of padding
// per 4 primes, so the four lane groups' primes (4 apart) read distinct banks.
__host__ __device__ constexpr uint32_t crt_ot_slot(uint32_t i) { return 2u * i + (i >> 2); }

// The outer inverse stages of the fused layout on a loaded group: for each of
// the lane's 16 primes its four values (positions x, x + n/4, x + n/2,
// x + 3n/4 of the tile's scaled output, in [0,2p)) go through the radix-2 or
// radix-4 Gentleman-Sande stages of the separate outer pass, in the same
// order, and stay in [0,2p) (the CRT takes residues below 2p).  ot holds per
// prime {p, w1, w1', w2} {w2', w3, w3', 0}: the inverse twiddles 1..3.
__device__ __forceinline__ void crt_outer_inverse(uint32_t (&b)[4][4][4], const uint4* ot, uint32_t k, uint32_t g,
                                                  uint32_t outer) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t i = 16u * s + 4u * g + j;
            i = i < k ? i : k - 1u;  // padding rows meet zero columns of the digit matrix
            const uint4 c0 = ot[crt_ot_slot(i)], c1 = ot[crt_ot_slot(i) + 1u];
            const uint32_t p = c0.x, p2 = 2u * p;
            uint32_t x0 = b[s][0][j], x1 = b[s][1][j], x2 = b[s][2][j], x3 = b[s][3][j];
            if (outer >= 2u) {
                gs_bfly(x0, x1, Tw{c0.w, c1.x}, p, p2);
                gs_bfly(x2, x3, Tw{c1.y, c1.z}, p, p2);
            }
            gs_bfly(x0, x2, Tw{c0.y, c0.z}, p, p2);
            gs_bfly(x1, x3, Tw{c0.y, c0.z}, p, p2);
            b[s][0][j] = x0;
            b[s][1][j] = x1;
            b[s][2][j] = x2;
            b[s][3][j] = x3;
            // one prime's constants at a time: left to itself the compiler
            // loads all 16 primes' ahead and spills
            __builtin_amdgcn_sched_barrier(0);
        }
}

__device__ __forceinline__ void crt_mfma_group(const Tables& tb, uint32_t (&bop)[4][4][4], const uint4* a_lds,
                                               const int4* corr_lds, uint8_t* wave_lds, uint32_t lane, uint32_t n,
                                               uint32_t c0, uint32_t* __restrict__ out, uint32_t cstride = 16u) {
    const uint32_t g = lane >> 4, col = lane & 15u;
    uint32_t* lo = reinterpret_cast<uint32_t*>(wave_lds);
    int16_t* hi = reinterpret_cast<int16_t*>(wave_lds + kCrtWords * kCrtLdsStride * 4);
    v4i b[4][4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
            b[s][ct] = v4i{int(bop[s][ct][0]), int(bop[s][ct][1]), int(bop[s][ct][2]), int(bop[s][ct][3])} ^
                       v4i{int(0x80808080u), int(0x80808080u), int(0x80808080u), int(0x80808080u)};
#pragma unroll
    for (int rt = 0; rt < 8; ++rt) {
        const uint32_t word = 4u * rt + g;
        const int4 corr = corr_lds[word < kCrtWords ? word : 0u];
        v4i acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = v4i{corr.x, corr.y, corr.z, corr.w};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint4 au = a_lds[(rt * 4 + s) * 64 + lane];
            const v4i a = v4i{int(au.x), int(au.y), int(au.z), int(au.w)};
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[ct] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[s][ct], acc[ct], 0, 0, 0);
        }
        if (word < kCrtWords) {
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                // |D| <= 64 primes * 128 * (3*255 + 127) < 2^22.8 (top byte of t < 2p < 2^31 is
                // below 128), so d0 + 256 d1 and d2 + 256 d3 fit in int32 and the word
                // partial u + 2^16 w stays below 2^47 in magnitude
                const int32_t u = acc[ct][0] + acc[ct][1] * 256;
                const int32_t w = acc[ct][2] + acc[ct][3] * 256;
                const int64_t part = int64_t(u) + int64_t(w) * 65536;
                const uint32_t at = word * kCrtLdsStride + 16u * ct + col;
                lo[at] = uint32_t(part);
                hi[at] = int16_t(part >> 32);
            }
        }
    }
    // the same wave reads back: LDS instructions of one wave execute in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t c = c0 + (lane >> 4) * cstride + (lane & 15u);  // lane = 16 * column tile + column
    uint32_t x[29];
    int64_t carry = 0;
#pragma unroll
    for (int w = 0; w < 29; ++w) {
        const int64_t v = int64_t(uint64_t(lo[w * kCrtLdsStride + lane])) +
                          (int64_t(hi[w * kCrtLdsStride + lane]) << 32) + carry;
        x[w] = uint32_t(v);
        carry = v >> 32;  // arithmetic: partials are signed
    }
    const int64_t e0 = int64_t(uint64_t(lo[29 * kCrtLdsStride + lane])) + (int64_t(hi[29 * kCrtLdsStride + lane]) << 32);
    const int64_t e1 = int64_t(uint64_t(lo[30 * kCrtLdsStride + lane])) + (int64_t(hi[30 * kCrtLdsStride + lane]) << 32);
    const uint64_t est = uint64_t(e0 + (e1 << 32));
    crt_reduce(tb, x, est);
    // The results leave through the same LDS rows: lane L's 112 bytes at
    // 112 L, so the group's column tiles (lanes 16 ct .. 16 ct + 15, 16
    // consecutive coefficients) are 1792 contiguous bytes each, and store i
    // moves 16-byte chunk 64 i + lane: every store instruction writes whole
    // runs of lines instead of 16 bytes at a 112-byte stride.
    uint4* stage = reinterpret_cast<uint4*>(wave_lds);
#pragma unroll
    for (int w = 0; w < 28; w += 4) stage[lane * 7u + uint32_t(w / 4)] = make_uint4(x[w], x[w + 1], x[w + 2], x[w + 3]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (uint32_t i = 0; i < 7; ++i) {
        const uint32_t idx = 64u * i + lane, ct = idx / 112u, off = idx % 112u;
        const uint32_t cbase = c0 + ct * cstride;
        const uint4 v = stage[idx];
        if (cbase + off / 7u < n) reinterpret_cast<uint4*>(out + size_t(cbase) * kLimbs)[off] = v;
    }
    // the next group of this wave rewrites the same LDS rows
    __builtin_amdgcn_wave_barrier();
}

// Shared operand tables to LDS (whole block), then a barrier.  ot_lds: per
// prime the modulus and inverse twiddles 1..3 of the fused outer stages.
__device__ __forceinline__ void crt_tables_to_lds(const Tables& tb, uint4* a_lds, int4* corr_lds, uint4* ot_lds) {
    for (uint32_t i = threadIdx.x; i < 8u * 4u * 64u; i += blockDim.x) a_lds[i] = tb.crt_a[i];
    for (uint32_t i = threadIdx.x; i < 32u; i += blockDim.x) corr_lds[i] = tb.crt_corr[i];
    for (uint32_t i = threadIdx.x; i < 64u; i += blockDim.x) {
        const uint32_t pi = i < tb.k ? i : 0u;
        const Tw* inv = tb.inv + (size_t(pi) << tb.max_logn);
        Tw w[3] = {};
        for (uint32_t j = 0; j < 3; ++j)
            if ((1u + j) < (1u << tb.max_logn)) w[j] = inv[1u + j];
        ot_lds[crt_ot_slot(i)] = make_uint4(tb.primes[pi], w[0].value, w[0].shoup, w[1].value);
        ot_lds[crt_ot_slot(i) + 1u] = make_uint4(w[1].shoup, w[2].value, w[2].shoup, 0u);
    }
    __syncthreads();
}

// LDS of a CRT block with W waves: A operands, row corrections, outer-stage
// constants, per-wave rows
constexpr size_t kCrtSharedLds =
    8u * 4u * 64u * sizeof(uint4) + 32u * sizeof(int4) + (crt_ot_slot(63) + 2u) * sizeof(uint4);
template <int W>
constexpr size_t crt_block_lds() {
    return kCrtSharedLds + size_t(W) * kCrtWaveLds;
}

// ---------------------------------------------------------------------------
// quartic backend: q = 2^868 - delta, four 217-bit components, four channels
// A(r i^k) per prime (role of quartic/solve.cu:289-316 and 906-1040)
// ---------------------------------------------------------------------------
// canonical a + b, a - b for a, b in [0,p)
__device__ __forceinline__ uint32_t add_p(uint32_t a, uint32_t b, uint32_t p) { return csub(a + b, p); }
__device__ __forceinline__ uint32_t sub_p(uint32_t a, uint32_t b, uint32_t p) { return csub(a - b + p, p); }

// Channels of coefficient c (28 words at coef + 28 c) into rows 4i + j of
// res (row stride n): the components a_j = bits [217 j, 217 j + 217), their
// residues from 7 Shoup products with 2^(32 l) mod p each, z_j = a_j r^j and
// the 4-point transform over i (channel j = sum_m z_m i^(j m)), canonical.
__device__ __forceinline__ void quartic_convert_one(const Tables& tb, const uint32_t* __restrict__ coef,
                                                    uint32_t n, uint32_t c, uint32_t* __restrict__ res) {
    uint32_t w[kLimbs];
    const uint4* src = reinterpret_cast<const uint4*>(coef + size_t(c) * kLimbs);
#pragma unroll
    for (int v = 0; v < 7; ++v) {
        const uint4 x = src[v];
        w[4 * v] = x.x;
        w[4 * v + 1] = x.y;
        w[4 * v + 2] = x.z;
        w[4 * v + 3] = x.w;
    }
    uint32_t part[4][7];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        constexpr uint32_t kBits = 217;
        const uint32_t shift = kBits * uint32_t(j), base = shift / 32, sh = shift % 32;
#pragma unroll
        for (int l = 0; l < 7; ++l) {
            uint32_t v = w[base + l] >> sh;
            if (sh && base + l + 1 < kLimbs) v |= w[base + l + 1] << (32 - sh);
            part[j][l] = l == 6 ? (v & ((1u << 25) - 1u)) : v;
        }
    }
    for (uint32_t i = 0; i < tb.qk; ++i) {
        const uint32_t p = tb.qprimes[i], p2 = 2u * p;
        const Tw* pw = tb.q_limb_pow + size_t(i) * 7;
        uint32_t z[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t acc = 0;
#pragma unroll
            for (int l = 0; l < 7; ++l) acc = csub(acc + shoup(part[j][l], pw[l], p), p2);  // [0,2p)
            acc = csub(acc, p);
            z[j] = j ? csub(shoup(acc, tb.q_rpow[i * 4 + j], p), p) : acc;
        }
        const uint32_t u = add_p(z[0], z[2], p), v = sub_p(z[0], z[2], p);
        const uint32_t s = add_p(z[1], z[3], p), t = csub(shoup(z[1] - z[3] + p, tb.q_iota[i], p), p);
        uint32_t* row = res + size_t(4 * i) * n + c;
        row[0] = add_p(u, s, p);
        row[n] = add_p(v, t, p);
        row[2 * size_t(n)] = sub_p(u, s, p);
        row[3 * size_t(n)] = sub_p(v, t, p);
    }
}

// Y_j = X_j + (qk - K_j) P for component J (see quartic_crt_one), added
// into S at bit 217 J
template <int J>
__device__ __forceinline__ void quartic_crt_component(const Tables& tb, const uint32_t (&t)[kQuarticMaxPrimes][4],
                                                      uint32_t (&S)[37]) {
    constexpr int j = J;
    const uint32_t qk = tb.qk;
    uint64_t acc[kQuarticLimbs30];
#pragma unroll
    for (int l = 0; l < int(kQuarticLimbs30); ++l) acc[l] = 0;
    uint64_t est = 0;
#pragma unroll
    for (uint32_t i = 0; i < kQuarticMaxPrimes; ++i) {
        if (i >= qk) continue;
        const uint64_t ti = t[i][j];
        est += ti * tb.q_round[i];
        const uint32_t* m = tb.q_m30 + size_t(i) * kQuarticLimbs30;
#pragma unroll
        for (int l = 0; l < int(kQuarticLimbs30); ++l) acc[l] += ti * m[l];  // each product < 2^60
        if (i == 7) {
            // 17 products could pass 2^64: carry the first 8 up, leaving
            // each limb below 2^30 + 2^34 before the rest
#pragma unroll
            for (int l = 0; l < int(kQuarticLimbs30) - 1; ++l) {
                acc[l + 1] += acc[l] >> 30;
                acc[l] &= 0x3FFFFFFFull;
            }
        }
    }
    // K_j (the fixed-point error is below 17 * 2^30 / 2^42 and R_j / P in (-1/4, 1/4))
    uint32_t K = uint32_t((est + (uint64_t(1) << 41)) >> 42);
    K = K <= qk ? K : qk;
    const uint32_t* kp = tb.q_kp30 + size_t(qk - K) * kQuarticLimbs30;
#pragma unroll
    for (int l = 0; l < int(kQuarticLimbs30); ++l) acc[l] += kp[l];
    uint32_t limb[kQuarticLimbs30];  // exact 30-bit limbs of Y_j (< 2^515)
#pragma unroll
    for (int l = 0; l < int(kQuarticLimbs30) - 1; ++l) {
        acc[l + 1] += acc[l] >> 30;
        limb[l] = uint32_t(acc[l]) & 0x3FFFFFFFu;
    }
    limb[kQuarticLimbs30 - 1] = uint32_t(acc[kQuarticLimbs30 - 1]);
    uint32_t y[17];  // 32-bit words
#pragma unroll
    for (int w = 0; w < 17; ++w) {
        const int L0 = (32 * w) / 30, s0 = (32 * w) % 30;
        y[w] = (limb[L0] >> s0) | (L0 + 1 < int(kQuarticLimbs30) ? (limb[L0 + 1] << (30 - s0)) : 0u) |
               (s0 > 28 && L0 + 2 < int(kQuarticLimbs30) ? (limb[L0 + 2] << (60 - s0)) : 0u);
    }
    // S += Y_j 2^(217 j)
    const int off = (217 * j) / 32, sh = (217 * j) % 32;
    uint32_t cy = 0;
#pragma unroll
    for (int w = 0; w < 18; ++w) {
        const uint32_t lo = w < 17 ? y[w] : 0u, prev = w > 0 ? y[w - 1] : 0u;
        const uint32_t part = sh ? ((lo << sh) | (prev >> (32 - sh))) : lo;
        S[off + w] = __builtin_addc(S[off + w], part, cy, &cy);
    }
#pragma unroll
    for (int w = off + 18; w < 37; ++w) S[w] = __builtin_addc(S[w], 0u, cy, &cy);
}

// The 28-word product of coefficient c from the tile's scaled channel values
// (rows 4i + j of t_res, in [0,p), already times (4N)^-1 (M_i mod p_i)^-1):
// per prime the inverse 4-point transform and r^-j give t_ij = R_j (M_i mod
// p_i)^-1 mod p_i; per component the CRT sum X_j = sum_i t_ij M_i and its
// nearest-integer quotient K_j = round(sum_i t_ij / p_i) give Y_j = X_j +
// (qk - K_j) P = R_j + qk P > 0; S = sum_j Y_j 2^(217 j) is folded modulo
// q = 2^868 - delta, plus the constant (-qk P (1 + B + B^2 + B^3)) mod q.
__device__ __forceinline__ void quartic_crt_one(const Tables& tb, const uint32_t* __restrict__ t_res, uint32_t n,
                                                uint32_t c, uint32_t* __restrict__ out) {
    const uint32_t qk = tb.qk;
    uint32_t t[kQuarticMaxPrimes][4];
#pragma unroll
    for (uint32_t i = 0; i < kQuarticMaxPrimes; ++i) {
        t[i][0] = t[i][1] = t[i][2] = t[i][3] = 0;
        if (i >= qk) continue;  // (qk is the same for every lane)
        const uint32_t p = tb.qprimes[i];
        const uint32_t* row = t_res + size_t(4 * i) * n + c;
        const uint32_t f0 = row[0], f1 = row[n], f2 = row[2 * size_t(n)], f3 = row[3 * size_t(n)];
        const uint32_t u = add_p(f0, f2, p), v = add_p(f1, f3, p);
        const uint32_t d0 = sub_p(f0, f2, p), d1 = csub(shoup(f1 - f3 + p, tb.q_iota[i], p), p);
        t[i][0] = add_p(u, v, p);
        t[i][1] = csub(shoup(sub_p(d0, d1, p), tb.q_rinv[i * 4 + 1], p), p);
        t[i][2] = csub(shoup(sub_p(u, v, p), tb.q_rinv[i * 4 + 2], p), p);
        t[i][3] = csub(shoup(add_p(d0, d1, p), tb.q_rinv[i * 4 + 3], p), p);
    }
    uint32_t S[37];
#pragma unroll
    for (int w = 0; w < 37; ++w) S[w] = 0;
    quartic_crt_component<0>(tb, t, S);
    quartic_crt_component<1>(tb, t, S);
    quartic_crt_component<2>(tb, t, S);
    quartic_crt_component<3>(tb, t, S);
    // S mod (2^868 - delta): low 868 bits + high * delta + the constant
    uint32_t T[29];
#pragma unroll
    for (int w = 0; w < 27; ++w) T[w] = S[w];
    T[27] = S[27] & 0xFu;
    T[28] = 0;
    uint32_t H[10];
#pragma unroll
    for (int w = 0; w < 10; ++w) H[w] = (S[27 + w] >> 4) | (28 + w < 37 ? (S[28 + w] << 28) : 0u);
    uint64_t carry = 0;
#pragma unroll
    for (int w = 0; w < 29; ++w) {
        const uint64_t v = uint64_t(T[w]) + (w < 10 ? uint64_t(H[w]) * tb.qdelta : 0u) +
                           (w < int(kLimbs) ? tb.q_negc[w] : 0u) + carry;
        T[w] = uint32_t(v);
        carry = v >> 32;
    }
    // T < 2^870: fold once more (T = h 2^868 + l, h <= 3), then one
    // conditional subtraction of q
    const uint32_t h = (T[27] >> 4) | (T[28] << 28);
    T[27] &= 0xFu;
    uint32_t cin = 0;
    T[0] = __builtin_addc(T[0], h * tb.qdelta, cin, &cin);
#pragma unroll
    for (int w = 1; w < 28; ++w) T[w] = __builtin_addc(T[w], 0u, cin, &cin);
    // T >= q  <=>  T + delta >= 2^868, and then T - q = (T + delta) mod 2^868
    uint32_t U[28];
    cin = 0;
    U[0] = __builtin_addc(T[0], tb.qdelta, cin, &cin);
#pragma unroll
    for (int w = 1; w < 28; ++w) U[w] = __builtin_addc(T[w], 0u, cin, &cin);
    const bool ge = (U[27] >> 4) != 0;
    U[27] &= 0xFu;
    uint4* dst = reinterpret_cast<uint4*>(out + size_t(c) * kLimbs);
#pragma unroll
    for (int w = 0; w < 28; w += 4)
        dst[w / 4] = ge ? make_uint4(U[w], U[w + 1], U[w + 2], U[w + 3]) : make_uint4(T[w], T[w + 1], T[w + 2], T[w + 3]);
}

__global__ void __launch_bounds__(256) k_convert_quartic(Tables tb, Job job) {
    const uint32_t n = 1u << job.logn, c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) quartic_convert_one(tb, blockIdx.y ? job.b : job.a, n, c, blockIdx.y ? job.rb : job.ra);
}

__global__ void __launch_bounds__(256) k_crt_quartic(Tables tb, Job job) {
    const uint32_t n = 1u << job.logn, c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) quartic_crt_one(tb, job.ra, n, c, job.c);
}

// batch: conversion items of kConvItemCoefs coefficients, CRT items of kBlock
__global__ void __launch_bounds__(256) k_items_conv_quartic(Tables tb, BatchArgs ba, uint32_t first) {
    const Item it = ba.items[first + blockIdx.x];
    const Job job = ba.jobs[it.job];
    const uint32_t n = 1u << job.logn, end = min(n, (it.a1 + 1u) * kConvItemCoefs);
    for (uint32_t c = it.a1 * kConvItemCoefs + threadIdx.x; c < end; c += blockDim.x)
        quartic_convert_one(tb, it.a0 ? job.b : job.a, n, c, it.a0 ? job.rb : job.ra);
}

__global__ void __launch_bounds__(256) k_items_crt_quartic(Tables tb, BatchArgs ba, uint32_t first) {
    const Item it = ba.items[first + blockIdx.x];
    const Job job = ba.jobs[it.job];
    const uint32_t n = 1u << job.logn, end = min(n, (it.a0 + 1u) * kBlock);
    for (uint32_t c = it.a0 * kBlock + threadIdx.x; c < end; c += blockDim.x) quartic_crt_one(tb, job.ra, n, c, job.c);
}

// ---------------------------------------------------------------------------
// standalone kernels (plain path)
// ---------------------------------------------------------------------------
// Coefficients [c_first, c_end) of a (poly 0) and/or b (1): blockIdx.y is
// poly - poly_first; blockIdx.z picks one of gridDim.z slices of the primes,
// so a single product still spreads over many waves.
__global__ void __launch_bounds__(kBlock) k_convert(Tables tb, Job job, uint32_t c_first, uint32_t c_end,
                                                    uint32_t poly_first) {
    __shared__ ConvLds L;
    const uint32_t per = (tb.quads + gridDim.z - 1) / gridDim.z;
    const uint32_t quad_begin = blockIdx.z * per, quad_end = quad_begin + per;
    conv_tables_to_lds(tb, L, quad_begin, quad_end);
    const uint32_t n = 1u << job.logn;
    const uint32_t poly = poly_first + blockIdx.y;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t c0 = c_first + blockIdx.x * blockDim.x + wave * kConvPerWave;  // one coefficient per thread
    if (c0 >= c_end) return;
    convert_wave(tb, poly ? job.b : job.a, n, c_end, c0, poly ? job.rb : job.ra, L, lane, quad_begin, quad_end);
}

// The plain path with the outer stages fused: wave w of the grid converts
// slot w (columns 32w .. 32w + 31 of each quarter) of a (blockIdx.y = 0)
// or b; blockIdx.z picks a slice of the primes.
__global__ void __launch_bounds__(kBlock) k_convert_fused(Tables tb, Job job) {
    __shared__ ConvLds L;
    const uint32_t per = (tb.quads + gridDim.z - 1) / gridDim.z;
    const uint32_t quad_begin = blockIdx.z * per, quad_end = quad_begin + per;
    conv_tables_to_lds(tb, L, quad_begin, quad_end);
    const uint32_t n = 1u << job.logn, lane = threadIdx.x & 63u;
    const uint32_t slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (slot * kConvSlotCoefs >= n) return;
    const uint32_t* coef = blockIdx.y ? job.b : job.a;
    const uint32_t cA = 32u * slot;
    v4u a0[4], a1[4], b0[4], b1[4];
    conv_load(coef, n, cA, lane, a0, a1, n / 4u);
    conv_load(coef, n, cA + 16u, lane, b0, b1, n / 4u);
    conv_slot_compute(tb, a0, a1, b0, b1, n, cA, n / 4u, true, job.outer, blockIdx.y ? job.rb : job.ra, L, lane,
                      quad_begin, quad_end);
}

__global__ void __launch_bounds__(256) k_convert_scalar(Tables tb, Job job) {
    const uint32_t n = 1u << job.logn;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const bool poly = blockIdx.y != 0;
    convert_scalar_one(tb, poly ? job.b : job.a, n, c, poly ? job.rb : job.ra);
}

// units = k << u sub-blocks; unit -> (prime = unit >> u, s = unit & (2^u-1))
template <int TLOG>
__global__ void __launch_bounds__(tile_block_threads(TLOG)) k_conv_tile(Tables tb, Job job, uint32_t u,
                                                                        uint32_t units) {
    extern __shared__ uint32_t lds[];
    using G = Geo<TLOG>;
    const uint32_t tiles_per_block = blockDim.x >> G::TPTLOG;
    const uint32_t tile = threadIdx.x >> G::TPTLOG, tau = threadIdx.x & (G::TPT - 1);
    const uint32_t unit = blockIdx.x * tiles_per_block + tile;
    const bool valid = unit < units;
    const uint32_t prime = valid ? (unit >> u) : 0u, s = valid ? (unit & ((1u << u) - 1u)) : 0u;
    const size_t n = size_t(1) << job.logn;
    const size_t off = size_t(prime) * n + (size_t(s) << TLOG);
    conv_tile<TLOG>(tb, job.ra + off, job.rb + off, job.ra + off, prime, (1u << u) + s, job.logn,
                    lds + tile * G::STRIDE, tau, valid);
}

// outer CT stages for one (poly, prime) channel, columns strided by `stride`
template <int U>
__device__ __forceinline__ void outer_forward_columns(const Tables& tb, uint32_t* base, uint32_t tlog,
                                                      uint32_t prime, uint32_t x0, uint32_t stride) {
    const uint32_t T = 1u << tlog;
    const uint32_t p = tb.primes[prime], p2 = 2u * p;
    const Tw* tw = tb.fwd + (size_t(prime) << tb.max_logn);
    constexpr int R = 1 << U;
    for (uint32_t x = x0; x < T; x += stride) {
        uint32_t r[R];
#pragma unroll
        for (int j = 0; j < R; ++j) r[j] = base[(size_t(j) << tlog) + x];
#pragma unroll
        for (int bit = U - 1; bit >= 0; --bit) {
            const uint32_t m = 1u << (U - 1 - bit);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                if (j & (1 << bit)) continue;
                ct_bfly(r[j], r[j | (1 << bit)], tw[m + (uint32_t(j) >> (bit + 1))], p, p2);
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) base[(size_t(j) << tlog) + x] = r[j];
    }
}

template <int U>
__device__ __forceinline__ void outer_inverse_columns(const Tables& tb, uint32_t* base, uint32_t tlog,
                                                      uint32_t prime, uint32_t x0, uint32_t stride) {
    const uint32_t T = 1u << tlog;
    const uint32_t p = tb.primes[prime], p2 = 2u * p;
    const Tw* tw = tb.inv + (size_t(prime) << tb.max_logn);
    constexpr int R = 1 << U;
    for (uint32_t x = x0; x < T; x += stride) {
        uint32_t r[R];
#pragma unroll
        for (int j = 0; j < R; ++j) r[j] = base[(size_t(j) << tlog) + x];
#pragma unroll
        for (int bit = 0; bit < U; ++bit) {
            const uint32_t m = 1u << (U - 1 - bit);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                if (j & (1 << bit)) continue;
                gs_bfly(r[j], r[j | (1 << bit)], tw[m + (uint32_t(j) >> (bit + 1))], p, p2);
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) base[(size_t(j) << tlog) + x] = csub(r[j], p);
    }
}

// ---------------------------------------------------------------------------
// batch work-list items
// ---------------------------------------------------------------------------
template <int TLOG, int EL = tile_elog(TLOG)>
__device__ __forceinline__ void tile_item(const Tables& tb, const BatchArgs& ba, const Item& it, uint32_t* lds) {
    using G = Geo<TLOG, EL>;
    const uint32_t tile = threadIdx.x >> G::TPTLOG, tau = threadIdx.x & (G::TPT - 1);
    const uint32_t count = it.a0 >> 8;
    const bool valid = tile < count;
    const Unit unit = ba.units[it.a1 + (valid ? tile : 0u)];  // padding reads the item's first tile
    const Job job = ba.jobs[unit.job];
    const uint32_t prime = unit.prime_s >> 2, s = unit.prime_s & 3u;
    const size_t off = (size_t(prime) << job.logn) + (size_t(s) << TLOG);
    conv_tile<TLOG, EL>(tb, job.ra + off, job.rb + off, job.ra + off, prime, (1u << job.outer) + s, job.logn,
                        lds + tile * G::STRIDE, tau, valid);
}

// Outer forward CT stages (bits logn-1 .. logn-u) for a and b: one thread per
// column x of the (2^u x T) view; grid.y = 2*k (poly, prime).
template <int U>
__global__ void __launch_bounds__(256) k_outer_forward(Tables tb, Job job) {
    const uint32_t logn = job.logn, tlog = logn - U;
    const uint32_t T = 1u << tlog;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= T) return;
    const uint32_t poly = blockIdx.y / tb.k, prime = blockIdx.y % tb.k;
    const uint32_t p = tb.primes[prime], p2 = 2u * p;
    const Tw* tw = tb.fwd + (size_t(prime) << tb.max_logn);
    uint32_t* v = (poly ? job.rb : job.ra) + (size_t(prime) << logn) + x;
    constexpr int R = 1 << U;
    uint32_t r[R];
#pragma unroll
    for (int j = 0; j < R; ++j) r[j] = v[size_t(j) << tlog];
#pragma unroll
    for (int bit = U - 1; bit >= 0; --bit) {
        const uint32_t m = 1u << (U - 1 - bit);  // groups at this stage
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & (1 << bit)) continue;
            ct_bfly(r[j], r[j | (1 << bit)], tw[m + (uint32_t(j) >> (bit + 1))], p, p2);
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) v[size_t(j) << tlog] = r[j];
}

template <int U>
__global__ void __launch_bounds__(256) k_outer_inverse(Tables tb, Job job) {
    const uint32_t logn = job.logn, tlog = logn - U;
    const uint32_t T = 1u << tlog;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= T) return;
    const uint32_t prime = blockIdx.y;
    const uint32_t p = tb.primes[prime], p2 = 2u * p;
    const Tw* tw = tb.inv + (size_t(prime) << tb.max_logn);
    uint32_t* v = job.ra + (size_t(prime) << logn) + x;
    constexpr int R = 1 << U;
    uint32_t r[R];
#pragma unroll
    for (int j = 0; j < R; ++j) r[j] = v[size_t(j) << tlog];
#pragma unroll
    for (int bit = 0; bit < U; ++bit) {
        const uint32_t m = 1u << (U - 1 - bit);
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & (1 << bit)) continue;
            gs_bfly(r[j], r[j | (1 << bit)], tw[m + (uint32_t(j) >> (bit + 1))], p, p2);
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) v[size_t(j) << tlog] = csub(r[j], p);
}

__global__ void __launch_bounds__(kCrtPerBlock) k_crt(Tables tb, Job job) {
    const uint32_t n = 1u << job.logn;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    crt_one(tb, job.ra, n, c, job.c);
}

// ---------------------------------------------------------------------------
// phased batch: one launch per phase over a contiguous range of the work list
// (the launches order the phases; block b runs item first + b)
// ---------------------------------------------------------------------------
// Conversion items (kConvItemCoefs coefficients of one polynomial each, 16
// slots of 128) run by resident blocks: wave w of block b takes slots b*W + w,
// + grid*W, ...; the weights go to LDS once per block.  A slot's two halves
// are loaded together (64 b128 loads per lane in flight), then converted.
// The conversion items [first, first + count) by the waves of the grid, the
// tables already in L (each wave independent: no block barrier inside).
__device__ __forceinline__ void conv_items_loop(const Tables& tb, const BatchArgs& ba, uint32_t first,
                                                uint32_t count, const ConvLds& L) {
    constexpr uint32_t per_item = kConvItemCoefs / kConvSlotCoefs;
    // the slot index and all that follows from it (item, product, pointers)
    // in scalar registers: blockDim.x comes from the dispatch packet through a
    // vector load, and a stride left in a VGPR makes the loop, the pointers
    // and the buffer descriptors per-lane (waterfall loops around every load)
    const uint32_t waves = __builtin_amdgcn_readfirstlane(blockDim.x >> 6), lane = threadIdx.x & 63u;
    const uint32_t total = count * per_item, stride = __builtin_amdgcn_readfirstlane(gridDim.x * waves);
    for (uint32_t g = __builtin_amdgcn_readfirstlane(blockIdx.x * waves + (threadIdx.x >> 6)); g < total;
         g += stride) {
        const Item it = const_entry(ba.items, first + g / per_item);
        const Job job = const_entry(ba.jobs, it.job);
        const uint32_t n = __builtin_amdgcn_readfirstlane(1u << job.logn);
        const uint32_t slot = __builtin_amdgcn_readfirstlane(it.a1) * per_item + g % per_item;
        if (slot * kConvSlotCoefs >= n) continue;  // the item's tail past a small n
        const bool second = __builtin_amdgcn_readfirstlane(it.a0) != 0;
        const uint32_t* coef = wave_uniform(second ? job.b : job.a);
        uint32_t* res = wave_uniform(second ? job.rb : job.ra);
        const bool fused = __builtin_amdgcn_readfirstlane(job.fused) != 0;
        // fused: slot s covers columns 32s .. 32s + 31 of each quarter
        const uint32_t cA = fused ? 32u * slot : kConvSlotCoefs * slot, cstride = fused ? n / 4u : 16u;
        const uint32_t cB = fused ? cA + 16u : cA + 64u;
        v4u a0[4], a1[4], b0[4], b1[4];
        conv_load(coef, n, cA, lane, a0, a1, cstride);
        conv_load(coef, n, cB, lane, b0, b1, cstride);
        conv_slot_compute(tb, a0, a1, b0, b1, n, cA, cstride, fused, __builtin_amdgcn_readfirstlane(job.outer), res, L,
                          lane);
    }
}

__global__ void __launch_bounds__(kBlock) k_items_conv(Tables tb, BatchArgs ba, uint32_t first, uint32_t count) {
    __shared__ ConvLds L;
    conv_tables_to_lds(tb, L);
    conv_items_loop(tb, ba, first, count, L);
}

__global__ void __launch_bounds__(kBlock) k_items_conv_scalar(Tables tb, BatchArgs ba, uint32_t first) {
    const Item it = ba.items[first + blockIdx.x];
    const Job job = ba.jobs[it.job];
    const uint32_t n = 1u << job.logn;
    for (uint32_t c = it.a1 * kConvItemCoefs + threadIdx.x; c < n && c < (it.a1 + 1u) * kConvItemCoefs;
         c += blockDim.x)
        convert_scalar_one(tb, it.a0 ? job.b : job.a, n, c, it.a0 ? job.rb : job.ra);
}

__global__ void __launch_bounds__(kBlock) k_items_outer(Tables tb, BatchArgs ba, uint32_t first) {
    const Item it = ba.items[first + blockIdx.x];
    const Job job = ba.jobs[it.job];
    const uint32_t tlog = job.logn - job.outer;
    if (it.type == kItemOuterFwd) {
        const uint32_t poly = it.a0 / tb.k, prime = it.a0 % tb.k;
        uint32_t* base = (poly ? job.rb : job.ra) + (size_t(prime) << job.logn);
        if (job.outer == 1) outer_forward_columns<1>(tb, base, tlog, prime, threadIdx.x, kBlock);
        else outer_forward_columns<2>(tb, base, tlog, prime, threadIdx.x, kBlock);
    } else {
        uint32_t* base = job.ra + (size_t(it.a0) << job.logn);
        if (job.outer == 1) outer_inverse_columns<1>(tb, base, tlog, it.a0, threadIdx.x, kBlock);
        else outer_inverse_columns<2>(tb, base, tlog, it.a0, threadIdx.x, kBlock);
    }
}

template <int TLOG, int EL>
__global__ void __launch_bounds__(tile_block_threads(TLOG)) k_items_tile(Tables tb, BatchArgs ba, uint32_t first) {
    extern __shared__ uint32_t lds[];
    const Item it = ba.items[first + blockIdx.x];
    tile_item<TLOG, EL>(tb, ba, it, lds);
}

// One group of 64 coefficients of the CRT: residues t ([k][n]), output, its
// first coefficient and column-tile stride, and the fused outer stages.
// A group with nothing to store has c0 >= n and cstride 16.
struct CrtGroup {
    const uint32_t* t;
    uint32_t* out;
    uint32_t n, c0, cstride, outer;
};

// Groups of 64 coefficients, looped by W-wave blocks (wave w of block b takes
// groups b*W + w, + grid*W, ...); the next group's residues load while the
// current one is computed.
template <int W, typename GroupFn>
__device__ __forceinline__ void crt_group_loop(const Tables& tb, uint32_t groups, GroupFn where) {
    extern __shared__ uint8_t crt_lds[];
    uint4* a_lds = reinterpret_cast<uint4*>(crt_lds);
    int4* corr_lds = reinterpret_cast<int4*>(crt_lds + 8u * 4u * 64u * sizeof(uint4));
    uint4* ot_lds = reinterpret_cast<uint4*>(crt_lds + 8u * 4u * 64u * sizeof(uint4) + 32u * sizeof(int4));
    uint8_t* rows = crt_lds + kCrtSharedLds;
    crt_tables_to_lds(tb, a_lds, corr_lds, ot_lds);
    // the wave index is the same in every lane: make it (and all that follows
    // from it: group, product, pointers) scalar
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
    const uint32_t stride = gridDim.x * W;
    uint32_t gi = blockIdx.x * W + wave;
    if (gi >= groups) return;  // whole wave; no block barriers follow
    CrtGroup cur_g;
    where(gi, cur_g);
    uint32_t cur[4][4][4];
    crt_load_group(cur_g.t, cur_g.n, tb.k, cur_g.c0, lane, cur, cur_g.cstride);
    uint8_t* my_rows = rows + wave * kCrtWaveLds;
    for (;;) {
        // the fused outer stages on this group before the next one's loads
        // are issued: both sets of 64 values at once do not fit the registers
        if (cur_g.outer) {
            crt_outer_inverse(cur, ot_lds, tb.k, lane >> 4, cur_g.outer);
            __builtin_amdgcn_sched_barrier(0);
        }
        const uint32_t gn = gi + stride;
        const bool more = gn < groups;
        CrtGroup nxt_g;
        where(more ? gn : gi, nxt_g);
        uint32_t nxt[4][4][4];
        crt_load_group(nxt_g.t, nxt_g.n, tb.k, nxt_g.c0, lane, nxt, nxt_g.cstride);  // a repeat when none is left
        crt_mfma_group(tb, cur, a_lds, corr_lds, my_rows, lane, cur_g.n, cur_g.c0, cur_g.out, cur_g.cstride);
        if (!more) break;
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int j = 0; j < 4; ++j) cur[s][ct][j] = nxt[s][ct][j];
        gi = gn;
        cur_g = nxt_g;
    }
}

// coefficients [c_first, c_end) with c_first a multiple of 64 and c_end a
// multiple of 64 or n; job.c may be device memory or mapped page-locked host
// memory
__global__ void __launch_bounds__(kCrtPlainWaves * 64) k_crt_mfma(Tables tb, Job job, uint32_t c_first,
                                                                  uint32_t c_end) {
    const uint32_t n = 1u << job.logn;
    crt_group_loop<kCrtPlainWaves>(tb, (c_end - c_first + kCrtWaveCoefs - 1) / kCrtWaveCoefs,
                                  [&](uint32_t gi, CrtGroup& w) {
                                      // fused: whole range, group gi at quarter offsets
                                      w = job.fused ? CrtGroup{job.ra, job.c, n, 16u * gi, n / 4u, job.outer}
                                                    : CrtGroup{job.ra, job.c, n, c_first + gi * kCrtWaveCoefs, 16u, 0u};
                                  });
}

// CRT items cover kBlock coefficients = kBlock / 64 groups each; groups past a
// product's n (small N) load clamped data and store nothing.  W waves per
// block: 8 fill a CU's LDS (about 154 KB) and its VGPRs (224 per lane, two
// waves per SIMD) with one block; 4 or 2 spread a small batch over more CUs.
template <int W>
__device__ __forceinline__ void crt_items_loop(const Tables& tb, const BatchArgs& ba, uint32_t first,
                                               uint32_t count) {
    constexpr uint32_t per_item = kBlock / kCrtWaveCoefs;
    crt_group_loop<W>(tb, count * per_item, [&](uint32_t gi, CrtGroup& w) {
        const Item it = ba.items[first + gi / per_item];
        const Job job = ba.jobs[it.job];
        const uint32_t n = 1u << job.logn, c0 = it.a0 * kBlock + (gi % per_item) * kCrtWaveCoefs;
        // fused: group c0/64 at quarter offsets; a group past n stores nothing
        w = job.fused && c0 < n ? CrtGroup{job.ra, job.c, n, c0 / 4u, n / 4u, job.outer}
                                : CrtGroup{job.ra, job.c, n, c0, 16u, 0u};
    });
}

template <int W>
__global__ void __launch_bounds__(W * 64) k_items_crt_mfma(Tables tb, BatchArgs ba, uint32_t first,
                                                           uint32_t count) {
    crt_items_loop<W>(tb, ba, first, count);
}

__global__ void __launch_bounds__(kBlock) k_items_crt(Tables tb, BatchArgs ba, uint32_t first) {
    const Item it = ba.items[first + blockIdx.x];
    const Job job = ba.jobs[it.job];
    const uint32_t n = 1u << job.logn;
    const uint32_t c = it.a0 * kBlock + threadIdx.x;
    if (c < n) crt_one(tb, job.ra, n, c, job.c);
}

// ---------------------------------------------------------------------------
// persistent batch: the whole work list in one launch
// ---------------------------------------------------------------------------
// One block of kPersistentThreads per compute unit runs the phases of the
// phased graph in turn -- conversion, fused tiles (largest size class first),
// CRT -- as grid-stride loops over their items, with a grid barrier between
// phases.  Every block runs the same sequence (no early exit before the last
// phase), so each barrier is reached by all blocks; the grid is one block per
// compute unit and a block holds most of a CU's LDS, so all blocks are
// resident at once.  A barrier wait is bounded by wall-clock ticks and a spin
// count: running out latches kErrTimeout in the error word and the block goes
// on (the results are then wrong and the host reports the timeout), so no
// wait can last.  One kernel is register-allocated for its heaviest phase
// (the MFMA CRT), which is what this mode costs against the phased graph.
constexpr uint32_t kPersistentMaxSpins = 1u << 26;

__device__ __forceinline__ void grid_barrier(uint32_t* counter, const GridSync& gs) {
    __threadfence();  // this thread's stores visible at device scope (L2 written back)
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t start = wall_clock64();
        for (uint32_t spins = 0;; ++spins) {
            if (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= gridDim.x) break;
            if (__hip_atomic_load(gs.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
            if (wall_clock64() - start > gs.wait_ticks || spins >= kPersistentMaxSpins) {
                __hip_atomic_fetch_or(gs.error, kErrTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
    }
    __syncthreads();
    __threadfence();  // and the other blocks' stores visible to this thread's loads (caches invalidated)
}

// Tile items of class TLOG and then of the smaller classes, block-strided;
// the block's threads are an item's tiles x threads per tile (the work list
// is built for it).
template <int TLOG>
__device__ __forceinline__ void tile_items_loop(const Tables& tb, const BatchArgs& ba, const PhaseRanges& r,
                                                uint32_t* lds) {
    for (uint32_t i = blockIdx.x; i < r.tile_count[TLOG]; i += gridDim.x) {
        __syncthreads();  // the previous item's last exchange may still be read by other waves
        tile_item<TLOG>(tb, ba, ba.items[r.tile_first[TLOG] + i], lds);
    }
    if constexpr (TLOG > 5) tile_items_loop<TLOG - 1>(tb, ba, r, lds);
}

__global__ void __launch_bounds__(kPersistentThreads) k_batch_persistent(Tables tb, BatchArgs ba, PhaseRanges r,
                                                                        GridSync gs) {
    extern __shared__ uint32_t lds[];
    ConvLds& L = *reinterpret_cast<ConvLds*>(lds);
    conv_tables_to_lds(tb, L);
    conv_items_loop(tb, ba, r.conv_first, r.conv_count, L);
    grid_barrier(gs.counters + 0, gs);
    tile_items_loop<kPersistentMaxTileLog>(tb, ba, r, lds);
    grid_barrier(gs.counters + 1, gs);
    crt_items_loop<kPersistentThreads / 64>(tb, ba, r.crt_first, r.crt_count);
}

__global__ void __launch_bounds__(64) k_clear_sync(GridSync gs) {
    if (threadIdx.x < kPersistentBarriers) gs.counters[threadIdx.x] = 0u;
    if (threadIdx.x == 0) *gs.error = 0u;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <int TLOG>
static hipError_t launch_tile(const Tables& tb, const Job& job, uint32_t u, hipStream_t stream) {
    using G = Geo<TLOG>;
    const uint32_t units = tb.k << u;
    uint32_t tiles = tile_block_threads(TLOG) >> G::TPTLOG;
    if (tiles > units) tiles = units;
    // fewer tiles per block while that leaves CUs idle: one product has only
    // k << u tiles
    while (tiles > 1 && (units + tiles - 1) / tiles < tb.cus && ((tiles >> 1) << G::TPTLOG) >= 64) tiles >>= 1;
    uint32_t threads = tiles << G::TPTLOG;
    if (threads < 64) {  // whole waves
        tiles = (64u >> G::TPTLOG) ? (64u >> G::TPTLOG) : 1u;
        threads = tiles << G::TPTLOG;
    }
    const uint32_t blocks = (units + tiles - 1) / tiles;
    const size_t lds = size_t(tiles) * G::STRIDE * sizeof(uint32_t);
    hipLaunchKernelGGL(k_conv_tile<TLOG>, dim3(blocks), dim3(threads), lds, stream, tb, job, u, units);
    return hipGetLastError();
}

static hipError_t launch_tile_dispatch(uint32_t tlog, const Tables& tb, const Job& job, uint32_t u,
                                       hipStream_t stream) {
    switch (tlog) {
        case 1: return launch_tile<1>(tb, job, u, stream);
        case 2: return launch_tile<2>(tb, job, u, stream);
        case 3: return launch_tile<3>(tb, job, u, stream);
        case 4: return launch_tile<4>(tb, job, u, stream);
        case 5: return launch_tile<5>(tb, job, u, stream);
        case 6: return launch_tile<6>(tb, job, u, stream);
        case 7: return launch_tile<7>(tb, job, u, stream);
        case 8: return launch_tile<8>(tb, job, u, stream);
        case 9: return launch_tile<9>(tb, job, u, stream);
        case 10: return launch_tile<10>(tb, job, u, stream);
        case 11: return launch_tile<11>(tb, job, u, stream);
        case 12: return launch_tile<12>(tb, job, u, stream);
        case 13: return launch_tile<13>(tb, job, u, stream);
        case 14: return launch_tile<14>(tb, job, u, stream);
        case 15: return launch_tile<15>(tb, job, u, stream);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// host-buffer path: conversion straight from page-locked host memory as the
// host fills it, and the result's return in chunks with ready flags
// ---------------------------------------------------------------------------
// Block 0 is a watcher: one lane polls the host-written ready flags in chunk
// order (a handful of PCIe reads per microsecond) and republishes each in
// device memory (io.mirror, one flag per 256 bytes so they sit in different
// L2 channels), where one lane per converting block polls it with growing
// sleeps -- thousands of waves polling one host word, or one L2 line, in a
// tight loop would crowd out the transfer itself.  The watcher publishes in
// chunk order, so a block whose waves straddle two chunks waits for the later.
// Every wait is bounded; a timeout latches kErrTimeout in the host-visible
// io.error and this product's seq in the device error word, and every later
// wait gives up at once.  Wave w of blocks 1.. converts group w of the
// concatenated a|b array (64 coefficients) once its chunk is published;
// job.a / job.b point at page-locked host memory.
__global__ void __launch_bounds__(kConvSmallBlock) k_ingest_convert(Tables tb, Job job, IoArgs io) {
    uint32_t* dev_error = io.mirror + kMaxIoChunks * kMirrorStride;
    if (blockIdx.x == 0) {
        if (threadIdx.x < kMaxIoChunks) io.counters[threadIdx.x] = 0u;  // egress chunk counters
        if (threadIdx.x == 0) {
            for (uint32_t j = 0; j < io.chunks; ++j) {
                const uint64_t start = wall_clock64();
                bool ok = true;
                // relaxed polls (an acquire per poll would invalidate caches
                // each time), one acquire once the flag is up
                while (__hip_atomic_load(io.ready + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != io.seq) {
                    if (wall_clock64() - start > io.wait_ticks) {
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                if (ok) (void)__hip_atomic_load(io.ready + j, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
                if (!ok) {
                    __hip_atomic_store(io.error, kErrTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    __hip_atomic_store(dev_error, io.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __hip_atomic_store(io.mirror + j * kMirrorStride, io.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        return;
    }
    __shared__ ConvLds L;
    __shared__ uint32_t s_ok;
    conv_tables_to_lds(tb, L);
    const uint32_t n = 1u << job.logn, per_poly = (n + kConvPerWave - 1) / kConvPerWave;
    const uint32_t groups = 2u * per_poly, waves = blockDim.x >> 6;
    const uint32_t g_first = (blockIdx.x - 1) * waves;
    const uint32_t g_last = min(g_first + waves, groups) - 1u;  // launch: g_first < groups
    const uint32_t* flag = io.mirror + (g_last / io.groups_per_chunk) * kMirrorStride;
    if (threadIdx.x == 0) {
        // the watcher's own waits are bounded by wait_ticks per chunk
        const uint64_t start = wall_clock64(), limit = io.wait_ticks * (io.chunks + 1);
        uint32_t ok = 1, nap = 1;
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != io.seq) {
            if (__hip_atomic_load(dev_error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == io.seq ||
                wall_clock64() - start > limit) {
                ok = 0;
                break;
            }
            // s_sleep takes an immediate: 64 x nap clocks, nap = 1, 2, 4 .. 32
            switch (nap) {
                case 1: __builtin_amdgcn_s_sleep(1); break;
                case 2: __builtin_amdgcn_s_sleep(2); break;
                case 4: __builtin_amdgcn_s_sleep(4); break;
                case 8: __builtin_amdgcn_s_sleep(8); break;
                case 16: __builtin_amdgcn_s_sleep(16); break;
                default: __builtin_amdgcn_s_sleep(32); break;
            }
            nap = nap < 32 ? 2 * nap : 32;
        }
        if (ok) (void)__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        s_ok = ok;
    }
    __syncthreads();  // the block's loads come after thread 0's acquire
    const uint32_t g = __builtin_amdgcn_readfirstlane(g_first + (threadIdx.x >> 6));
    if (!s_ok || g >= groups) return;
    const uint32_t poly = g / per_poly, c0 = (g % per_poly) * kConvPerWave;
    v4u v0[4], v1[4];
    conv_load_host(poly ? job.b : job.a, n, c0, threadIdx.x & 63u, v0, v1);
    conv_compute(tb, v0, v1, n, n, c0, poly ? job.rb : job.ra, L, threadIdx.x & 63u);
}

// Copies chunk `chunk` (chunk_words words, a multiple of 4) of src (device)
// to dst (page-locked host); the last block to finish sets done[chunk] = seq
// once every store of the chunk is visible to the host.  One launch per chunk,
// so chunks land in order.
// Gather copy of the serving path: block row y copies span segs[y], 16 bytes
// per lane where both ends are 16-byte aligned, else 4.  The span table is
// read where it lies (page-locked host memory is readable by kernels), so a
// launch needs no copy engine: every step of a served batch is a kernel on
// one queue.
__global__ void __launch_bounds__(256) k_copy_segments(const CopySeg* __restrict__ segs, uint32_t to_host) {
    const CopySeg sg = segs[blockIdx.y];
    const size_t lane0 = size_t(blockIdx.x) * blockDim.x + threadIdx.x, step = size_t(gridDim.x) * blockDim.x;
    if (((reinterpret_cast<uintptr_t>(sg.src) | reinterpret_cast<uintptr_t>(sg.dst)) & 15u) == 0) {
        const size_t n4 = sg.words / 4;
        const uint4* s4 = reinterpret_cast<const uint4*>(sg.src);
        uint4* d4 = reinterpret_cast<uint4*>(sg.dst);
        for (size_t i = lane0; i < n4; i += step) d4[i] = s4[i];
        for (size_t i = 4 * n4 + lane0; i < sg.words; i += step) sg.dst[i] = sg.src[i];
    } else {
        for (size_t i = lane0; i < sg.words; i += step) sg.dst[i] = sg.src[i];
    }
    if (to_host) __threadfence_system();  // the stores reach the host before the kernel's end is signalled
}

__global__ void __launch_bounds__(256) k_egress(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                size_t chunk_words, IoArgs io, uint32_t chunk) {
    const size_t first = size_t(chunk) * chunk_words / 4, count = chunk_words / 4;
    const uint4* s4 = reinterpret_cast<const uint4*>(src) + first;
    uint4* d4 = reinterpret_cast<uint4*>(dst) + first;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        d4[i] = s4[i];
    __threadfence_system();  // this wave's stores reach the host before the count
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t before = __hip_atomic_fetch_add(io.counters + chunk, 1u, __ATOMIC_ACQ_REL,
                                                       __HIP_MEMORY_SCOPE_AGENT);
        if (before + 1u == gridDim.x) {
            __threadfence_system();
            __hip_atomic_store(io.done + chunk, io.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

hipError_t launch_ingest_convert(const Tables& tb, const Job& job, const IoArgs& io, hipStream_t stream) {
    if (tb.quartic) return hipErrorInvalidValue;  // the quartic backend converts from device memory
    const uint32_t n = 1u << job.logn;
    const uint32_t groups = 2u * ((n + kConvPerWave - 1) / kConvPerWave);
    const uint32_t waves_per_block = kConvSmallBlock / kWaveSize;
    const uint32_t blocks = 1 + (groups + waves_per_block - 1) / waves_per_block;  // + the watcher
    if (!io.groups_per_chunk || io.chunks > kMaxIoChunks ||
        (groups + io.groups_per_chunk - 1) / io.groups_per_chunk != io.chunks)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ingest_convert, dim3(blocks), dim3(kConvSmallBlock), 0, stream, tb, job, io);
    return hipGetLastError();
}

hipError_t launch_copy_segments(const CopySeg* segs, uint32_t n_segs, uint64_t max_words, bool to_host,
                                hipStream_t stream) {
    if (n_segs == 0) return hipSuccess;
    if (!segs || n_segs > 65535u) return hipErrorInvalidValue;
    // up to 64 blocks of 1024 words per pass for the longest span
    uint64_t blocks = (max_words + 1023) / 1024;
    blocks = blocks < 1 ? 1 : (blocks > 64 ? 64 : blocks);
    hipLaunchKernelGGL(k_copy_segments, dim3(uint32_t(blocks), n_segs), dim3(256), 0, stream, segs,
                       to_host ? 1u : 0u);
    return hipGetLastError();
}

hipError_t launch_egress(const uint32_t* src, uint32_t* dst_host, size_t chunk_words, uint32_t chunks,
                         const IoArgs& io, hipStream_t stream) {
    if (!chunks || chunks > kMaxIoChunks || chunk_words % 4 ||
        ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst_host)) & 15u))
        return hipErrorInvalidValue;
    // ~64 blocks keep a PCIe transfer at its rate; fewer for small chunks
    size_t blocks = (chunk_words / 4 + 1023) / 1024;
    blocks = blocks < 1 ? 1 : (blocks > 64 ? 64 : blocks);
    for (uint32_t j = 0; j < chunks; ++j) {
        hipLaunchKernelGGL(k_egress, dim3(uint32_t(blocks)), dim3(256), 0, stream, src, dst_host, chunk_words, io, j);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Slices of the primes (over blockIdx.z) of a plain-path conversion of
// `waves` waves: the job's choice (1, 2 or 4), else up to 4 while that leaves
// fewer than ~4 waves per CU, so a single product (or one chunk of it) still
// spreads over the whole device
static uint32_t conv_slices(const Tables& tb, const Job& job, uint32_t waves) {
    if (job.conv_slices == 1 || job.conv_slices == 2 || job.conv_slices == 4)
        return job.conv_slices < tb.quads ? job.conv_slices : 1u;
    uint32_t slices = 1;
    while (slices < 4 && slices < tb.quads && waves * slices < 4 * tb.cus) slices *= 2;
    return slices;
}

hipError_t launch_convert_range(const Tables& tb, const Job& job, uint32_t c_first, uint32_t c_end,
                                uint32_t poly_first, uint32_t polys, hipStream_t stream) {
    const uint32_t n = 1u << job.logn;
    if (c_first >= c_end || c_end > n || polys < 1 || poly_first + polys > 2) return hipErrorInvalidValue;
    if (tb.quartic) {  // whole range, both polynomials, outer stages as passes of their own
        if (c_first != 0 || c_end != n || poly_first != 0 || polys != 2 || job.fused) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_convert_quartic, dim3((n + 255) / 256, 2), dim3(256), 0, stream, tb, job);
        return hipGetLastError();
    }
    if (job.fused && (c_first != 0 || c_end != n || !outer_fusable(job.logn, job.outer))) return hipErrorInvalidValue;
    if (job.fused) {
        if (poly_first != 0 || polys != 2) return hipErrorInvalidValue;
        // 4-wave blocks of one slot (128 coefficients) per wave, the primes
        // split over conv_slices slices
        const uint32_t slots = n / kConvSlotCoefs, threads = kConvSmallBlock;
        const uint32_t per_block = threads / kWaveSize, blocks = (slots + per_block - 1) / per_block;
        const uint32_t waves = blocks * polys * per_block;
        const uint32_t slices = conv_slices(tb, job, waves);
        hipLaunchKernelGGL(k_convert_fused, dim3(blocks, 2, slices), dim3(threads), 0, stream, tb, job);
        return hipGetLastError();
    }
    const uint32_t count = c_end - c_first;
    // 4-wave blocks, the primes split over conv_slices slices
    uint32_t threads = kConvSmallBlock;
    if (count < threads) threads = ((count + kConvPerWave - 1) / kConvPerWave) * kWaveSize;
    const uint32_t blocks = (count + threads - 1) / threads;
    const uint32_t waves = blocks * polys * (threads / kWaveSize);
    const uint32_t slices = conv_slices(tb, job, waves);
    hipLaunchKernelGGL(k_convert, dim3(blocks, polys, slices), dim3(threads), 0, stream, tb, job, c_first, c_end,
                       poly_first);
    return hipGetLastError();
}

hipError_t launch_convert(const Tables& tb, const Job& job, hipStream_t stream) {
    return launch_convert_range(tb, job, 0, 1u << job.logn, 0, 2, stream);
}

hipError_t launch_ntt(const Tables& tb, const Job& job, uint32_t u, hipStream_t stream) {
    if (u > job.logn || job.logn - u > 15 || u > 2) return hipErrorInvalidValue;
    if (job.fused && (u != job.outer || !outer_fusable(job.logn, u))) return hipErrorInvalidValue;
    const uint32_t tlog = job.logn - u;
    if (u > 0 && !job.fused) {  // fused: the conversion and the CRT run the outer stages
        const uint32_t T = 1u << tlog;
        const dim3 grid((T + 255) / 256, 2 * tb.k);
        if (u == 1) hipLaunchKernelGGL(k_outer_forward<1>, grid, dim3(256), 0, stream, tb, job);
        else hipLaunchKernelGGL(k_outer_forward<2>, grid, dim3(256), 0, stream, tb, job);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipError_t e = launch_tile_dispatch(tlog, tb, job, u, stream);
    if (e != hipSuccess) return e;
    if (u > 0 && !job.fused) {
        const uint32_t T = 1u << tlog;
        const dim3 grid((T + 255) / 256, tb.k);
        if (u == 1) hipLaunchKernelGGL(k_outer_inverse<1>, grid, dim3(256), 0, stream, tb, job);
        else hipLaunchKernelGGL(k_outer_inverse<2>, grid, dim3(256), 0, stream, tb, job);
        e = hipGetLastError();
    }
    return e;
}

// blocks of `tiles` tiles (whole waves; padding tiles, if any, store nothing):
// fewer tiles per block leave room for several blocks per CU, so one block's
// loads and barriers overlap another's arithmetic
template <int TLOG, int EL>
static hipError_t launch_items_tile(const Tables& tb, const BatchArgs& ba, uint32_t first, uint32_t count,
                                    uint32_t tiles, hipStream_t stream) {
    using G = Geo<TLOG, EL>;
    while ((tiles << G::TPTLOG) < kWaveSize) tiles <<= 1;
    if ((tiles << G::TPTLOG) > tile_block_threads(TLOG)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_items_tile<TLOG, EL>), dim3(count), dim3(tiles << G::TPTLOG),
                       size_t(tiles) * G::STRIDE * sizeof(uint32_t), stream, tb, ba, first);
    return hipSuccess;
}

// a tile class at 32 residues per thread, or 16 (kind bit 20) where compiled
template <int TLOG>
static hipError_t launch_items_tile_rpt(const Tables& tb, const BatchArgs& ba, uint32_t first, uint32_t count,
                                        uint32_t tiles, bool sixteen, hipStream_t stream) {
    if constexpr (TLOG >= kRpt16MinLog && TLOG <= kRpt16MaxLog)
        if (sixteen) return launch_items_tile<TLOG, 4>(tb, ba, first, count, tiles, stream);
    if (sixteen) return hipErrorInvalidValue;
    return launch_items_tile<TLOG, tile_elog(TLOG)>(tb, ba, first, count, tiles, stream);
}

// W-wave CRT blocks, as many as the LDS lets share a CU (two of 2 waves,
// one of 4 or 8), looping over the groups
template <int W>
static void launch_items_crt(const Tables& tb, const BatchArgs& ba, uint32_t first, uint32_t count,
                             uint32_t groups, hipStream_t stream) {
    constexpr uint32_t per_cu = crt_block_lds<W>() * 2 <= kLdsPerCu ? 2u : 1u;
    uint32_t grid = (groups + W - 1) / W;
    grid = grid < per_cu * tb.cus ? grid : per_cu * tb.cus;
    hipLaunchKernelGGL(k_items_crt_mfma<W>, dim3(grid), dim3(W * 64), crt_block_lds<W>(), stream, tb, ba, first,
                       count);
}

hipError_t launch_phase(const Tables& tb, const BatchArgs& ba, uint32_t kind, uint32_t first, uint32_t count,
                        hipStream_t stream) {
    if (!count) return hipSuccess;
    if (tb.quartic) {  // the quartic backend's own conversion and CRT; tiles and outer passes are shared
        switch (kind & 0xFFu) {
            case kPhaseConv:
            case kPhaseConvScalar:
                hipLaunchKernelGGL(k_items_conv_quartic, dim3(count), dim3(256), 0, stream, tb, ba, first);
                return hipGetLastError();
            case kPhaseCrt:
            case kPhaseCrtScalar:
                hipLaunchKernelGGL(k_items_crt_quartic, dim3(count), dim3(256), 0, stream, tb, ba, first);
                return hipGetLastError();
            default:
                break;
        }
    }
    switch (kind & 0xFFu) {
        case kPhaseConv: {
            // 4-wave blocks, up to 8 per CU, each looping over slots: measured
            // 2 % faster for the whole batch graph than 1024-thread blocks
            const uint32_t grid = count * 4 < 8 * tb.cus ? count * 4 : 8 * tb.cus;
            hipLaunchKernelGGL(k_items_conv, dim3(grid), dim3(kConvSmallBlock), 0, stream, tb, ba, first, count);
            break;
        }
        case kPhaseConvScalar:
            hipLaunchKernelGGL(k_items_conv_scalar, dim3(count), dim3(kBlock), 0, stream, tb, ba, first);
            break;
        case kPhaseOuterFwd:
        case kPhaseOuterInv:
            hipLaunchKernelGGL(k_items_outer, dim3(count), dim3(kBlock), 0, stream, tb, ba, first);
            break;
        case kPhaseCrt: {
            const uint32_t groups = count * (kBlock / kCrtWaveCoefs);
            switch ((kind >> 16) & 0xFu) {
                case 1: launch_items_crt<2>(tb, ba, first, count, groups, stream); break;
                case 2: launch_items_crt<4>(tb, ba, first, count, groups, stream); break;
                case 0:  // the default
                case 3: launch_items_crt<8>(tb, ba, first, count, groups, stream); break;
                default: return hipErrorInvalidValue;
            }
            break;
        }
        case kPhaseCrtScalar:
            hipLaunchKernelGGL(k_items_crt, dim3(count), dim3(kBlock), 0, stream, tb, ba, first);
            break;
        case kPhaseTile: {
            const uint32_t tiles = 1u << ((kind >> 16) & 0xFu);
            const bool r16 = ((kind >> 20) & 1u) != 0;
            hipError_t e = hipSuccess;
            switch ((kind >> 8) & 0xFFu) {
                case 5: e = launch_items_tile_rpt<5>(tb, ba, first, count, tiles, r16, stream); break;
                case 6: e = launch_items_tile_rpt<6>(tb, ba, first, count, tiles, r16, stream); break;
                case 7: e = launch_items_tile_rpt<7>(tb, ba, first, count, tiles, r16, stream); break;
                case 8: e = launch_items_tile_rpt<8>(tb, ba, first, count, tiles, r16, stream); break;
                case 9: e = launch_items_tile_rpt<9>(tb, ba, first, count, tiles, r16, stream); break;
                case 10: e = launch_items_tile_rpt<10>(tb, ba, first, count, tiles, r16, stream); break;
                case 11: e = launch_items_tile_rpt<11>(tb, ba, first, count, tiles, r16, stream); break;
                case 12: e = launch_items_tile_rpt<12>(tb, ba, first, count, tiles, r16, stream); break;
                case 13: e = launch_items_tile_rpt<13>(tb, ba, first, count, tiles, r16, stream); break;
                case 14: e = launch_items_tile_rpt<14>(tb, ba, first, count, tiles, r16, stream); break;
                case 15: e = launch_items_tile_rpt<15>(tb, ba, first, count, tiles, r16, stream); break;
                default: return hipErrorInvalidValue;
            }
            if (e != hipSuccess) return e;
            break;
        }
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

std::vector<Phase> phase_list(const PhaseRanges& r, bool scalar_conversion, bool scalar_crt, uint32_t crt_waves) {
    uint32_t crt_lg = 0;  // log2 of the CRT's waves per block (0: the default)
    if (crt_waves == 2 || crt_waves == 4 || crt_waves == 8) crt_lg = uint32_t(__builtin_ctz(crt_waves));
    std::vector<Phase> out;
    if (r.conv_count) out.push_back({scalar_conversion ? kPhaseConvScalar : kPhaseConv, r.conv_first, r.conv_count});
    if (r.fwd_count) out.push_back({kPhaseOuterFwd, r.fwd_first, r.fwd_count});
    for (uint32_t tlog = 15; tlog >= 5; --tlog) {
        if (!r.tile_count[tlog]) continue;
        uint32_t lg = 0;  // tiles per block: the class's largest item, rounded up to a power of two
        while ((1u << lg) < r.tile_units[tlog]) ++lg;
        out.push_back({kPhaseTile | (tlog << 8) | (lg << 16) | ((r.tile_rpt16[tlog] ? 1u : 0u) << 20),
                       r.tile_first[tlog], r.tile_count[tlog]});
    }
    if (r.inv_count) out.push_back({kPhaseOuterInv, r.inv_first, r.inv_count});
    if (r.crt_count)
        out.push_back({scalar_crt ? kPhaseCrtScalar : kPhaseCrt | (crt_lg << 16), r.crt_first, r.crt_count});
    return out;
}

hipError_t launch_phased(const Tables& tb, const BatchArgs& ba, const PhaseRanges& r, bool scalar_conversion,
                         bool scalar_crt, uint32_t crt_waves, hipStream_t stream) {
    for (const Phase& ph : phase_list(r, scalar_conversion, scalar_crt, crt_waves)) {
        const hipError_t e = launch_phase(tb, ba, ph.kind, ph.first, ph.count, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_convert_scalar(const Tables& tb, const Job& job, hipStream_t stream) {
    if (tb.quartic) return launch_convert(tb, job, stream);  // the quartic conversion is scalar already
    const uint32_t n = 1u << job.logn;
    hipLaunchKernelGGL(k_convert_scalar, dim3((n + 255) / 256, 2), dim3(256), 0, stream, tb, job);
    return hipGetLastError();
}

hipError_t launch_crt_scalar(const Tables& tb, const Job& job, hipStream_t stream) {
    if (tb.quartic) return launch_crt(tb, job, stream);  // one CRT on the quartic backend
    const uint32_t n = 1u << job.logn;
    const uint32_t threads = n < kCrtPerBlock ? ((n + 63) / 64) * 64 : kCrtPerBlock;
    hipLaunchKernelGGL(k_crt, dim3((n + threads - 1) / threads), dim3(threads), 0, stream, tb, job);
    return hipGetLastError();
}

hipError_t launch_crt_range(const Tables& tb, const Job& job, uint32_t c_first, uint32_t c_end,
                            hipStream_t stream) {
    const uint32_t n = 1u << job.logn;
    if (c_first >= c_end || c_end > n || c_first % kCrtWaveCoefs || (c_end != n && c_end % kCrtWaveCoefs))
        return hipErrorInvalidValue;
    if (job.fused && (c_first != 0 || c_end != n || !outer_fusable(job.logn, job.outer))) return hipErrorInvalidValue;
    if (tb.quartic) {
        if (c_first != 0 || c_end != n || job.fused) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_crt_quartic, dim3((n + 255) / 256), dim3(256), 0, stream, tb, job);
        return hipGetLastError();
    }
    const uint32_t groups = (c_end - c_first + kCrtWaveCoefs - 1) / kCrtWaveCoefs;
    uint32_t grid = (groups + kCrtPlainWaves - 1) / kCrtPlainWaves;
    grid = grid < tb.cus ? grid : tb.cus;
    hipLaunchKernelGGL(k_crt_mfma, dim3(grid), dim3(kCrtPlainWaves * 64), crt_block_lds<kCrtPlainWaves>(), stream,
                       tb, job, c_first, c_end);
    return hipGetLastError();
}

hipError_t launch_crt(const Tables& tb, const Job& job, hipStream_t stream) {
    return launch_crt_range(tb, job, 0, 1u << job.logn, stream);
}

// LDS of a persistent block: the largest of the conversion tables, a tile
// phase's tiles (kPersistentThreads x 32 residues, padded) and the 8-wave CRT
template <int TLOG>
constexpr uint32_t persistent_tile_words() {
    using G = Geo<TLOG>;
    const uint32_t here = (kPersistentThreads >> G::TPTLOG) * uint32_t(G::STRIDE);
    if constexpr (TLOG > 5) {
        const uint32_t below = persistent_tile_words<TLOG - 1>();
        return here > below ? here : below;
    } else {
        return here;
    }
}
constexpr size_t persistent_lds() {
    size_t m = sizeof(ConvLds);
    const size_t t = size_t(persistent_tile_words<kPersistentMaxTileLog>()) * sizeof(uint32_t);
    m = t > m ? t : m;
    const size_t c = crt_block_lds<kPersistentThreads / 64>();
    return c > m ? c : m;
}
static_assert(persistent_lds() <= kLdsPerCu, "a persistent block fits the LDS of a CU");
static_assert(kPersistentThreads / 64 == 8, "the persistent CRT phase runs 8-wave groups");

hipError_t persistent_blocks(uint32_t cus, uint32_t* blocks) {
    int per_cu = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(&k_batch_persistent), int(kPersistentThreads), persistent_lds());
    if (e != hipSuccess) return e;
    // one block per CU: the LDS holds no second one, and more blocks would
    // only wait at the barriers
    *blocks = per_cu >= 1 ? cus : 0u;
    return hipSuccess;
}

hipError_t launch_persistent(const Tables& tb, const BatchArgs& ba, const PhaseRanges& r, const GridSync& gs,
                             uint32_t blocks, hipStream_t stream) {
    if (!blocks || tb.quartic || r.fwd_count || r.inv_count || r.tile_count[15]) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_clear_sync, dim3(1), dim3(64), 0, stream, gs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_persistent, dim3(blocks), dim3(kPersistentThreads), persistent_lds(), stream, tb, ba,
                       r, gs);
    return hipGetLastError();
}

}  // namespace pmd

namespace pmd {
// dynamic LDS of the largest tile block: 4 tiles of 2^13 (each lds_pad<13>(8191) + 1
// words) or one tile of 2^15
constexpr uint32_t kTileLdsMaxWords =
    uint32_t(4 * Geo<13>::STRIDE) > kLdsWords ? uint32_t(4 * Geo<13>::STRIDE) : kLdsWords;
static_assert(kTileLdsMaxWords * 4 <= kLdsPerCu, "tile blocks fit the LDS of a CU");
static_assert(crt_block_lds<8>() <= kLdsPerCu, "an 8-wave CRT block fits the LDS of a CU");
template <int TLOG>
static hipError_t allow_lds() {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv_tile<TLOG>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(kTileLdsMaxWords * sizeof(uint32_t)));
}
template <int TLOG>
static hipError_t allow_items_lds() {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_items_tile<TLOG, tile_elog(TLOG)>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(kTileLdsMaxWords * sizeof(uint32_t)));
    if constexpr (TLOG >= kRpt16MinLog && TLOG <= kRpt16MaxLog)
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_items_tile<TLOG, 4>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(kTileLdsMaxWords * sizeof(uint32_t)));
    return e;
}

hipError_t configure_kernels() {
    hipError_t e = hipSuccess;
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_crt_mfma),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 int(crt_block_lds<kCrtPlainWaves>()))) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_items_crt_mfma<2>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(crt_block_lds<2>()))) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_items_crt_mfma<4>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(crt_block_lds<4>()))) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_items_crt_mfma<8>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(crt_block_lds<8>()))) != hipSuccess)
        return e;
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_persistent),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, int(persistent_lds()))) != hipSuccess)
        return e;
    if ((e = allow_lds<1>()) != hipSuccess) return e;
    if ((e = allow_lds<2>()) != hipSuccess) return e;
    if ((e = allow_lds<3>()) != hipSuccess) return e;
    if ((e = allow_lds<4>()) != hipSuccess) return e;
    if ((e = allow_lds<5>()) != hipSuccess) return e;
    if ((e = allow_lds<6>()) != hipSuccess) return e;
    if ((e = allow_lds<7>()) != hipSuccess) return e;
    if ((e = allow_lds<8>()) != hipSuccess) return e;
    if ((e = allow_lds<9>()) != hipSuccess) return e;
    if ((e = allow_lds<10>()) != hipSuccess) return e;
    if ((e = allow_lds<11>()) != hipSuccess) return e;
    if ((e = allow_lds<12>()) != hipSuccess) return e;
    if ((e = allow_lds<13>()) != hipSuccess) return e;
    if ((e = allow_lds<14>()) != hipSuccess) return e;
    if ((e = allow_lds<15>()) != hipSuccess) return e;
    if ((e = allow_items_lds<5>()) != hipSuccess) return e;
    if ((e = allow_items_lds<6>()) != hipSuccess) return e;
    if ((e = allow_items_lds<7>()) != hipSuccess) return e;
    if ((e = allow_items_lds<8>()) != hipSuccess) return e;
    if ((e = allow_items_lds<9>()) != hipSuccess) return e;
    if ((e = allow_items_lds<10>()) != hipSuccess) return e;
    if ((e = allow_items_lds<11>()) != hipSuccess) return e;
    if ((e = allow_items_lds<12>()) != hipSuccess) return e;
    if ((e = allow_items_lds<13>()) != hipSuccess) return e;
    if ((e = allow_items_lds<14>()) != hipSuccess) return e;
    if ((e = allow_items_lds<15>()) != hipSuccess) return e;
    return hipSuccess;
}
}  // namespace pmd
