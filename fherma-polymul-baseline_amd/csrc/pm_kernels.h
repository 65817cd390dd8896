// Launchers of the gfx950 kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pm_device.h"

namespace pmd {

// host-buffer path: conversion of a|b from page-locked host memory, group by
// group as the host flags input chunks ready (also clears io.counters)
hipError_t launch_ingest_convert(const Tables& tb, const Job& job, const IoArgs& io, hipStream_t stream);
// device src -> page-locked host dst in `chunks` chunks, io.done[j] = io.seq
// once chunk j is visible to the host
hipError_t launch_egress(const uint32_t* src, uint32_t* dst_host, size_t chunk_words, uint32_t chunks,
                         const IoArgs& io, hipStream_t stream);
// copies every span of segs[0 .. n_segs) (segs may be page-locked host
// memory: the kernel reads it); max_words bounds the longest span.  With
// to_host the stores are made visible to the host before the kernel ends.
hipError_t launch_copy_segments(const CopySeg* segs, uint32_t n_segs, uint64_t max_words, bool to_host,
                                hipStream_t stream);
// residues of a and b for all primes (MFMA conversion)
hipError_t launch_convert(const Tables& tb, const Job& job, hipStream_t stream);
// the same for coefficients [c_first, c_end) of polynomials poly_first ..
// poly_first + polys - 1 (0 = a, 1 = b)
hipError_t launch_convert_range(const Tables& tb, const Job& job, uint32_t c_first, uint32_t c_end,
                                uint32_t poly_first, uint32_t polys, hipStream_t stream);
// residues by scalar limb reduction (comparison path, same output contract)
hipError_t launch_convert_scalar(const Tables& tb, const Job& job, hipStream_t stream);
// forward NTTs, pointwise product, inverse NTT, scale; u outer radix-2 stages
// run as separate passes (tile size 2^(logn-u) <= 2^15)
hipError_t launch_ntt(const Tables& tb, const Job& job, uint32_t u, hipStream_t stream);
// CRT and reduction modulo q into 28-limb coefficients (int8 MFMA sums)
hipError_t launch_crt(const Tables& tb, const Job& job, hipStream_t stream);
// the same for coefficients [c_first, c_end): c_first a multiple of 64, c_end a
// multiple of 64 or n (c may be mapped host memory)
hipError_t launch_crt_range(const Tables& tb, const Job& job, uint32_t c_first, uint32_t c_end,
                            hipStream_t stream);
// the same by 64-bit multiply-accumulate on the vector ALUs (comparison path)
hipError_t launch_crt_scalar(const Tables& tb, const Job& job, hipStream_t stream);
// Contiguous item ranges of a work list by phase (items are ordered by type).
struct PhaseRanges {
    uint32_t conv_first = 0, conv_count = 0;
    uint32_t fwd_first = 0, fwd_count = 0;
    uint32_t tile_first[16] = {}, tile_count[16] = {};
    uint32_t tile_units[16] = {};  // most units (tiles) in one item of the class: tiles per block
    uint32_t tile_rpt16[16] = {};  // 1: the class runs 16 residues per thread (else 32)
    uint32_t inv_first = 0, inv_count = 0;
    uint32_t crt_first = 0, crt_count = 0;
};
// phase kinds (tile phases carry the tile size class, log2 of the tiles
// per block and 16 residues per thread: kPhaseTile | tlog << 8 |
// log2(tiles) << 16 | rpt16 << 20)
enum PhaseKind : uint32_t {
    kPhaseConv = 0, kPhaseOuterFwd = 1, kPhaseTile = 2, kPhaseOuterInv = 3, kPhaseCrt = 4, kPhaseConvScalar = 5,
    kPhaseCrtScalar = 6
};
struct Phase {
    uint32_t kind, first, count;
};
// crt_waves: waves per block of the MFMA CRT (2, 4 or 8; anything else: 8),
// carried in the CRT phase's kind as kPhaseCrt | log2(waves) << 16
std::vector<Phase> phase_list(const PhaseRanges& ranges, bool scalar_conversion, bool scalar_crt,
                              uint32_t crt_waves = 0);
hipError_t launch_phase(const Tables& tb, const BatchArgs& ba, uint32_t kind, uint32_t first, uint32_t count,
                        hipStream_t stream);
// one launch per phase (and per tile size class) over the same work list
hipError_t launch_phased(const Tables& tb, const BatchArgs& ba, const PhaseRanges& ranges, bool scalar_conversion,
                         bool scalar_crt, uint32_t crt_waves, hipStream_t stream);
// The persistent batch: clears gs, then one launch of `blocks` blocks runs the
// conversion, tile and CRT phases of the work list with grid barriers between
// them (the work list built for kPersistentThreads-thread tile blocks, tiles
// of at most 2^kPersistentMaxTileLog, outer stages fused, general backend).
hipError_t launch_persistent(const Tables& tb, const BatchArgs& ba, const PhaseRanges& ranges, const GridSync& gs,
                             uint32_t blocks, hipStream_t stream);
// blocks of the persistent kernel that are resident at once on `cus` compute
// units (0 when none fits)
hipError_t persistent_blocks(uint32_t cus, uint32_t* blocks);
// device address of the static residue scratch (kScratchWords words) on
// `device`, which must be the current device; looked up once per device
hipError_t scratch_address(int device, uint32_t** out);
// raise the dynamic LDS limit of the tile kernels (call once per device)
hipError_t configure_kernels();

}  // namespace pmd
