// Launchers of the gfx950 kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "pm_device.h"

namespace pmd {

// residues of a and b for all primes (MFMA conversion)
hipError_t launch_convert(const Tables& tb, const Job& job, hipStream_t stream);
// forward NTTs, pointwise product, inverse NTT, scale; u outer radix-2 stages
// run as separate passes (tile size 2^(logn-u) <= 2^15)
hipError_t launch_ntt(const Tables& tb, const Job& job, uint32_t u, hipStream_t stream);
// CRT and reduction modulo q into 28-limb coefficients
hipError_t launch_crt(const Tables& tb, const Job& job, hipStream_t stream);
// reset counters, then one persistent launch over the batch work list
hipError_t launch_persistent(const Tables& tb, const BatchArgs& ba, uint32_t* reset_words, uint32_t reset_count,
                             uint32_t blocks, hipStream_t stream);
// Contiguous item ranges of a work list by phase (items are ordered by type).
struct PhaseRanges {
    uint32_t conv_first = 0, conv_count = 0;
    uint32_t fwd_first = 0, fwd_count = 0;
    uint32_t tile_first[16] = {}, tile_count[16] = {};
    uint32_t inv_first = 0, inv_count = 0;
    uint32_t crt_first = 0, crt_count = 0;
};
// one launch per phase (and per tile size class) over the same work list
hipError_t launch_phased(const Tables& tb, const BatchArgs& ba, const PhaseRanges& ranges, hipStream_t stream);
// raise the dynamic LDS limit of the tile kernels (call once per device)
hipError_t configure_kernels();

}  // namespace pmd
