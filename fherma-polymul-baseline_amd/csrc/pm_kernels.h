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
// raise the dynamic LDS limit of the tile kernels (call once per device)
hipError_t configure_kernels();

}  // namespace pmd
