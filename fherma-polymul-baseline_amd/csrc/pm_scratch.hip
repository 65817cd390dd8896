// The static residue scratch of pm_multiply_device_noalloc: a file-scope
// device array (each GPU the library runs on holds its own copy until the
// process ends), so that entry point never allocates.  Host code never passes
// the array's name to a kernel or a copy -- on the host that name is only a
// stand-in -- but the device address hipGetSymbolAddress returns for the
// current device, looked up once per device and kept.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pm_device.h"
#include "pm_kernels.h"

namespace pmd {

static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_address(int device, uint32_t** out) {
    if (!out || device < 0 || device >= kMaxDevices) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int current = -1;
        hipError_t e = hipGetDevice(&current);
        if (e == hipSuccess && current != device) e = hipErrorInvalidDevice;  // the caller makes `device` current
        void* ptr = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<uint32_t*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace pmd
