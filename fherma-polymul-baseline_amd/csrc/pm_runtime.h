// Internal definition of the opaque pm_context and shared runtime helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "pm_device.h"
#include "pm_params.h"

struct pm_context {
    int device = 0;
    uint32_t max_logn = 0;
    pm::Params params;
    pmd::Tables tables{};
    std::vector<void*> allocations;
    uint64_t table_bytes = 0, workspace_bytes = 0;
    hipStream_t stream = nullptr;
    // plain-path workspace
    uint32_t *d_a = nullptr, *d_b = nullptr, *d_c = nullptr, *d_ra = nullptr, *d_rb = nullptr;
    // outer radix-2 stages run as separate passes, per log2 N (plain path)
    uint32_t outer[pm::kMaxLogN + 1] = {};
    std::string last_error;

    void* device_alloc(size_t bytes);
};

int pm_fail(pm_context* ctx, int status, const std::string& why);
int pm_hip(pm_context* ctx, hipError_t e, const char* what);
int pm_check_n(pm_context* ctx, uint32_t n, uint32_t* logn);
int pm_run_plain(pm_context* ctx, const pmd::Job& job, uint32_t outer, hipStream_t stream);
uint32_t pm_default_outer(uint32_t logn);
