// ctx.hpp -- host-side internals shared by the C ABI translation units:
// HIP status helpers, growable device buffers and the context and plan
// objects behind the opaque handles of openpiv_amd.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "openpiv_amd.h"
#include "status.hpp"

namespace opiv::detail {

inline int hip_fail(hipError_t e, const char* what)
{
    return fail(OPIV_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(call)                                        \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(e_, #call);    \
    } while (0)

// device buffer that grows on demand
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

inline bool is_device_ptr(const void* ptr)
{
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();   // unregistered host memory: clear the sticky error
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

} // namespace opiv::detail

struct opiv_ctx {
    int device = 0;
    int cu_count = 0;
    std::string arch;
    hipStream_t stream = nullptr;
    opiv::detail::DevBuf stage_a, stage_b, stage_out;
    opiv::detail::DevBuf stage_fwd, stage_cot, stage_ga, stage_gb, stage_var;
    opiv::detail::DevBuf adj_coef;   // VJP tap weights, 32 B per window
    void* scratch_coef = nullptr;    // this device's copy of the allocation-free VJP's scratch array
    // background subtraction: staged host frames and outputs, mean, pointer
    // tables, per-frame range keys
    opiv::detail::DevBuf bg_frames, bg_out, bg_mean, bg_table, bg_range;
};

struct opiv_plan {
    opiv_ctx* ctx = nullptr;
    opiv_plan_desc desc{};
    uint32_t off_x = 0, off_y = 0;
    uint32_t nx = 0, ny = 0;
    int32_t x_start = 0, y_start = 0;
};
