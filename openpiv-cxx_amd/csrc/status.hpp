// status.hpp -- the per-thread error text and status helpers of the C ABI,
// with no HIP dependency, so that the host-only parts of the library (image
// decoding, descriptor checks) also build as plain C++ (sanitizer builds).
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "openpiv_amd.h"

namespace opiv::detail {

inline thread_local std::string g_last_error;

inline int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
inline int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

inline bool is_pow2(uint64_t v) { return v && (v == (v & (~v + 1))); }   // util.h:67-70

inline size_t dtype_size(int dt)
{
    switch (dt) {
    case OPIV_U8: return 1;
    case OPIV_U16: return 2;
    case OPIV_F32: return 4;
    case OPIV_F64: return 8;
    default: return 0;
    }
}

// relative gap under which two FP32 correlation values count as tied for
// OPIV_FLAG_NEAR_TIE
constexpr float kTieTol = 2e-6f;

} // namespace opiv::detail
