// Runtime support shared by every emitted mechanism library.
//
// Included once by each generated crs_<mech>.hip: per-thread error text,
// argument checks and launch helpers behind the C ABI of include/crs_hip.h.
// Device code here is mechanism independent.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>

#include "crs_hip.h"

namespace crs {

// Threads per block for the per-state kernels: four wave64 wavefronts.
constexpr int kBlock = 256;

inline char *error_buffer()
{
    static thread_local char buf[512] = {0};
    return buf;
}

inline int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(error_buffer(), 512, fmt, ap);
    va_end(ap);
    return code;
}

inline void clear_error() { error_buffer()[0] = '\0'; }

// Checks shared by all batched entry points.  A null pointer is only an
// error when there is work to do.
inline int check_sizes(const char *fn, int64_t n, int64_t ld)
{
    if (n < 0)
        return fail(CRS_ERR_ARG, "%s: n must be >= 0 (got %lld)", fn,
                    (long long)n);
    if (ld < n)
        return fail(CRS_ERR_ARG, "%s: ld must be >= n (ld=%lld, n=%lld)", fn,
                    (long long)ld, (long long)n);
    return CRS_OK;
}

inline int check_ptr(const char *fn, const void *p, const char *what, int64_t n)
{
    if (n > 0 && p == nullptr)
        return fail(CRS_ERR_ARG, "%s: %s is NULL", fn, what);
    return CRS_OK;
}

inline int launched(const char *fn)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(CRS_ERR_HIP, "%s: kernel launch failed: %s", fn,
                    hipGetErrorString(e));
    return CRS_OK;
}

inline dim3 grid_for(int64_t n)
{
    return dim3((unsigned)((n + kBlock - 1) / kBlock));
}

}  // namespace crs

#define CRS_TRY(expr)                 \
    do {                              \
        int crs_rc_ = (expr);         \
        if (crs_rc_ != CRS_OK)        \
            return crs_rc_;           \
    } while (0)
