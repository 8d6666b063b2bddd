// errors.hpp — status reporting shared by the host translation units of libcubo_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/cubo_hip.h"

namespace cubo {

inline thread_local std::string g_last_error;

inline int fail(int status, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

inline int hip_fail(hipError_t e, const char *what) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString(e));
    return fail(CUBO_HIP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace cubo

#define CUBO_HIP_TRY(expr, what)                                  \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) return ::cubo::hip_fail(_e, what);  \
    } while (0)
