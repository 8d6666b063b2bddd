// cubo_hip.hip — the C ABI of include/cubo_hip.h: devices, the HBM-resident cube, and the fused
// reduction's host side (weight preparation, kernel choice, scratch, tile loop, results).
//
// Host behaviour follows the reference's backend (Sources/CuboVirtual/MetalReducers.swift:392-490 and
// ResidentCube.swift:31-128): per-tile passes in scan order, results in Double, refusals never fall back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/cubo_hip.h"
#include "cube.hpp"
#include "errors.hpp"
#include "reduce_kernels.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::g_last_error;
using cubo::hip_fail;
using cubo::element_size;
using cubo::rows_range;
#define HIP_TRY CUBO_HIP_TRY

namespace {

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}


bool valid_shape(const int64_t *s) {
    if (!s) return false;
    __int128 total = 1;
    for (int k = 0; k < 4; ++k) {
        if (s[k] <= 0) return false;
        total *= s[k];
        if (total > ((__int128)1 << 60)) return false;
    }
    return true;
}

}  // namespace


extern "C" {

int cubo_hip_last_error(char *buffer, size_t size) {
    if (!buffer || size == 0) return CUBO_HIP_ERR_INVALID_ARGUMENT;
    snprintf(buffer, size, "%s", g_last_error.c_str());
    return CUBO_HIP_OK;
}

int cubo_hip_device_count(int *count) {
    if (!count) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "count is NULL");
    *count = device_count();
    return CUBO_HIP_OK;
}

int cubo_hip_get_device_info(int device, cubo_hip_device_info *info) {
    if (!info) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "info is NULL");
    const int n = device_count();
    if (n == 0) return fail(CUBO_HIP_ERR_NO_DEVICE, "no HIP device is available");
    if (device < 0 || device >= n) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "device %d of %d", device, n);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    DeviceGuard g(device);
    size_t freeb = 0, total = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &total), "hipMemGetInfo");
    memset(info, 0, sizeof *info);
    info->hbm_bytes = total;
    info->free_bytes = freeb;
    info->max_alloc_bytes = total;
    info->compute_units = prop.multiProcessorCount;
    info->wavefront_size = prop.warpSize;
    snprintf(info->name, sizeof info->name, "%s", prop.name);
    snprintf(info->arch, sizeof info->arch, "%s", prop.gcnArchName);
    return CUBO_HIP_OK;
}

int cubo_hip_availability(int device, const int64_t shape[4], int dtype, uint64_t hbm_bytes_override, int *reason,
                          uint64_t *bytes, uint64_t *cap) {
    if (!reason || !bytes || !cap) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL output");
    if (!valid_shape(shape)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalid cube shape");
    const int es = element_size(dtype);
    if (!es) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    const uint64_t b = (uint64_t)shape[0] * shape[1] * shape[2] * shape[3] * (uint64_t)es;
    *bytes = b;
    const int n = device_count();
    uint64_t hbm = hbm_bytes_override;
    if (hbm == 0 && n > 0 && device >= 0 && device < n) {
        DeviceGuard g(device);
        size_t freeb = 0, total = 0;
        if (hipMemGetInfo(&freeb, &total) == hipSuccess) hbm = total;
    }
    *cap = (uint64_t)((double)hbm * CUBO_HIP_MEMORY_CAP_FRACTION);
    // Order of the reference's checks: the cap, then the device, then the largest buffer.
    if (hbm > 0 && b > *cap) { *reason = CUBO_HIP_UNAVAILABLE_EXCEEDS_MEMORY_CAP; return CUBO_HIP_OK; }
    if (n == 0 || device < 0 || device >= n) { *reason = CUBO_HIP_UNAVAILABLE_NO_DEVICE; return CUBO_HIP_OK; }
    if (hbm == 0) { *reason = CUBO_HIP_UNAVAILABLE_NO_DEVICE; return CUBO_HIP_OK; }
    DeviceGuard g(device);
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) == hipSuccess && hbm_bytes_override == 0 && b > freeb) {
        *reason = CUBO_HIP_UNAVAILABLE_EXCEEDS_BUFFER_LIMIT;
        return CUBO_HIP_OK;
    }
    *reason = CUBO_HIP_AVAILABLE;
    return CUBO_HIP_OK;
}

int cubo_hip_cube_create(int device, const int64_t shape[4], int dtype, cubo_hip_cube **out) {
    if (!out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    const int n = device_count();
    if (n == 0) return fail(CUBO_HIP_ERR_NO_DEVICE, "no HIP device is available; run the CPU path");
    if (device < 0 || device >= n) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "device %d of %d", device, n);
    if (!valid_shape(shape)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalid cube shape");
    const int es = element_size(dtype);
    if (!es) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    int reason = 0;
    uint64_t bytes = 0, cap = 0;
    int st = cubo_hip_availability(device, shape, dtype, 0, &reason, &bytes, &cap);
    if (st) return st;
    if (reason == CUBO_HIP_UNAVAILABLE_EXCEEDS_MEMORY_CAP)
        return fail(CUBO_HIP_ERR_EXCEEDS_CAP, "the cube is %.1f GB; a resident cube may use at most %.1f GB of HBM",
                    bytes / 1e9, cap / 1e9);
    if (reason == CUBO_HIP_UNAVAILABLE_EXCEEDS_BUFFER_LIMIT)
        return fail(CUBO_HIP_ERR_EXCEEDS_BUFFER_LIMIT, "the cube is %.1f GB; that much HBM is not free", bytes / 1e9);
    DeviceGuard g(device);
    auto *c = new (std::nothrow) cubo_hip_cube();
    if (!c) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    c->device = device;
    memcpy(c->shape, shape, sizeof c->shape);
    c->dtype = dtype;
    c->npos = shape[0] * shape[1];
    c->ndet = shape[2] * shape[3];
    c->bytes = bytes;
    hipError_t e = hipMalloc(&c->data, bytes);
    if (e != hipSuccess) { delete c; return hip_fail(e, "allocating the resident cube"); }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { (void)hipFree(c->data); delete c; return hip_fail(e, "hipStreamCreate"); }
    *out = c;
    return CUBO_HIP_OK;
}

int cubo_hip_cube_destroy(cubo_hip_cube *c) {
    if (!c) return CUBO_HIP_OK;
    DeviceGuard g(c->device);
    (void)hipStreamSynchronize(c->stream);
    cubo::free_feed(c);
    c->timer.release();
    if (c->data) (void)hipFree(c->data);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return CUBO_HIP_OK;
}

int cubo_hip_cube_clone(const cubo_hip_cube *src, cubo_hip_cube **out) {
    if (!src || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = cubo::require_resident(src, "cloning")) return st;
    if (int st = cubo_hip_cube_create(src->device, src->shape, src->dtype, out)) return st;
    cubo_hip_cube *d = *out;
    DeviceGuard g(src->device);
    hipError_t e = cubo::launch_copy_bytes(src->data, d->data, src->bytes, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) {
        cubo_hip_cube_destroy(d);
        *out = nullptr;
        return hip_fail(e, "copying the cube");
    }
    return CUBO_HIP_OK;
}

int cubo_hip_cube_byte_count(const cubo_hip_cube *c, uint64_t *bytes) {
    if (!c || !bytes) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *bytes = c->bytes;
    return CUBO_HIP_OK;
}

int cubo_hip_cube_upload_rows(cubo_hip_cube *c, int64_t row0, int64_t nrows, const void *host) {
    if (!c || (!host && nrows > 0)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = cubo::require_resident(c, "uploading rows (a streamed cube is fed)")) return st;
    if (int st = rows_range(c, row0, nrows)) return st;
    DeviceGuard g(c->device);
    ++c->generation;
    const uint64_t row_bytes = (uint64_t)c->shape[1] * c->ndet * element_size(c->dtype);
    HIP_TRY(hipMemcpyAsync((char *)c->data + row0 * row_bytes, host, nrows * row_bytes, hipMemcpyHostToDevice,
                           c->stream), "uploading scan rows");
    HIP_TRY(hipStreamSynchronize(c->stream), "uploading scan rows");
    return CUBO_HIP_OK;
}

int cubo_hip_cube_read_rows(cubo_hip_cube *c, int64_t row0, int64_t nrows, void *host) {
    if (!c || (!host && nrows > 0)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = cubo::require_resident(c, "reading rows")) return st;
    if (int st = rows_range(c, row0, nrows)) return st;
    DeviceGuard g(c->device);
    const uint64_t row_bytes = (uint64_t)c->shape[1] * c->ndet * element_size(c->dtype);
    HIP_TRY(hipMemcpyAsync(host, (const char *)c->data + row0 * row_bytes, nrows * row_bytes, hipMemcpyDeviceToHost,
                           c->stream), "reading scan rows");
    HIP_TRY(hipStreamSynchronize(c->stream), "reading scan rows");
    return CUBO_HIP_OK;
}

int cubo_hip_cube_read_patterns(cubo_hip_cube *c, const int64_t *indices, int64_t n, void *host) {
    if (!c || (n > 0 && (!indices || !host))) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n < 0) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "%lld patterns", (long long)n);
    if (int st = cubo::require_resident(c, "reading patterns")) return st;
    for (int64_t i = 0; i < n; ++i)                      // requirePosition (CubeShape.swift), before any copy
        if (indices[i] < 0 || indices[i] >= c->npos)
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "position %lld outside the scan's %lld positions",
                        (long long)indices[i], (long long)c->npos);
    DeviceGuard g(c->device);
    const uint64_t pb = (uint64_t)c->ndet * element_size(c->dtype);
    for (int64_t i = 0; i < n; ++i)
        HIP_TRY(hipMemcpyAsync((char *)host + i * pb, (const char *)c->data + indices[i] * pb, pb,
                               hipMemcpyDeviceToHost, c->stream), "reading patterns");
    HIP_TRY(hipStreamSynchronize(c->stream), "reading patterns");
    return CUBO_HIP_OK;
}

int cubo_hip_cube_kernel_time(cubo_hip_cube *c, double *total_ms, int64_t *launches) {
    if (!c || !total_ms || !launches) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceGuard g(c->device);
    HIP_TRY(c->timer.total(total_ms, launches), "timing the launches");
    return CUBO_HIP_OK;
}

}  // extern "C"
