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
#define HIP_TRY CUBO_HIP_TRY

namespace {

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int element_size(int dtype) { return dtype == CUBO_HIP_U16 ? 2 : dtype == CUBO_HIP_F32 ? 4 : 0; }

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


struct cubo_hip_reduction {
    cubo_hip_cube *cube;
    int path;               // 1 integer, 2 FP64
    int n_masks;
    bool want_cols, want_com;
    bool com_ones;          // integer path: arithmetic CoM rows
    int nw;                 // weight rows in the kernel (padded)
    int nv;                 // row outputs per position
    int chunks;             // integer kernel chunks per lane
    int slice_px, nslices;
    int64_t rows_done;
    int64_t next_row;
    // device buffers
    void *weights;          // packed uint16 pairs (int) or doubles (f64)
    void *rowpart;
    void *colsum_part;
    void *colmax_part;
    double *rows_out;       // [nv][npos]
    void *col_sum;          // uint64 or double [ndet]
    void *col_max;          // uint32 or float [ndet]
    double *mean_out, *max_out;
    int64_t max_tile_pos;
    int max_groups;
};

namespace {

void free_reduction(cubo_hip_reduction *r) {
    if (!r) return;
    DeviceGuard g(r->cube->device);
    void *bufs[] = {r->weights, r->rowpart, r->colsum_part, r->colmax_part, r->rows_out, r->col_sum, r->col_max,
                    r->mean_out, r->max_out};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete r;
}

// Integral, non-negative and at most `limit`.
bool integral_in(const float *w, int64_t n, double limit, double *maxw) {
    double m = 0;
    for (int64_t k = 0; k < n; ++k) {
        const double v = w[k];
        if (!(v >= 0) || v > limit || v != std::floor(v)) return false;
        m = std::max(m, v);
    }
    *maxw = m;
    return true;
}

int padded_int_nw(int nw) {
    for (int c : {0, 1, 2, 4, 8, 16})
        if (nw <= c) return c;
    return -1;
}

int padded_f64_nw(int nw) {
    for (int c : {1, 2, 3, 4, 5, 6, 8, 11, 16})
        if (nw <= c) return c;
    return -1;
}

// Position groups for one launch: enough blocks to fill the device, at most 65536 positions per group
// (the integer column partials are uint32: 65535 * 65536 < 2^32).
void plan_groups(int64_t npos, int nslices, int cus, int *ngroups, int *ppg, int *ppw) {
    const int64_t target_blocks = (int64_t)cus * 8;
    int64_t g = std::max<int64_t>(1, (target_blocks + nslices - 1) / nslices);
    g = std::min<int64_t>(g, std::max<int64_t>(1, npos / cubo::kWavesPerBlock));
    int64_t per = (npos + g - 1) / g;
    per = std::min<int64_t>(per, 65536);
    per = ((per + cubo::kWavesPerBlock - 1) / cubo::kWavesPerBlock) * cubo::kWavesPerBlock;
    *ppg = (int)per;
    *ngroups = (int)((npos + per - 1) / per);
    *ppw = (int)(per / cubo::kWavesPerBlock);
}

int compute_units(int dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    return cus;
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
    for (auto &ev : c->timed) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
    (void)hipFree(c->data);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return CUBO_HIP_OK;
}

int cubo_hip_cube_byte_count(const cubo_hip_cube *c, uint64_t *bytes) {
    if (!c || !bytes) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *bytes = c->bytes;
    return CUBO_HIP_OK;
}

static int rows_range(const cubo_hip_cube *c, int64_t row0, int64_t nrows) {
    if (row0 < 0 || nrows < 0 || row0 + nrows > c->shape[0])
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "rows [%lld, %lld) outside the scan's %lld rows", (long long)row0,
                    (long long)(row0 + nrows), (long long)c->shape[0]);
    return CUBO_HIP_OK;
}

int cubo_hip_cube_upload_rows(cubo_hip_cube *c, int64_t row0, int64_t nrows, const void *host) {
    if (!c || (!host && nrows > 0)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = rows_range(c, row0, nrows)) return st;
    DeviceGuard g(c->device);
    const uint64_t row_bytes = (uint64_t)c->shape[1] * c->ndet * element_size(c->dtype);
    HIP_TRY(hipMemcpyAsync((char *)c->data + row0 * row_bytes, host, nrows * row_bytes, hipMemcpyHostToDevice,
                           c->stream), "uploading scan rows");
    HIP_TRY(hipStreamSynchronize(c->stream), "uploading scan rows");
    return CUBO_HIP_OK;
}

int cubo_hip_cube_read_rows(cubo_hip_cube *c, int64_t row0, int64_t nrows, void *host) {
    if (!c || (!host && nrows > 0)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = rows_range(c, row0, nrows)) return st;
    DeviceGuard g(c->device);
    const uint64_t row_bytes = (uint64_t)c->shape[1] * c->ndet * element_size(c->dtype);
    HIP_TRY(hipMemcpyAsync(host, (const char *)c->data + row0 * row_bytes, nrows * row_bytes, hipMemcpyDeviceToHost,
                           c->stream), "reading scan rows");
    HIP_TRY(hipStreamSynchronize(c->stream), "reading scan rows");
    return CUBO_HIP_OK;
}

int cubo_hip_cube_kernel_time(cubo_hip_cube *c, double *total_ms, int64_t *launches) {
    if (!c || !total_ms || !launches) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceGuard g(c->device);
    double t = 0;
    for (auto &ev : c->timed) {
        HIP_TRY(hipEventSynchronize(ev.second), "hipEventSynchronize");
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ev.first, ev.second), "hipEventElapsedTime");
        t += ms;
    }
    *total_ms = t;
    *launches = (int64_t)c->timed.size();
    for (auto &ev : c->timed) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
    c->timed.clear();
    return CUBO_HIP_OK;
}

// ---- the fused reduction ----------------------------------------------------------------------------

static int choose_path(const cubo_hip_cube *c, const cubo_hip_reduce_request *q, cubo_hip_reduction *r) {
    const int qy = (int)c->shape[2], qx = (int)c->shape[3];
    const int64_t ndet = c->ndet;
    r->n_masks = q->n_masks;
    r->want_cols = q->want_mean_max != 0;
    r->want_com = q->want_com != 0;
    r->path = 2;
    if (c->dtype != CUBO_HIP_U16 || ndet % 8 != 0) return CUBO_HIP_OK;
    double maxw = 0, m = 0;
    for (int k = 0; k < q->n_masks; ++k) {
        if (!integral_in(q->masks + (size_t)k * ndet, ndet, 65535, &m)) return CUBO_HIP_OK;
        maxw = std::max(maxw, m);
    }
    bool com_ones = false;
    int nw = q->n_masks;
    double comw = 0;
    if (r->want_com) {
        if (!q->com_weights && qx % 8 == 0 && qx <= 4096 && qy <= 4096) {
            com_ones = true;
            comw = (double)std::max(qy - 1, qx - 1);
        } else {
            const float *w = q->com_weights;
            double wm = 1;
            if (w && !integral_in(w, ndet, 65535, &wm)) return CUBO_HIP_OK;
            if (wm * std::max(qy - 1, qx - 1) > 65535) return CUBO_HIP_OK;
            comw = wm * std::max(qy - 1, qx - 1);
            nw += 3;
        }
    }
    const int pnw = padded_int_nw(nw);
    if (pnw < 0) return CUBO_HIP_OK;
    // Largest chunk count whose per-lane uint32 partials cannot overflow: 8*chunks * 65535 * weight.
    const double wbound = std::max({maxw, comw, 1.0});
    for (int ch : {4, 2, 1}) {
        if ((double)ch * 8.0 * 65535.0 * wbound >= 4294967296.0) continue;
        if (!cubo::supported_int(pnw, ch, com_ones)) continue;
        r->path = 1;
        r->com_ones = com_ones;
        r->nw = pnw;
        r->nv = pnw + (com_ones ? 3 : 0);
        r->chunks = ch;
        r->slice_px = cubo::int_slice_pixels(ch);
        return CUBO_HIP_OK;
    }
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_path(cubo_hip_cube *c, const cubo_hip_reduce_request *q, int *path) {
    if (!c || !q || !path) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    cubo_hip_reduction tmp{};
    tmp.cube = c;
    if (int st = choose_path(c, q, &tmp)) return st;
    *path = tmp.path;
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_begin(cubo_hip_cube *c, const cubo_hip_reduce_request *q, cubo_hip_reduction **out) {
    if (!c || !q || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (q->n_masks < 0 || (q->n_masks > 0 && !q->masks))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "masks: %d rows but no data", q->n_masks);
    if (q->n_masks == 0 && !q->want_mean_max && !q->want_com)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "nothing requested");
    const int64_t ndet = c->ndet;
    for (int64_t k = 0; k < (int64_t)q->n_masks * ndet; ++k)
        if (!std::isfinite(q->masks[k])) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "a mask needs finite weights");
    if (q->com_weights)
        for (int64_t k = 0; k < ndet; ++k)
            if (!std::isfinite(q->com_weights[k]))
                return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the centre-of-mass weights must be finite");

    auto *r = new (std::nothrow) cubo_hip_reduction();
    if (!r) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    r->cube = c;
    if (int st = choose_path(c, q, r)) { delete r; return st; }
    const int qy = (int)c->shape[2], qx = (int)c->shape[3];
    if (r->path == 2) {
        const int nw = q->n_masks + (r->want_com ? 3 : 0);
        const int pnw = padded_f64_nw(std::max(nw, 1));
        if (pnw < 0) {
            delete r;
            return fail(CUBO_HIP_ERR_UNSUPPORTED, "%d weight rows in one pass (at most 16); split the masks", nw);
        }
        r->nw = pnw;
        r->nv = pnw;
        r->slice_px = cubo::kF64Slice;
    } else if (r->nv == 0) {
        r->nv = 0;
    }
    r->nslices = (int)((ndet + r->slice_px - 1) / r->slice_px);

    DeviceGuard g(c->device);
    // Weight rows on the device.
    hipError_t e = hipSuccess;
    if (r->path == 1) {
        const int64_t pairs = ndet / 2;
        std::vector<uint32_t> pk((size_t)std::max(r->nw, 1) * pairs, 0u);
        auto put = [&](int row, int64_t px, uint32_t v) {
            uint32_t &word = pk[(size_t)row * pairs + px / 2];
            word |= (v & 0xffffu) << (16 * (px & 1));
        };
        for (int k = 0; k < q->n_masks; ++k)
            for (int64_t px = 0; px < ndet; ++px) put(k, px, (uint32_t)q->masks[(size_t)k * ndet + px]);
        if (r->want_com && !r->com_ones) {
            for (int64_t px = 0; px < ndet; ++px) {
                const uint32_t w = q->com_weights ? (uint32_t)q->com_weights[px] : 1u;
                put(q->n_masks, px, w);
                put(q->n_masks + 1, px, w * (uint32_t)(px / qx));
                put(q->n_masks + 2, px, w * (uint32_t)(px % qx));
            }
        }
        e = hipMalloc(&r->weights, pk.size() * 4);
        if (e == hipSuccess) e = hipMemcpy(r->weights, pk.data(), pk.size() * 4, hipMemcpyHostToDevice);
    } else {
        std::vector<double> wd((size_t)r->nw * ndet, 0.0);
        for (int k = 0; k < q->n_masks; ++k)
            for (int64_t px = 0; px < ndet; ++px) wd[(size_t)k * ndet + px] = (double)q->masks[(size_t)k * ndet + px];
        if (r->want_com) {
            for (int64_t px = 0; px < ndet; ++px) {
                const double w = q->com_weights ? (double)q->com_weights[px] : 1.0;
                wd[(size_t)q->n_masks * ndet + px] = w;                             // w, w*i, w*j in Double
                wd[(size_t)(q->n_masks + 1) * ndet + px] = w * (double)(px / qx);   // (CenterOfMass.swift:58-59)
                wd[(size_t)(q->n_masks + 2) * ndet + px] = w * (double)(px % qx);
            }
        }
        e = hipMalloc(&r->weights, wd.size() * 8);
        if (e == hipSuccess) e = hipMemcpy(r->weights, wd.data(), wd.size() * 8, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) { free_reduction(r); return hip_fail(e, "uploading the weight rows"); }

    // Scratch for the largest tile (the whole cube).
    const int cus = compute_units(c->device);
    int ngroups = 0, ppg = 0, ppw = 0;
    plan_groups(c->npos, r->nslices, cus, &ngroups, &ppg, &ppw);
    r->max_tile_pos = c->npos;
    r->max_groups = ngroups;
    const size_t acc = 8;
    const size_t rowpart_bytes = (size_t)r->nslices * c->npos * std::max(r->nv, 1) * acc;
    const size_t colsum_bytes = (size_t)ngroups * ndet * (r->path == 1 ? 4 : 8);
    const size_t colmax_bytes = (size_t)ngroups * ndet * 4;
    struct { void **p; size_t n; } allocs[] = {
        {&r->rowpart, rowpart_bytes}, {&r->colsum_part, colsum_bytes}, {&r->colmax_part, colmax_bytes},
        {(void **)&r->rows_out, (size_t)std::max(r->nv, 1) * c->npos * 8}, {&r->col_sum, (size_t)ndet * 8},
        {&r->col_max, (size_t)ndet * 4}, {(void **)&r->mean_out, (size_t)ndet * 8},
        {(void **)&r->max_out, (size_t)ndet * 8}};
    for (auto &a : allocs) {
        e = hipMalloc(a.p, std::max<size_t>(a.n, 8));
        if (e != hipSuccess) { free_reduction(r); return hip_fail(e, "allocating reduction scratch"); }
    }
    if (r->path == 1)
        e = cubo::launch_fill_cols_init((uint64_t *)r->col_sum, (uint32_t *)r->col_max, nullptr, nullptr, (int)ndet,
                                        c->stream);
    else
        e = cubo::launch_fill_cols_init(nullptr, nullptr, (double *)r->col_sum, (float *)r->col_max, (int)ndet,
                                        c->stream);
    if (e != hipSuccess) { free_reduction(r); return hip_fail(e, "initialising the column sums"); }
    *out = r;
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_rows(cubo_hip_reduction *r, int64_t row0, int64_t nrows) {
    if (!r) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL reduction");
    cubo_hip_cube *c = r->cube;
    if (int st = rows_range(c, row0, nrows)) return st;
    if (row0 != r->next_row)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "tiles must come in scan order: expected row %lld, got %lld",
                    (long long)r->next_row, (long long)row0);
    if (nrows == 0) return CUBO_HIP_OK;
    DeviceGuard g(c->device);
    const int64_t pos0 = row0 * c->shape[1];
    const int64_t npos = nrows * c->shape[1];
    const int ndet = (int)c->ndet;
    int ngroups = 0, ppg = 0, ppw = 0;
    plan_groups(npos, r->nslices, compute_units(c->device), &ngroups, &ppg, &ppw);
    if (ngroups > r->max_groups) return fail(CUBO_HIP_ERR_HIP, "internal: group plan exceeds scratch");
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&t1), "hipEventCreate");
    HIP_TRY(hipEventRecord(t0, c->stream), "hipEventRecord");
    hipError_t e;
    if (r->path == 1) {
        cubo::IntLaunch a{};
        a.cube = (const uint16_t *)c->data + pos0 * c->ndet;
        a.npos = npos;
        a.ndet = ndet;
        a.qx = (int)c->shape[3];
        a.nslices = r->nslices;
        a.ngroups = ngroups;
        a.ppg = ppg;
        a.ppw = ppw;
        a.wpk = (const uint32_t *)r->weights;
        a.rowpart = (uint64_t *)r->rowpart;
        a.colsum = (uint32_t *)r->colsum_part;
        a.colmax = (uint32_t *)r->colmax_part;
        e = cubo::launch_int(a, r->nw, r->chunks, r->com_ones, c->stream);
    } else {
        cubo::F64Launch a{};
        a.cube = (const char *)c->data + (size_t)pos0 * c->ndet * element_size(c->dtype);
        a.npos = npos;
        a.ndet = ndet;
        a.nslices = r->nslices;
        a.ngroups = ngroups;
        a.ppg = ppg;
        a.ppw = ppw;
        a.w = (const double *)r->weights;
        a.rowpart = (double *)r->rowpart;
        a.colsum = (double *)r->colsum_part;
        a.colmax = (float *)r->colmax_part;
        e = cubo::launch_f64(a, c->dtype == CUBO_HIP_U16, r->nw, c->stream);
    }
    if (e != hipSuccess) return hip_fail(e, "launching the fused reduction");
    HIP_TRY(hipEventRecord(t1, c->stream), "hipEventRecord");
    c->timed.emplace_back(t0, t1);
    if (r->nv > 0) {
        e = r->path == 1 ? cubo::launch_finalize_rows_u64((const uint64_t *)r->rowpart, r->nslices, npos, r->nv,
                                                          r->rows_out, pos0, c->npos, c->stream)
                         : cubo::launch_finalize_rows_f64((const double *)r->rowpart, r->nslices, npos, r->nv,
                                                          r->rows_out, pos0, c->npos, c->stream);
        if (e != hipSuccess) return hip_fail(e, "launching the row totals");
    }
    e = r->path == 1 ? cubo::launch_accumulate_cols_int((const uint32_t *)r->colsum_part,
                                                        (const uint32_t *)r->colmax_part, ngroups, ndet,
                                                        (uint64_t *)r->col_sum, (uint32_t *)r->col_max, c->stream)
                     : cubo::launch_accumulate_cols_f64((const double *)r->colsum_part, (const float *)r->colmax_part,
                                                        ngroups, ndet, (double *)r->col_sum, (float *)r->col_max,
                                                        c->stream);
    if (e != hipSuccess) return hip_fail(e, "launching the column totals");
    r->next_row = row0 + nrows;
    r->rows_done += nrows;
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_cancel(cubo_hip_reduction *r) {
    if (!r) return CUBO_HIP_OK;
    DeviceGuard g(r->cube->device);
    (void)hipStreamSynchronize(r->cube->stream);
    free_reduction(r);
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_results(cubo_hip_reduction *r, cubo_hip_reduce_result *res) {
    if (!r || !res) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    cubo_hip_cube *c = r->cube;
    if (r->rows_done != c->shape[0]) {
        int st = fail(CUBO_HIP_ERR_INCOMPLETE, "the reduction covered %lld of %lld scan rows", (long long)r->rows_done,
                      (long long)c->shape[0]);
        return st;
    }
    DeviceGuard g(c->device);
    const int64_t npos = c->npos, ndet = c->ndet;
    int status = CUBO_HIP_OK;
    auto copy = [&](double *dst, const double *src, int64_t n) {
        if (!dst || status) return;
        hipError_t e = hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) status = hip_fail(e, "reading results");
    };
    if (r->want_cols) {
        hipError_t e = r->path == 1
            ? cubo::launch_finish_cols_int((const uint64_t *)r->col_sum, (const uint32_t *)r->col_max, (int)ndet,
                                           (double)npos, r->mean_out, r->max_out, c->stream)
            : cubo::launch_finish_cols_f64((const double *)r->col_sum, (const float *)r->col_max, (int)ndet,
                                           (double)npos, r->mean_out, r->max_out, c->stream);
        if (e != hipSuccess) status = hip_fail(e, "launching the column finish");
        copy(res->mean, r->mean_out, ndet);
        copy(res->max, r->max_out, ndet);
    }
    for (int k = 0; k < r->n_masks && res->images; ++k) copy(res->images + (size_t)k * npos, r->rows_out + (size_t)k * npos, npos);
    std::vector<double> s, mr, mc;
    if (r->want_com) {
        // The integer kernel's arithmetic CoM rows follow its weight rows; in every other layout the three
        // CoM rows directly follow the masks.
        const int base = (r->path == 1 && r->com_ones) ? r->nw : r->n_masks;
        s.resize(npos); mr.resize(npos); mc.resize(npos);
        copy(s.data(), r->rows_out + (size_t)base * npos, npos);
        copy(mr.data(), r->rows_out + (size_t)(base + 1) * npos, npos);
        copy(mc.data(), r->rows_out + (size_t)(base + 2) * npos, npos);
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess && !status) status = hip_fail(e, "waiting for the reduction");
    if (!status && r->want_com) {
        for (int64_t p = 0; p < npos; ++p) {
            if (res->com_intensity) res->com_intensity[p] = s[p];
            if (res->com_moment_row) res->com_moment_row[p] = mr[p];
            if (res->com_moment_column) res->com_moment_column[p] = mc[p];
            if (res->com_row) res->com_row[p] = mr[p] / s[p];       // NaN where S = 0 (CenterOfMass.swift:44-45)
            if (res->com_column) res->com_column[p] = mc[p] / s[p];
        }
    }
    return status;
}

int cubo_hip_reduce_reset(cubo_hip_reduction *r) {
    if (!r) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL reduction");
    cubo_hip_cube *c = r->cube;
    DeviceGuard g(c->device);
    r->rows_done = 0;
    r->next_row = 0;
    hipError_t e = r->path == 1
        ? cubo::launch_fill_cols_init((uint64_t *)r->col_sum, (uint32_t *)r->col_max, nullptr, nullptr, (int)c->ndet,
                                      c->stream)
        : cubo::launch_fill_cols_init(nullptr, nullptr, (double *)r->col_sum, (float *)r->col_max, (int)c->ndet,
                                      c->stream);
    if (e != hipSuccess) return hip_fail(e, "initialising the column sums");
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_release(cubo_hip_reduction *r) { return cubo_hip_reduce_cancel(r); }

int cubo_hip_reduce_end(cubo_hip_reduction *r, cubo_hip_reduce_result *res) {
    if (!r) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL reduction");
    const int st = cubo_hip_reduce_results(r, res);
    cubo_hip_reduce_cancel(r);
    return st;
}



int cubo_hip_reduce(cubo_hip_cube *c, const cubo_hip_reduce_request *q, cubo_hip_reduce_result *res) {
    cubo_hip_reduction *r = nullptr;
    if (int st = cubo_hip_reduce_begin(c, q, &r)) return st;
    if (int st = cubo_hip_reduce_rows(r, 0, c->shape[0])) { cubo_hip_reduce_cancel(r); return st; }
    return cubo_hip_reduce_end(r, res);
}

}  // extern "C"
