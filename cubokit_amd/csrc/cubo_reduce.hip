// cubo_reduce.hip — host side of the fused reduction (include/cubo_hip.h, "the fused reduction").
//
// Chooses the kernel family, prepares the weight rows, plans position groups from the kernel's occupancy,
// feeds scan-row tiles in order (Sources/CuboVirtual/MetalReducers.swift:439-490: one pass per tile, the
// caller checks cancellation between tiles) and returns the Double results. All results of a pass live in
// one device buffer that comes back in one copy into pinned host memory; cubo_hip_reduce_view exposes it
// without a further copy, cubo_hip_reduce_results copies it into the caller's buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/cubo_hip.h"
#include "cube.hpp"
#include "errors.hpp"
#include "reduce_kernels.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::hip_fail;

struct cubo_hip_reduction {
    cubo_hip_cube *cube;
    int path;               // 1 exact integer, 2 FP64
    int n_masks;
    bool want_cols, want_com;
    bool com_ones;          // integer path: arithmetic CoM rows
    int com_kind;           // FP64 path: 0 no CoM, 1 unit weights, 2 weighted
    bool f64_vec;           // FP64 path: the vectorised kernel where it applies (CUBO_HIP_F64_KERNEL=scalar: never)
    int nw;                 // weight rows in the kernel (padded)
    int nv;                 // row outputs per position
    int chunks;             // integer kernel: 8-pixel chunks per lane
    int slice_px, nslices;
    int blocks_per_cu;
    int64_t rows_done, next_row;
    bool viewed;
    // device
    void *weights;          // packed uint16 pairs (integer) or Double rows (FP64)
    void *rowpart;          // [nslices][npos][nv]
    void *colsum_part;      // [groups][ndet] uint32 (integer) or double (FP64)
    void *colmax_part;      // [groups][ndet] packed uint16 pairs (integer) or float (FP64)
    void *col_sum;          // uint64 or double [ndet], running over tiles
    void *col_max;          // uint32 or float [ndet]
    bool full;              // integer kernel in the FULL layout (one block per group, rows written in-kernel)
    int max_groups;
    // Mapped pinned host memory, written by the kernels: rows [nv][npos], mean [ndet], max [ndet], then com
    // row [npos], com column [npos]. res_hd is its device address.
    double *res_host;
    double *res_hd;
    size_t res_count;       // doubles before the com rows
};

namespace {

int64_t result_doubles(const cubo_hip_reduction *r) {
    return (int64_t)std::max(r->nv, 1) * r->cube->npos + 2 * r->cube->ndet;
}

void free_reduction(cubo_hip_reduction *r) {
    if (!r) return;
    DeviceGuard g(r->cube->device);
    void *bufs[] = {r->weights, r->rowpart, r->colsum_part, r->colmax_part, r->col_sum, r->col_max};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (r->res_host) (void)hipHostFree(r->res_host);
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

// Position groups of one launch: one wave of blocks that fills every compute unit at the kernel's
// occupancy, at most 65536 positions per group (the integer column partials are uint32 before the
// 64-bit add: 65535 * 65536 < 2^32).
// FULL layout: one block per group, every wave walks the whole group.
void plan_groups_full(int64_t npos, int cus, int blocks_per_cu, int *ngroups, int *ppg, int *ppw) {
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * std::max(1, blocks_per_cu), npos));
    const int64_t per = std::min<int64_t>((npos + g - 1) / g, 65536);
    *ppg = (int)per;
    *ngroups = (int)((npos + per - 1) / per);
    *ppw = (int)per;
}

void plan_groups(int64_t npos, int nslices, int cus, int blocks_per_cu, int *ngroups, int *ppg, int *ppw) {
    const int64_t target_blocks = (int64_t)cus * std::max(1, blocks_per_cu);
    int64_t g = std::max<int64_t>(1, (target_blocks + nslices - 1) / nslices);
    g = std::min<int64_t>(g, std::max<int64_t>(1, npos / cubo::kWavesPerBlock));
    int64_t per = (npos + g - 1) / g;
    per = std::min<int64_t>(per, 65536);
    per = ((per + cubo::kWavesPerBlock - 1) / cubo::kWavesPerBlock) * cubo::kWavesPerBlock;
    *ppg = (int)per;
    *ngroups = (int)((npos + per - 1) / per);
    *ppw = (int)(per / cubo::kWavesPerBlock);
}

int choose_path(const cubo_hip_cube *c, const cubo_hip_reduce_request *q, cubo_hip_reduction *r) {
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
            double wm = 1;
            if (q->com_weights && !integral_in(q->com_weights, ndet, 65535, &wm)) return CUBO_HIP_OK;
            if (wm * std::max(qy - 1, qx - 1) > 65535) return CUBO_HIP_OK;
            comw = wm * std::max(qy - 1, qx - 1);
            nw += 3;
        }
    }
    const int pnw = padded_int_nw(nw);
    if (pnw < 0) return CUBO_HIP_OK;
    // Largest chunk count whose per-lane uint32 partials cannot overflow: 8*chunks * 65535 * weight.
    const double wbound = std::max({maxw, comw, 1.0});
    // CUBO_HIP_MAX_CHUNKS caps the chunk count (a tuning knob: fewer chunks, fewer registers per lane).
    int max_chunks = 4;
    if (const char *env = getenv("CUBO_HIP_MAX_CHUNKS")) max_chunks = std::max(1, atoi(env));
    for (int ch : {4, 2, 1}) {
        if (ch > max_chunks) continue;
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


void plan(const cubo_hip_reduction *r, int64_t npos, int *ngroups, int *ppg, int *ppw) {
    const int cus = cubo::compute_units(r->cube->device);
    if (r->full) plan_groups_full(npos, cus, r->blocks_per_cu, ngroups, ppg, ppw);
    else plan_groups(npos, r->nslices, cus, r->blocks_per_cu, ngroups, ppg, ppw);
}

// Where a launch starting at position pos0 writes its row totals and centre of mass (mapped host memory).
cubo::RowsOut rows_out(const cubo_hip_reduction *r, int64_t pos0) {
    cubo::RowsOut o{};
    o.rows = r->res_hd;
    o.com = r->res_hd + r->res_count;
    o.pos0 = pos0;
    o.total_pos = r->cube->npos;
    o.nv = r->nv;
    // The integer kernel's arithmetic CoM rows follow its weight rows; in every other layout the three CoM
    // rows directly follow the masks.
    o.com_base = !r->want_com ? -1 : (r->path == 1 && r->com_ones) ? r->nw : r->n_masks;
    return o;
}

int check_request(const cubo_hip_cube *c, const cubo_hip_reduce_request *q) {
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
    return CUBO_HIP_OK;
}

// The FP64 launch fields that select its kernel (the vectorised one or the one-pixel-per-lane one).
cubo::F64Launch f64_probe(const cubo_hip_reduction *r) {
    cubo::F64Launch a{};
    a.cube = r->cube->data;
    a.ndet = (int32_t)r->cube->ndet;
    a.nmask = r->n_masks;
    a.com = r->com_kind;
    a.qx = (int32_t)r->cube->shape[3];
    a.vec_ok = r->f64_vec;
    return a;
}

// The kernel family, weight rows, slices, block layout and the position groups of a whole-cube tile (sets
// max_groups and res_count). Allocates nothing.
int configure(cubo_hip_cube *c, const cubo_hip_reduce_request *q, cubo_hip_reduction *r) {
    r->cube = c;
    if (int st = choose_path(c, q, r)) return st;
    const int64_t ndet = c->ndet;
    DeviceGuard g(c->device);
    if (r->path == 2) {
        const int nw = q->n_masks + (r->want_com ? 3 : 0);
        const int pnw = padded_f64_nw(std::max(nw, 1));
        if (pnw < 0)
            return fail(CUBO_HIP_ERR_UNSUPPORTED, "%d weight rows in one pass (at most 16); split the masks", nw);
        r->nw = pnw;
        r->nv = pnw;
        r->com_kind = !r->want_com ? 0 : q->com_weights ? 2 : 1;
        // CUBO_HIP_F64_KERNEL=scalar keeps the one-pixel-per-lane kernel (a test knob)
        const char *k = getenv("CUBO_HIP_F64_KERNEL");
        r->f64_vec = !(k && strcmp(k, "scalar") == 0);
        r->slice_px = cubo::f64_slice_pixels(f64_probe(r), c->dtype == CUBO_HIP_U16);
        // the vectorised kernel with one octet per pattern writes the row totals itself (one block per group)
        r->full = cubo::f64_kernel(f64_probe(r), c->dtype == CUBO_HIP_U16) == 2 && ndet <= r->slice_px;
        r->blocks_per_cu = cubo::blocks_per_cu_f64(f64_probe(r), c->dtype == CUBO_HIP_U16, pnw);
    }
    r->nslices = (int)((ndet + r->slice_px - 1) / r->slice_px);
    if (r->path == 1) {
        // CUBO_HIP_INT_LAYOUT=split forces the (slice, group) layout (a tuning and test knob).
        const char *layout = getenv("CUBO_HIP_INT_LAYOUT");
        r->full = r->nslices <= cubo::kFullSlices && !(layout && strcmp(layout, "split") == 0) &&
                  cubo::full_without_spills(r->nw, r->chunks, r->com_ones);
        r->blocks_per_cu = cubo::blocks_per_cu_int(r->nw, r->chunks, r->com_ones, r->full, r->nslices);
    }
    // Blocks per compute unit of the launch plan: the kernel's occupancy, or a tuned multiple of it.
    if (const char *env = getenv("CUBO_HIP_BLOCKS_PER_CU")) r->blocks_per_cu = std::max(1, atoi(env));
    int ngroups = 0, ppg = 0, ppw = 0;
    plan(r, c->npos, &ngroups, &ppg, &ppw);
    r->max_groups = ngroups;
    r->res_count = (size_t)result_doubles(r);
    return CUBO_HIP_OK;
}

// Bytes of the device scratch of a configured reduction: row partials (split layouts), column partials (sums,
// maxima), running column sums and maxima.
struct ScratchSizes {
    size_t n[5];
};
ScratchSizes scratch_sizes(const cubo_hip_reduction *r) {
    const cubo_hip_cube *c = r->cube;
    const size_t ndet = (size_t)c->ndet, ng = (size_t)r->max_groups;
    // (integer path: the group maxima as packed uint16 pairs)
    return ScratchSizes{{r->full ? 0 : (size_t)r->nslices * c->npos * std::max(r->nv, 1) * 8,
                         ng * ndet * (r->path == 2 ? 8 : 4), ng * ndet * (r->path == 2 ? 4 : 2), ndet * 8, ndet * 4}};
}

// One tile's launches: the fused kernel, the row totals of the split layouts, the column totals. timed: bracket
// the fused kernel with the cube's launch timer.
int launch_tile(cubo_hip_reduction *r, int64_t row0, int64_t nrows, const void *tile, bool timed) {
    cubo_hip_cube *c = r->cube;
    const int64_t pos0 = row0 * c->shape[1];
    const int64_t npos = nrows * c->shape[1];
    const int ndet = (int)c->ndet;
    int ngroups = 0, ppg = 0, ppw = 0;
    plan(r, npos, &ngroups, &ppg, &ppw);
    if (ngroups > r->max_groups) return fail(CUBO_HIP_ERR_HIP, "internal: group plan exceeds scratch");
    if (timed) CUBO_HIP_TRY(c->timer.start(c->stream), "hipEventRecord");
    hipError_t e;
    if (r->path == 1) {
        cubo::IntLaunch a{};
        a.cube = (const uint16_t *)tile;
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
        a.out = rows_out(r, pos0);
        e = cubo::launch_int(a, r->nw, r->chunks, r->com_ones, r->full, c->stream);
    } else {
        cubo::F64Launch a{};
        a.cube = tile;
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
        a.nmask = r->n_masks;
        a.com = r->com_kind;
        a.qx = (int)c->shape[3];
        a.vec_ok = r->f64_vec;
        a.out = rows_out(r, pos0);
        e = cubo::launch_f64(a, c->dtype == CUBO_HIP_U16, r->nw, c->stream);
    }
    if (e != hipSuccess) return hip_fail(e, "launching the fused reduction");
    if (timed) CUBO_HIP_TRY(c->timer.stop(c->stream), "hipEventRecord");
    // The integer kernel in the FULL layout writes its row totals itself.
    if (r->nv > 0 && !r->full) {
        e = r->path == 1 ? cubo::launch_finalize_rows_u64((const uint64_t *)r->rowpart, r->nslices, npos,
                                                          rows_out(r, pos0), c->stream)
                         : cubo::launch_finalize_rows_f64((const double *)r->rowpart, r->nslices, npos,
                                                          rows_out(r, pos0), c->stream);
        if (e != hipSuccess) return hip_fail(e, "launching the row totals");
    }
    if (r->want_cols) {
        const bool first = r->rows_done == 0, finish = row0 + nrows == c->shape[0];
        double *mean = r->res_hd + (size_t)std::max(r->nv, 1) * c->npos, *max = mean + c->ndet;
        e = r->path == 1 ? cubo::launch_accumulate_cols_int((const uint32_t *)r->colsum_part,
                                                            (const uint32_t *)r->colmax_part, ngroups, ndet,
                                                            (uint64_t *)r->col_sum, (uint32_t *)r->col_max, first,
                                                            finish, (double)c->npos, mean, max, c->stream)
                         : cubo::launch_accumulate_cols_f64((const double *)r->colsum_part,
                                                            (const float *)r->colmax_part, ngroups, ndet,
                                                            (double *)r->col_sum, (float *)r->col_max, first, finish,
                                                            (double)c->npos, mean, max, c->stream);
        if (e != hipSuccess) return hip_fail(e, "launching the column totals");
    }
    return CUBO_HIP_OK;
}

// ---- the reduction without device allocation (cubo_hip_reduce_noalloc) ----
//
// Its weight rows, column and row partials, running column totals and results live in `scratch`, one
// file-scope device array per device (held until the process ends). Host code never passes `scratch` itself
// (its name is a host-side stand-in): the first call on a device looks up the array's device address with
// hipGetSymbolAddress, with that device current, and keeps it. Calls on one device take turns under a mutex and
// wait for their stream before returning, so no two of them use the array at once.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocDevice {
    std::mutex mu;
    void *ptr = nullptr;    // the device address of `scratch` on this device
};
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

// The weight value of (row, px) of a request: the masks, then (com_rows) w, w*i, w*j, then zero rows.
template <typename T>
__device__ __forceinline__ T weight_at(const float *masks, const float *cw, int n_masks, int64_t ndet, int qx,
                                       int com_rows, int row, int64_t px) {
    if (row < n_masks) return (T)masks[(int64_t)row * ndet + px];
    const int j = row - n_masks;
    if (!com_rows || j > 2) return (T)0;
    const T w = cw ? (T)cw[px] : (T)1;
    return j == 0 ? w : j == 1 ? w * (T)(px / qx) : w * (T)(px % qx);
}

// Integer kernel rows: uint16 pairs, the lower pixel in the low half (as the host packing of
// cubo_hip_reduce_begin); rows x (ndet / 2) words.
__global__ void pack_weights_int(const float *masks, const float *cw, int n_masks, int64_t ndet, int qx, int com_rows,
                                 int rows, uint32_t *wpk) {
    const int64_t pairs = ndet / 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * pairs) return;
    const int row = (int)(i / pairs);
    const int64_t px = 2 * (i % pairs);
    const uint32_t lo = weight_at<uint32_t>(masks, cw, n_masks, ndet, qx, com_rows, row, px);
    const uint32_t hi = weight_at<uint32_t>(masks, cw, n_masks, ndet, qx, com_rows, row, px + 1);
    wpk[i] = (lo & 0xffffu) | ((hi & 0xffffu) << 16);
}

// FP64 kernel rows in Double: rows x ndet.
__global__ void pack_weights_f64(const float *masks, const float *cw, int n_masks, int64_t ndet, int qx, int com_rows,
                                 int rows, double *w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * ndet) return;
    w[i] = weight_at<double>(masks, cw, n_masks, ndet, qx, com_rows, (int)(i / ndet), i % ndet);
}

}  // namespace

extern "C" {

int cubo_hip_reduce_noalloc(cubo_hip_cube *c, const cubo_hip_reduce_request *q, cubo_hip_reduce_result *res) {
    if (!c || !q || !res) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = cubo::require_resident(c, "cubo_hip_reduce_noalloc")) return st;
    if (c->device < 0 || c->device >= kNoAllocMaxDevices)
        return fail(CUBO_HIP_ERR_UNSUPPORTED, "device %d (the allocation-free reduction serves devices 0..%d)",
                    c->device, kNoAllocMaxDevices - 1);
    if (int st = check_request(c, q)) return st;
    cubo_hip_reduction r{};
    if (int st = configure(c, q, &r)) return st;
    const int64_t npos = c->npos, ndet = c->ndet;
    // The scratch layout, every piece 256-byte aligned: masks and CoM weights as given (Float32), the kernel's
    // weight rows, the partials, the results [rows][npos], mean, max, com row, com column.
    const ScratchSizes z = scratch_sizes(&r);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const int wrows = r.path == 1 ? std::max(r.nw, 1) : r.nw;
    const size_t o_masks = take((size_t)q->n_masks * ndet * 4);
    const size_t o_cw = take(q->com_weights ? (size_t)ndet * 4 : 0);
    const size_t o_w = take(r.path == 1 ? (size_t)wrows * (ndet / 2) * 4 : (size_t)wrows * ndet * 8);
    size_t o_part[5];
    for (int i = 0; i < 5; ++i) o_part[i] = take(z.n[i]);
    const size_t o_res = take((r.res_count + 2 * (size_t)npos) * 8);
    if (off > kNoAllocScratchBytes)
        return fail(CUBO_HIP_ERR_EXCEEDS_BUFFER_LIMIT,
                    "this request needs %.1f MiB of device scratch; the allocation-free reduction holds %zu MiB "
                    "(use cubo_hip_reduce)", off / 1048576.0, kNoAllocScratchBytes >> 20);
    DeviceGuard g(c->device);
    NoAllocDevice &d = g_noalloc[c->device];
    std::lock_guard<std::mutex> lock(d.mu);
    if (!d.ptr) CUBO_HIP_TRY(hipGetSymbolAddress(&d.ptr, HIP_SYMBOL(scratch)), "locating the device scratch");
    char *base = (char *)d.ptr;
    hipStream_t s = c->stream;
    const float *dmasks = q->n_masks > 0 ? (const float *)(base + o_masks) : nullptr;
    const float *dcw = q->com_weights ? (const float *)(base + o_cw) : nullptr;
    if (q->n_masks > 0)
        CUBO_HIP_TRY(hipMemcpyAsync(base + o_masks, q->masks, (size_t)q->n_masks * ndet * 4, hipMemcpyHostToDevice, s),
                     "uploading the masks");
    if (dcw) CUBO_HIP_TRY(hipMemcpyAsync(base + o_cw, q->com_weights, (size_t)ndet * 4, hipMemcpyHostToDevice, s),
                          "uploading the centre-of-mass weights");
    const int qx = (int)c->shape[3];
    if (r.path == 1) {
        const int64_t n = (int64_t)wrows * (ndet / 2);
        hipLaunchKernelGGL(pack_weights_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dmasks, dcw,
                           q->n_masks, ndet, qx, (int)(r.want_com && !r.com_ones), wrows, (uint32_t *)(base + o_w));
    } else {
        const int64_t n = (int64_t)wrows * ndet;
        hipLaunchKernelGGL(pack_weights_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dmasks, dcw,
                           q->n_masks, ndet, qx, (int)r.want_com, wrows, (double *)(base + o_w));
    }
    CUBO_HIP_TRY(hipGetLastError(), "packing the weight rows");
    r.weights = base + o_w;
    void **const bufs[] = {&r.rowpart, &r.colsum_part, &r.colmax_part, &r.col_sum, &r.col_max};
    for (int i = 0; i < 5; ++i) *bufs[i] = z.n[i] ? base + o_part[i] : nullptr;
    r.res_hd = (double *)(base + o_res);
    if (int st = launch_tile(&r, 0, c->shape[0], c->data, false)) return st;
    // The results into the caller's buffers, in the layout cubo_hip_reduce_view describes.
    const double *rows = r.res_hd;
    const double *mean = rows + (size_t)std::max(r.nv, 1) * npos, *max = mean + ndet;
    const double *crow = r.res_hd + r.res_count, *ccol = crow + npos;
    const int cb = (r.path == 1 && r.com_ones) ? r.nw : r.n_masks;
    struct { double *dst; const double *src; int64_t n; } outs[] = {
        {res->images, rows, (int64_t)r.n_masks * npos},
        {r.want_cols ? res->mean : nullptr, mean, ndet},
        {r.want_cols ? res->max : nullptr, max, ndet},
        {r.want_com ? res->com_intensity : nullptr, rows + (size_t)cb * npos, npos},
        {r.want_com ? res->com_moment_row : nullptr, rows + (size_t)(cb + 1) * npos, npos},
        {r.want_com ? res->com_moment_column : nullptr, rows + (size_t)(cb + 2) * npos, npos},
        {r.want_com ? res->com_row : nullptr, crow, npos},
        {r.want_com ? res->com_column : nullptr, ccol, npos}};
    for (auto &o : outs)
        if (o.dst && o.n > 0)
            CUBO_HIP_TRY(hipMemcpyAsync(o.dst, o.src, (size_t)o.n * 8, hipMemcpyDeviceToHost, s), "reading the results");
    CUBO_HIP_TRY(hipStreamSynchronize(s), "waiting for the reduction");
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
    if (int st = check_request(c, q)) return st;
    auto *r = new (std::nothrow) cubo_hip_reduction();
    if (!r) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    if (int st = configure(c, q, r)) { delete r; return st; }
    const int64_t ndet = c->ndet;
    const int qx = (int)c->shape[3];
    DeviceGuard g(c->device);
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
    const ScratchSizes z = scratch_sizes(r);
    void **const bufs[] = {&r->rowpart, &r->colsum_part, &r->colmax_part, &r->col_sum, &r->col_max};
    for (int i = 0; i < 5; ++i) {
        if (z.n[i] == 0) continue;
        e = hipMalloc(bufs[i], z.n[i]);
        if (e != hipSuccess) { free_reduction(r); return hip_fail(e, "allocating reduction scratch"); }
    }
    // Results go straight from the kernels to mapped, coherent pinned memory: no copy after the pass.
    e = hipHostMalloc((void **)&r->res_host, (r->res_count + 2 * (size_t)c->npos) * 8,
                      hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) { free_reduction(r); return hip_fail(e, "allocating pinned result memory"); }
    e = hipHostGetDevicePointer((void **)&r->res_hd, r->res_host, 0);
    if (e != hipSuccess) { free_reduction(r); return hip_fail(e, "mapping the result memory"); }
    *out = r;
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_rows(cubo_hip_reduction *r, int64_t row0, int64_t nrows) {
    if (!r) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL reduction");
    cubo_hip_cube *c = r->cube;
    if (int st = cubo::rows_range(c, row0, nrows)) return st;
    if (row0 != r->next_row)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "tiles must come in scan order: expected row %lld, got %lld",
                    (long long)r->next_row, (long long)row0);
    if (nrows == 0) return CUBO_HIP_OK;
    DeviceGuard g(c->device);
    const void *tile = nullptr;   // the tile's first element: in the resident buffer or a streamed cube's ring
    if (int st = cubo::rows_data(c, row0, nrows, &tile)) return st;
    if (int st = launch_tile(r, row0, nrows, tile, true)) return st;
    r->next_row = row0 + nrows;
    r->rows_done += nrows;
    r->viewed = false;
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_view(cubo_hip_reduction *r, cubo_hip_result_view *v) {
    if (!r || !v) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    cubo_hip_cube *c = r->cube;
    if (r->rows_done != c->shape[0])
        return fail(CUBO_HIP_ERR_INCOMPLETE, "the reduction covered %lld of %lld scan rows", (long long)r->rows_done,
                    (long long)c->shape[0]);
    const int64_t npos = c->npos, ndet = c->ndet;
    double *rows = r->res_host;
    double *mean = rows + (size_t)std::max(r->nv, 1) * npos, *max = mean + ndet;
    double *crow = r->res_host + r->res_count, *ccol = crow + npos;
    // The integer kernel's arithmetic CoM rows follow its weight rows; in every other layout the three CoM
    // rows directly follow the masks.
    const int base = (r->path == 1 && r->com_ones) ? r->nw : r->n_masks;
    if (!r->viewed) {
        DeviceGuard g(c->device);
        CUBO_HIP_TRY(hipStreamSynchronize(c->stream), "waiting for the reduction");
        r->viewed = true;
    }
    memset(v, 0, sizeof *v);
    if (r->n_masks > 0) v->images = rows;
    if (r->want_cols) { v->mean = mean; v->max = max; }
    if (r->want_com) {
        v->com_intensity = rows + (size_t)base * npos;
        v->com_moment_row = rows + (size_t)(base + 1) * npos;
        v->com_moment_column = rows + (size_t)(base + 2) * npos;
        v->com_row = crow;
        v->com_column = ccol;
    }
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_results(cubo_hip_reduction *r, cubo_hip_reduce_result *res) {
    if (!r || !res) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    cubo_hip_result_view v;
    if (int st = cubo_hip_reduce_view(r, &v)) return st;
    const int64_t npos = r->cube->npos, ndet = r->cube->ndet;
    auto put = [](double *dst, const double *src, int64_t n) {
        if (dst && src) memcpy(dst, src, n * sizeof(double));
    };
    put(res->images, v.images, (int64_t)r->n_masks * npos);
    put(res->mean, v.mean, ndet);
    put(res->max, v.max, ndet);
    put(res->com_intensity, v.com_intensity, npos);
    put(res->com_moment_row, v.com_moment_row, npos);
    put(res->com_moment_column, v.com_moment_column, npos);
    put(res->com_row, v.com_row, npos);
    put(res->com_column, v.com_column, npos);
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_reset(cubo_hip_reduction *r) {
    if (!r) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL reduction");
    DeviceGuard g(r->cube->device);
    r->rows_done = 0;
    r->next_row = 0;
    r->viewed = false;
    return CUBO_HIP_OK;
}

int cubo_hip_reduce_cancel(cubo_hip_reduction *r) {
    if (!r) return CUBO_HIP_OK;
    DeviceGuard g(r->cube->device);
    (void)hipStreamSynchronize(r->cube->stream);
    free_reduction(r);
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
