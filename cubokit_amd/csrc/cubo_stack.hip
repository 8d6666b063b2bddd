// cubo_stack.hip — C ABI of the mask-stack contraction (include/cubo_hip.h, "mask stack").
//
// Prepares the weights (power-of-two column scales and the FP16 hi/lo split in MFMA operand order for the
// matrix-core kernel; a pixel-major FP32 copy for the plain kernel), launches on the cube's stream and
// returns Y = X W in Double.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/cubo_hip.h"
#include "cube.hpp"
#include "errors.hpp"
#include "maskstack_kernels.hpp"
#include "stack_prep.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::hip_fail;

struct cubo_hip_stack_plan {
    cubo_hip_cube *cube;
    int k;
    int method;          // CUBO_HIP_STACK_PLAIN or CUBO_HIP_STACK_MFMA_SPLIT
    int ntiles;          // MFMA: 32-mask tiles per pass (1 or 2)
    int passes;          // MFMA: passes of up to 64 masks
    int nchunks;         // MFMA: 64-pixel chunks
    int kpad;            // plain: masks per pass (padded)
    void *w;             // device weights
    double *inv_scale;   // MFMA: [k]
    int tile_lo[8];      // MFMA: a 32-mask tile with a non-zero lo part (4 passes x 2 tiles)
    double *y;           // device output [npos][k]
};

namespace {

// Masks per pass of the plain kernel: at most 32 keeps two waves per SIMD (64 FP32 + 64 FP64 accumulators
// would need every register of a wave); more masks take more passes over the cube.
int padded_plain(int k) {
    for (int c : {8, 16, 32})
        if (k <= c) return c;
    return 32;
}

void free_plan(cubo_hip_stack_plan *p) {
    if (!p) return;
    DeviceGuard g(p->cube->device);
    if (p->w) (void)hipFree(p->w);
    if (p->inv_scale) (void)hipFree(p->inv_scale);
    if (p->y) (void)hipFree(p->y);
    delete p;
}

bool mfma_ok(const cubo_hip_cube *c, int k) { return c->dtype == CUBO_HIP_U16 && c->ndet % 8 == 0 && k <= 256; }

}  // namespace

extern "C" {

int cubo_hip_stack_plan_create(cubo_hip_cube *c, const float *masks, int32_t k, int32_t method,
                               cubo_hip_stack_plan **out) {
    if (!c || !masks || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (int st = cubo::require_resident(c, "the mask stack")) return st;
    if (k <= 0) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the stack needs at least one mask");
    const int64_t ndet = c->ndet, npos = c->npos;
    for (int64_t i = 0; i < (int64_t)k * ndet; ++i)
        if (!std::isfinite(masks[i])) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "a mask needs finite weights");
    if (method == CUBO_HIP_STACK_AUTO) method = mfma_ok(c, k) ? CUBO_HIP_STACK_MFMA_SPLIT : CUBO_HIP_STACK_PLAIN;
    if (method == CUBO_HIP_STACK_MFMA_SPLIT && !mfma_ok(c, k))
        return fail(CUBO_HIP_ERR_UNSUPPORTED,
                    "the split-FP16 matrix-core stack needs a uint16 cube, a pattern length divisible by 8 and at "
                    "most 256 masks");
    if (method != CUBO_HIP_STACK_MFMA_SPLIT && method != CUBO_HIP_STACK_PLAIN)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "unknown mask-stack method %d", method);

    auto *p = new (std::nothrow) cubo_hip_stack_plan();
    if (!p) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    p->cube = c;
    p->k = k;
    p->method = method;
    DeviceGuard g(c->device);
    hipError_t e = hipMalloc((void **)&p->y, (size_t)npos * k * sizeof(double));
    if (e != hipSuccess) { free_plan(p); return hip_fail(e, "allocating the stack output"); }

    if (method == CUBO_HIP_STACK_MFMA_SPLIT) {
        cubo::SplitWeights sw = cubo::split_weights(masks, k, ndet);
        p->ntiles = sw.ntiles;
        p->passes = sw.passes;
        p->nchunks = sw.nchunks;
        for (size_t t = 0; t < sw.tile_lo.size() && t < 8; ++t) p->tile_lo[t] = sw.tile_lo[t];
        const std::vector<_Float16> &wb = sw.wb;
        const std::vector<double> &inv = sw.inv;
        e = hipMalloc(&p->w, wb.size() * sizeof(_Float16));
        if (e == hipSuccess) e = hipMemcpy(p->w, wb.data(), wb.size() * sizeof(_Float16), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc((void **)&p->inv_scale, k * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(p->inv_scale, inv.data(), k * sizeof(double), hipMemcpyHostToDevice);
    } else {
        // Pixel-major FP32 weights, one block of kpad columns per pass.
        p->kpad = padded_plain(k);
        const int passes = (k + p->kpad - 1) / p->kpad;
        std::vector<float> w((size_t)passes * ndet * p->kpad, 0.f);
        for (int m = 0; m < k; ++m) {
            const int pass = m / p->kpad, col = m % p->kpad;
            for (int64_t i = 0; i < ndet; ++i)
                w[((size_t)pass * ndet + i) * p->kpad + col] = masks[(size_t)m * ndet + i];
        }
        e = hipMalloc(&p->w, w.size() * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(p->w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) { free_plan(p); return hip_fail(e, "uploading the mask stack"); }
    *out = p;
    return CUBO_HIP_OK;
}

int cubo_hip_stack_plan_method(const cubo_hip_stack_plan *p, int32_t *method) {
    if (!p || !method) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *method = p->method;
    return CUBO_HIP_OK;
}

int cubo_hip_stack_plan_run(cubo_hip_stack_plan *p, double *y) {
    if (!p || !y) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    cubo_hip_cube *c = p->cube;
    DeviceGuard g(c->device);
    CUBO_HIP_TRY(c->timer.start(c->stream), "hipEventRecord");
    hipError_t e = hipSuccess;
    if (p->method == CUBO_HIP_STACK_MFMA_SPLIT) {
        const size_t per_pass = (size_t)p->ntiles * 4 * 2 * 64 * 8 * p->nchunks;
        for (int pass = 0; pass < p->passes && e == hipSuccess; ++pass) {
            cubo::StackLaunch a{};
            a.cube = (const uint16_t *)c->data;
            a.npos = c->npos;
            a.ndet = c->ndet;
            a.nchunks = p->nchunks;
            a.k = p->k;
            a.k0 = pass * 64;
            a.kn = std::min(64, p->k - pass * 64);
            a.wb = (const _Float16 *)p->w + pass * per_pass;
            a.inv_scale = p->inv_scale + pass * 64;
            a.tile_lo[0] = p->tile_lo[pass * 2];
            a.tile_lo[1] = p->tile_lo[pass * 2 + 1];
            a.y = p->y;
            e = cubo::launch_mask_stack_mfma(a, p->ntiles, c->stream);
        }
    } else {
        const int passes = (p->k + p->kpad - 1) / p->kpad;
        for (int pass = 0; pass < passes && e == hipSuccess; ++pass) {
            cubo::PlainStackLaunch a{};
            a.cube = c->data;
            a.npos = c->npos;
            a.ndet = c->ndet;
            a.k = p->k;
            a.k0 = pass * p->kpad;
            a.kn = std::min(p->kpad, p->k - pass * p->kpad);
            a.w = (const float *)p->w + (size_t)pass * c->ndet * p->kpad;
            a.y = p->y;
            e = cubo::launch_mask_stack_plain(a, c->dtype == CUBO_HIP_U16, p->kpad, c->stream);
        }
    }
    if (e != hipSuccess) return hip_fail(e, "launching the mask stack");
    CUBO_HIP_TRY(c->timer.stop(c->stream), "hipEventRecord");
    CUBO_HIP_TRY(hipMemcpyAsync(y, p->y, (size_t)c->npos * p->k * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                 "reading the stack");
    CUBO_HIP_TRY(hipStreamSynchronize(c->stream), "waiting for the stack");
    return CUBO_HIP_OK;
}

int cubo_hip_stack_plan_destroy(cubo_hip_stack_plan *p) {
    if (!p) return CUBO_HIP_OK;
    DeviceGuard g(p->cube->device);
    (void)hipStreamSynchronize(p->cube->stream);
    free_plan(p);
    return CUBO_HIP_OK;
}

int cubo_hip_mask_stack(cubo_hip_cube *c, const float *masks, int32_t k, int32_t method, double *y) {
    cubo_hip_stack_plan *p = nullptr;
    if (int st = cubo_hip_stack_plan_create(c, masks, k, method, &p)) return st;
    const int st = cubo_hip_stack_plan_run(p, y);
    cubo_hip_stack_plan_destroy(p);
    return st;
}

}  // extern "C"
