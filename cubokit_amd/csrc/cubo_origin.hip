// cubo_origin.hip — C ABI of the origins (include/cubo_hip.h, "origins"): ProbeSize.estimate on the host, and the
// per-position origins of a tile of scan rows on the device by py4DSTEM's get_origin or the iterated window.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/cubo_hip.h"
#include "cube.hpp"
#include "errors.hpp"
#include "host_math.hpp"
#include "origin_kernels.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::hip_fail;

namespace {

// GaussianFilter.init's checks (OriginMeasurement.swift:115-124) with the default maximum radius of 10^6.
bool valid_sigma(double s) {
    return s > 0 && std::isfinite(s) && s * s >= 2.2250738585072014e-308 && 4.0 * s + 0.5 < 1e6;
}

// SciPyKernels.gaussianWeights (SciPyKernels.swift:71-78): exp(-0.5/sigma^2 x^2) over NumPy's pairwise sum.
std::vector<double> gaussian_weights(double sigma) {
    const int R = (int)(4.0 * sigma + 0.5);
    const double sigma2 = sigma * sigma;
    std::vector<double> w(2 * R + 1);
    for (int x = -R; x <= R; ++x) w[x + R] = std::exp(-0.5 / sigma2 * (double)(x * x));
    const double s = cubo::pairwise_sum(w.data(), (int64_t)w.size());
    for (double &v : w) v = v / s;
    return w;
}

}  // namespace

extern "C" {

int cubo_hip_probe_size(const double *v, int32_t rows, int32_t cols, double lower, double upper, int32_t count,
                        cubo_hip_probe_estimate *out) {
    if (!v || !out || rows <= 0 || cols <= 0) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument or empty pattern");
    if (count < 2) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "probeNotFound: needs at least two thresholds");
    const int64_t n = (int64_t)rows * cols;
    // NumPy's linspace: k * step + start, the last value set to stop
    const double step = (upper - lower) / (double)(count - 1);
    std::vector<double> th(count), radii(count), grad(count);
    for (int k = 0; k < count; ++k) th[k] = (double)k * step + lower;
    th[count - 1] = upper;
    double maximum = -INFINITY;
    for (int64_t i = 0; i < n; ++i) maximum = (std::isnan(maximum) || std::isnan(v[i])) ? NAN : std::max(maximum, v[i]);
    for (int k = 0; k < count; ++k) {
        const double limit = maximum * th[k];
        int64_t c = 0;
        for (int64_t i = 0; i < n; ++i) c += v[i] > limit;
        radii[k] = std::sqrt((double)c / M_PI);
    }
    // np.gradient with unit spacing
    grad[0] = radii[1] - radii[0];
    grad[count - 1] = radii[count - 1] - radii[count - 2];
    for (int k = 1; k < count - 1; ++k) grad[k] = (radii[k + 1] - radii[k - 1]) / 2.0;
    std::vector<double> sorted = grad;
    std::sort(sorted.begin(), sorted.end());
    const double median = count % 2 ? sorted[count / 2] : (sorted[count / 2 - 1] + sorted[count / 2]) / 2;
    std::vector<double> tr, tt;
    for (int k = 0; k < count; ++k)
        if (grad[k] <= 0 && grad[k] >= 2 * median) { tr.push_back(radii[k]); tt.push_back(th[k]); }
    const double radius = tr.empty() ? NAN : cubo::pairwise_sum(tr.data(), (int64_t)tr.size()) / (double)tr.size();
    const double threshold = tt.empty() ? NAN : cubo::pairwise_sum(tt.data(), (int64_t)tt.size()) / (double)tt.size();
    // get_CoM(DP * mask) with NumPy's sums
    const double limit = maximum * threshold;
    std::vector<double> m(n), mr(n), mc(n);
    for (int64_t i = 0; i < n; ++i) {
        m[i] = v[i] * (v[i] > limit ? 1.0 : 0.0);
        mr[i] = (double)(i / cols) * m[i];
        mc[i] = (double)(i % cols) * m[i];
    }
    const double total = cubo::pairwise_sum(m.data(), n);
    const double row = cubo::pairwise_sum(mr.data(), n) / total, col = cubo::pairwise_sum(mc.data(), n) / total;
    if (!std::isfinite(radius) || !std::isfinite(row) || !std::isfinite(col))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "probeNotFound: %s",
                    tr.empty() ? "no trusted threshold" : "the masked pattern sums to zero");
    out->radius = radius;
    out->center_row = row;
    out->center_column = col;
    out->threshold = threshold;
    return CUBO_HIP_OK;
}

int cubo_hip_origins(cubo_hip_cube *c, const cubo_hip_origin_request *q, int64_t row0, int64_t nrows, double *out_row,
                     double *out_col) {
    if (!c || !q || !out_row || !out_col) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (int st = cubo::rows_range(c, row0, nrows)) return st;
    const double r = q->radius, scale = q->radius_scale;
    if (q->estimator == CUBO_HIP_ORIGIN_PY4DSTEM) {
        if (!valid_sigma(r)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalidRadius: %g", r);
        if (!(scale > 0) || !std::isfinite(scale)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalidRadius: %g", scale);
    } else if (q->estimator == CUBO_HIP_ORIGIN_ITERATED_WINDOW) {
        // OriginMeasurement.measureOrigins (IteratedWindowOrigin.swift:200-213)
        if (!(r > 0 && r < 1e6)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalidRadius: %g", r);
        if (!(scale > 0 && scale < 1e6 && r * scale < 1e6))
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalidRadius: %g", scale);
        if (q->passes < 0 || q->passes > 1000)
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the pass count must be from 0 to 1000 (got %d)", q->passes);
        if (!(q->minimum_margin >= 0) || !std::isfinite(q->minimum_margin))
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the minimum margin must be finite and not negative");
        if (q->scan_mask || q->fast_center)
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "a scan mask and the fast centre belong to py4DSTEM's estimator");
    } else {
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "unknown estimator %d", q->estimator);
    }
    if (nrows == 0) return CUBO_HIP_OK;
    const void *tile = nullptr;
    if (int st = cubo::rows_data(c, row0, nrows, &tile)) return st;
    DeviceGuard g(c->device);
    hipStream_t s = c->stream;
    const int64_t pos0 = row0 * c->shape[1], npos = nrows * c->shape[1];
    double *d_out = nullptr;
    CUBO_HIP_TRY(hipMalloc((void **)&d_out, (size_t)npos * 16), "allocating the origins");
    hipError_t e = hipSuccess;
    double *d_w = nullptr, *scratch = nullptr;
    uint8_t *d_take = nullptr;
    // host data of the copies below stays alive until the stream is synchronised at the end
    const std::vector<double> w = q->estimator == CUBO_HIP_ORIGIN_PY4DSTEM ? gaussian_weights(r) : std::vector<double>();
    if (q->estimator == CUBO_HIP_ORIGIN_PY4DSTEM) {
        const int blocks = cubo::py4_blocks(cubo::compute_units(c->device), npos);
        e = hipMalloc((void **)&d_w, w.size() * 8);
        if (e == hipSuccess) e = hipMalloc((void **)&scratch, (size_t)blocks * c->ndet * 8);
        if (e == hipSuccess && q->scan_mask) e = hipMalloc((void **)&d_take, (size_t)npos);
        if (e == hipSuccess) e = hipMemcpyAsync(d_w, w.data(), w.size() * 8, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && q->scan_mask)
            e = hipMemcpyAsync(d_take, q->scan_mask + pos0, (size_t)npos, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            cubo::Py4Launch a{};
            a.cube = tile;
            a.u16 = c->dtype == CUBO_HIP_U16;
            a.npos = npos;
            a.qy = (int)c->shape[2];
            a.qx = (int)c->shape[3];
            a.weights = d_w;
            a.R = (int)(w.size() / 2);
            a.limit = r * scale;
            a.take = d_take;
            a.fast_center = q->fast_center;
            a.scratch = scratch;
            a.out_row = d_out;
            a.out_col = d_out + npos;
            a.filtered = nullptr;
            e = c->timer.start(s);     // the launch's device time, reported by cubo_hip_cube_kernel_time
            if (e == hipSuccess) e = cubo::launch_py4_origin(a, blocks, s);
            if (e == hipSuccess) e = c->timer.stop(s);
        }
    } else {
        cubo::IWLaunch a{};
        a.cube = tile;
        a.u16 = c->dtype == CUBO_HIP_U16;
        a.npos = npos;
        a.qy = (int)c->shape[2];
        a.qx = (int)c->shape[3];
        a.radius = (float)r;
        a.radius_scale = (float)scale;
        a.margin = (float)q->minimum_margin;
        a.passes = q->passes;
        a.binned = q->binned_coarse;
        a.out_row = d_out;
        a.out_col = d_out + npos;
        e = c->timer.start(s);
        if (e == hipSuccess) e = cubo::launch_iterated_window(a, s);
        if (e == hipSuccess) e = c->timer.stop(s);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out_row + pos0, d_out, (size_t)npos * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(out_col + pos0, d_out + npos, (size_t)npos * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    else (void)hipStreamSynchronize(s);
    for (void *b : {(void *)d_out, (void *)d_w, (void *)scratch, (void *)d_take})
        if (b) (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(e, "measuring the origins");
    return CUBO_HIP_OK;
}

int cubo_hip_gaussian_filter_pattern(cubo_hip_cube *c, int64_t row, int64_t col, double sigma, double *out) {
    if (!c || !out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (row < 0 || row >= c->shape[0] || col < 0 || col >= c->shape[1])
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "position (%lld, %lld) outside the scan", (long long)row, (long long)col);
    if (!valid_sigma(sigma)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalidRadius: %g", sigma);
    if (int st = cubo::require_resident(c, "filtering one pattern")) return st;
    DeviceGuard g(c->device);
    hipStream_t s = c->stream;
    const std::vector<double> w = gaussian_weights(sigma);
    const int64_t ndet = c->ndet;
    double *d_w = nullptr, *scratch = nullptr, *d_out = nullptr, *filtered = nullptr;
    hipError_t e = hipMalloc((void **)&d_w, w.size() * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&scratch, (size_t)ndet * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&filtered, (size_t)ndet * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&d_out, 16);
    if (e == hipSuccess) e = hipMemcpyAsync(d_w, w.data(), w.size() * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        cubo::Py4Launch a{};
        a.cube = (const char *)c->data + (size_t)(row * c->shape[1] + col) * ndet * cubo::element_size(c->dtype);
        a.u16 = c->dtype == CUBO_HIP_U16;
        a.npos = 1;
        a.qy = (int)c->shape[2];
        a.qx = (int)c->shape[3];
        a.weights = d_w;
        a.R = (int)(w.size() / 2);
        a.limit = 1.0;
        a.fast_center = 1;
        a.scratch = scratch;
        a.out_row = d_out;
        a.out_col = d_out + 1;
        a.filtered = filtered;
        e = cubo::launch_py4_origin(a, 1, s);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, filtered, (size_t)ndet * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    else (void)hipStreamSynchronize(s);
    for (void *b : {(void *)d_w, (void *)scratch, (void *)d_out, (void *)filtered})
        if (b) (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(e, "filtering the pattern");
    return CUBO_HIP_OK;
}

}  // extern "C"
