// cubo_origin.hip — C ABI of the origins (include/cubo_hip.h, "origins"): ProbeSize.estimate on the host, and the
// per-position origins of a tile of scan rows on the device by py4DSTEM's get_origin or the iterated window.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
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

// The matrix-core get_origin's filter matrix (py4_origin_mfma): G[a][b] = the sum of the weights of the taps of
// output a that mode nearest maps to input b, scaled by a power of two s (largest entry below 60000) and split into
// FP16 hi + lo (lo parts below FP16's normal range dropped), as [4][n][n]: hi and lo, then both again with each
// row's columns in the order of the second product's k-steps: column 16 k + 8 h + e of the copy holds column
// 32 (k / 2) + 16 (k % 2) + 8 (e / 4) + 4 h + e % 4 of G. Returns false when the representation's relative error
// would not leave the candidates' margin at least twice the filter's relative error bound.
struct MfmaFilter {
    std::vector<_Float16> g;
    float pass2_scale = 0, margin = 0, slack = 0;
};
bool mfma_filter(const std::vector<double> &w, int n, MfmaFilter *f) {
    const int R = (int)(w.size() / 2);
    std::vector<double> G((size_t)n * n, 0.0);
    for (int a = 0; a < n; ++a)
        for (int t = -R; t <= R; ++t) {
            const int b = a + t < 0 ? 0 : (a + t >= n ? n - 1 : a + t);
            G[(size_t)a * n + b] += w[R + t];
        }
    double gmax = 0, rowmax = 0;
    for (int a = 0; a < n; ++a) {
        double rs = 0;
        for (int b = 0; b < n; ++b) { gmax = std::max(gmax, G[(size_t)a * n + b]); rs += G[(size_t)a * n + b]; }
        rowmax = std::max(rowmax, rs);
    }
    if (!(gmax > 0)) return false;
    const double s = std::ldexp(1.0, (int)std::floor(std::log2(60000.0 / gmax)));
    const double tiny = std::ldexp(1.0, -14);             // FP16's smallest normal
    f->g.assign((size_t)4 * n * n, (_Float16)0.f);
    double rep = 0;                                        // largest relative error of hi + lo
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            const double v = s * G[(size_t)a * n + b];
            _Float16 hi = (_Float16)v;
            if (std::fabs((double)hi) < tiny) hi = (_Float16)0.f;
            _Float16 lo = (_Float16)(v - (double)hi);
            if (std::fabs((double)lo) < tiny) lo = (_Float16)0.f;
            if (v > 0) rep = std::max(rep, std::fabs((double)hi + (double)lo - v) / v);
            f->g[(size_t)a * n + b] = hi;
            f->g[(size_t)n * n + (size_t)a * n + b] = lo;
        }
    for (int a = 0; a < n; ++a)
        for (int c = 0; c < n; ++c) {
            const int k = c / 16, hh = (c / 8) % 2, e = c % 8;
            const int b = 32 * (k / 2) + 16 * (k % 2) + 8 * (e / 4) + 4 * hh + e % 4;
            f->g[(size_t)2 * n * n + (size_t)a * n + c] = f->g[(size_t)a * n + b];
            f->g[(size_t)3 * n * n + (size_t)a * n + c] = f->g[(size_t)n * n + (size_t)a * n + b];
        }
    // the first product's sums reach s * 65535 * (largest row sum): scaled into FP16's range for the second
    const double b1max = s * 65535.0 * rowmax * (1 + 1e-3);
    const double p2 = std::ldexp(1.0, -(int)std::ceil(std::log2(b1max / 30000.0)));
    // e: the two FP32 sums (at most 512 terms each), G's representation, B1's split and the dropped lo x lo term
    const double e = 2 * 512 * std::ldexp(1.0, -24) + rep + 2 * std::ldexp(1.0, -22);
    // the candidates' margin: twice the bound with room to spare (rounded up to 5e-4, or 1e-3)
    const double margin = 2 * e < 2.5e-4 ? 5e-4 : 1e-3;
    if (!(2 * e < margin)) return false;
    f->pass2_scale = (float)p2;
    f->margin = (float)margin;
    // second-pass inputs below FP16's normal range: at most 2^-13 each, times a row of the scaled G, both sides
    f->slack = (float)(2 * s * rowmax * (1 + 1e-3) * std::ldexp(1.0, -13));
    return true;
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
    _Float16 *d_g = nullptr;
    int32_t *d_undecided = nullptr;      // [npos] positions, then their count
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
        // uint16 patterns of 128 x 128: the matrix-core kernel, the exact one for the patterns it leaves undecided
        // (CUBO_HIP_ORIGIN_KERNEL=exact keeps the exact kernel for every pattern: a test knob)
        MfmaFilter mf;
        const char *knob = getenv("CUBO_HIP_ORIGIN_KERNEL");
        const bool mfma = cubo::py4_mfma_supported(a) && !(knob && strcmp(knob, "exact") == 0) && mfma_filter(w, a.qx, &mf);
        if (e == hipSuccess && mfma) {
            e = hipMalloc((void **)&d_g, mf.g.size() * sizeof(_Float16));
            if (e == hipSuccess) e = hipMalloc((void **)&d_undecided, (size_t)(npos + 1) * 4);
            if (e == hipSuccess)
                e = hipMemcpyAsync(d_g, mf.g.data(), mf.g.size() * sizeof(_Float16), hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipMemsetAsync(d_undecided + npos, 0, 4, s);
        }
        if (e == hipSuccess) {
            e = c->timer.start(s);     // the launches' device time, reported by cubo_hip_cube_kernel_time
            if (e == hipSuccess && mfma) {
                cubo::Py4Launch m = a;
                m.gmat = d_g;
                m.pass2_scale = mf.pass2_scale;
                m.candidate_margin = mf.margin;
                m.slack = mf.slack;
                m.undecided = d_undecided;
                m.n_undecided = d_undecided + npos;
                e = cubo::launch_py4_origin_mfma(m, (int)std::min<int64_t>(cubo::compute_units(c->device), npos), s);
                a.list = d_undecided;
                a.nlist = d_undecided + npos;
            }
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
    for (void *b : {(void *)d_out, (void *)d_w, (void *)scratch, (void *)d_take, (void *)d_g, (void *)d_undecided})
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
