// cubo_phase.hip — C ABI of the CuboPhase field preparation on the device, in Double
// (include/cubo_hip.h, "centre-of-mass reference and rotation").
//
// CenterOfMassReference.normalize (Sources/CuboPhase/CenterOfMassReference.swift:27-84): a least-squares plane
// a + b u + c v over the finite positions (normal equations on centred coordinates, degenerate when
// det <= tol * max(suu svv, 1)), the constant mean, a given field, or nothing, subtracted; NaN -> 0, +-inf ->
// +-max finite; times the reciprocal sampling. RotationSolver.solve (Sources/CuboPhase/RotationSolver.swift:
// 48-150): for each angle and both orientations (components swapped first when transposed) the mean over the
// scan's interior of |d_row y' - d_col x'| (or of the divergence when maximising); the first best angle wins and
// the plain orientation wins ties; a forced angle and/or orientation narrows what is scored. Each (angle,
// orientation) is one block group on the device; the sums of the blocks are added on the host in block order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "../../include/cubo_hip.h"
#include "errors.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::hip_fail;

namespace {

constexpr int kThreads = 256;
constexpr int kSplit = 16;            // blocks per (angle, orientation)

__device__ __forceinline__ double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double out = red[0];
    __syncthreads();
    return out;
}

// mode 0: count, sum u, sum v of the finite entries; mode 1: suu, suv, svv, sum m; mode 2: sum u (m - a), sum v (m - a)
__global__ __launch_bounds__(kThreads) void plane_sums(const double *m, int64_t n, int cols, int mode, double rmean,
                                                       double cmean, double a, double *partial) {
    __shared__ double red[kThreads];
    double s[4] = {0, 0, 0, 0};
    for (int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k < n; k += (int64_t)gridDim.x * kThreads) {
        const double v = m[k];
        if (!isfinite(v)) continue;
        const double r = (double)(k / cols), c = (double)(k % cols);
        if (mode == 0) { s[0] += 1; s[1] += r; s[2] += c; }
        else {
            const double u = r - rmean, w = c - cmean;
            if (mode == 1) { s[0] += u * u; s[1] += u * w; s[2] += w * w; s[3] += v; }
            else { s[0] += u * (v - a); s[1] += w * (v - a); }
        }
    }
    for (int i = 0; i < 4; ++i) {
        const double t = block_sum(s[i], red);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * 4 + i] = t;
    }
}

__global__ void plane_eval(double *out, int64_t n, int cols, double a, double b, double c, double rmean,
                           double cmean) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = a + b * ((double)(k / cols) - rmean) + c * ((double)(k % cols) - cmean);
}

__device__ __forceinline__ double clean(double v) {
    if (isnan(v)) return 0.0;
    if (v == INFINITY) return DBL_MAX;
    if (v == -INFINITY) return -DBL_MAX;
    return v;
}

__global__ void normalize_apply(const double *mx, const double *my, const double *fx, const double *fy, int64_t n,
                                double sampling, double *nx, double *ny) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    nx[k] = clean(mx[k] - fx[k]) * sampling;
    ny[k] = clean(my[k] - fy[k]) * sampling;
}

// blockIdx.x = split, blockIdx.y = angle, blockIdx.z = the z-th orientation scored, t_first + z (1: components
// swapped first)
__global__ __launch_bounds__(kThreads) void rotation_sums(const double *nx, const double *ny, int R, int C,
                                                          const double *cosv, const double *sinv, int maximize,
                                                          int t_first, double *partial) {
    __shared__ double red[kThreads];
    const int angle = blockIdx.y, t = t_first + (int)blockIdx.z;
    const double co = cosv[angle], si = sinv[angle];
    const double *A = t ? ny : nx, *B = t ? nx : ny;
    auto xr = [&](int64_t k) { return co * A[k] - si * B[k]; };   // rotated components (RotationSolver.swift:56-62)
    auto yr = [&](int64_t k) { return si * A[k] + co * B[k]; };
    const int64_t ni = (int64_t)(R - 2) * (C - 2);
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < ni; i += (int64_t)kSplit * kThreads) {
        const int r = 1 + (int)(i / (C - 2)), c = 1 + (int)(i % (C - 2));
        const int64_t k = (int64_t)r * C + c;
        double v;
        if (maximize) v = fabs((xr(k + C) - xr(k - C)) + (yr(k + 1) - yr(k - 1)));
        else v = fabs((yr(k + C) - yr(k - C)) - (xr(k + 1) - xr(k - 1)));
        s += v;
    }
    const double tot = block_sum(s, red);
    if (threadIdx.x == 0) partial[((int64_t)blockIdx.z * gridDim.y + angle) * kSplit + blockIdx.x] = tot;
}

__global__ void rotate_apply(const double *nx, const double *ny, int64_t n, double co, double si, int t, double *cx,
                             double *cy) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double a = t ? ny[k] : nx[k], b = t ? nx[k] : ny[k];
    cx[k] = co * a - si * b;
    cy[k] = si * a + co * b;
}

unsigned blocks(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// The stream and device buffers of the phase calls on one device, kept between calls: a buffer only grows, so
// a warm call allocates and frees nothing (hipFree can wait for the whole device). Calls on one device take
// turns; the workspace lives until the process ends.
struct PhaseWorkspace {
    std::mutex mu;
    hipStream_t s = nullptr;
    std::vector<std::pair<void *, size_t>> slots;    // the k-th buffer a call asks for, and its capacity in bytes
};

PhaseWorkspace &workspace(int device) {
    static std::mutex map_mu;
    static std::map<int, PhaseWorkspace *> spaces;
    std::lock_guard<std::mutex> l(map_mu);
    PhaseWorkspace *&w = spaces[device];
    if (!w) w = new PhaseWorkspace();
    return *w;
}

// One call's hold on its device's workspace (construct with the device current).
struct Dev {
    int device;
    PhaseWorkspace &w;
    std::unique_lock<std::mutex> lock;
    hipStream_t s = nullptr;
    size_t next = 0;
    explicit Dev(int dev) : device(dev), w(workspace(dev)), lock(w.mu) {}
    ~Dev() {
        DeviceGuard g(device);
        if (s) (void)hipStreamSynchronize(s);        // the next call may reuse or regrow the buffers
    }
    hipError_t init() {
        if (!w.s) {
            if (hipError_t e = hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking)) return e;
        }
        s = w.s;
        return hipSuccess;
    }
    template <typename T>
    hipError_t alloc(T **p, size_t n) {
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        if (next == w.slots.size()) w.slots.push_back({nullptr, 0});
        std::pair<void *, size_t> &sl = w.slots[next++];
        if (sl.second < bytes) {
            if (sl.first) (void)hipFree(sl.first);
            sl = {nullptr, 0};
            if (hipError_t e = hipMalloc(&sl.first, bytes)) return e;
            sl.second = bytes;
        }
        *p = (T *)sl.first;
        return hipSuccess;
    }
};

int device_ok(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return fail(CUBO_HIP_ERR_NO_DEVICE, "no HIP device %d", device);
    return CUBO_HIP_OK;
}

// The planes of the two components m[0], m[1] (device) into out[0], out[1] (device);
// CUBO_HIP_ERR_INVALID_ARGUMENT when degenerate. Each of the three reduction passes runs for both components
// before the host reads their block sums (in block order) once. partial holds 2 x nb x 4 doubles.
int planes(Dev &d, const double *const m[2], int R, int C, double tol, double *const out[2], double *partial, int nb) {
    const int64_t n = (int64_t)R * C;
    std::vector<double> h((size_t)2 * nb * 4);
    auto sums = [&](int mode, const double (&rm)[2], const double (&cm)[2], const double (&a)[2],
                    double (&tot)[2][4]) -> hipError_t {
        for (int f = 0; f < 2; ++f)
            hipLaunchKernelGGL(plane_sums, dim3(nb), dim3(kThreads), 0, d.s, m[f], n, C, mode, rm[f], cm[f], a[f],
                               partial + (size_t)f * nb * 4);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), partial, h.size() * 8, hipMemcpyDeviceToHost, d.s);
        if (e == hipSuccess) e = hipStreamSynchronize(d.s);
        for (int f = 0; f < 2; ++f)
            for (int i = 0; i < 4; ++i) {
                tot[f][i] = 0;
                for (int b = 0; b < nb; ++b) tot[f][i] += h[((size_t)f * nb + b) * 4 + i];
            }
        return e;
    };
    const double zero[2] = {0, 0};
    double t0[2][4], t1[2][4], t2[2][4], rm[2], cm[2], a[2], cnt[2];
    CUBO_HIP_TRY(sums(0, zero, zero, zero, t0), "plane sums");
    for (int f = 0; f < 2; ++f) {
        if (t0[f][0] < 3) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the plane fit needs 3 finite positions");
        cnt[f] = t0[f][0];
        rm[f] = t0[f][1] / cnt[f];
        cm[f] = t0[f][2] / cnt[f];
    }
    CUBO_HIP_TRY(sums(1, rm, cm, zero, t1), "plane sums");
    for (int f = 0; f < 2; ++f) a[f] = t1[f][3] / cnt[f];
    CUBO_HIP_TRY(sums(2, rm, cm, a, t2), "plane sums");
    for (int f = 0; f < 2; ++f) {
        const double suu = t1[f][0], suv = t1[f][1], svv = t1[f][2];
        const double det = suu * svv - suv * suv;
        if (!(det > tol * std::fmax(suu * svv, 1.0)))
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the plane fit is degenerate (the positions are collinear)");
        const double b = (svv * t2[f][0] - suv * t2[f][1]) / det, c = (suu * t2[f][1] - suv * t2[f][0]) / det;
        hipLaunchKernelGGL(plane_eval, dim3(blocks(n)), dim3(kThreads), 0, d.s, out[f], n, C, a[f], b, c, rm[f], cm[f]);
        CUBO_HIP_TRY(hipGetLastError(), "plane evaluation");
    }
    return CUBO_HIP_OK;
}

}  // namespace

extern "C" {

int cubo_hip_phase_normalize(int device, const double *mx, const double *my, int32_t rows, int32_t cols,
                             int32_t reference, double sampling, double tol, const double *given_row,
                             const double *given_col, double *fx, double *fy, double *nx, double *ny) {
    if (!mx || !my || !fx || !fy || !nx || !ny) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (rows < 1 || cols < 1) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "empty scan");
    if (reference < CUBO_HIP_REFERENCE_NONE || reference > CUBO_HIP_REFERENCE_GIVEN)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "reference %d (0 none, 1 plane, 2 constant, 3 given)", reference);
    if (reference == CUBO_HIP_REFERENCE_GIVEN && (!given_row || !given_col))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the given reference needs both components");
    if (!(tol >= 0) || !std::isfinite(tol))                  // CenterOfMassReference.swift:30-32
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the degeneracy tolerance must be finite and not negative (%g)", tol);
    if (int st = device_ok(device)) return st;
    DeviceGuard g(device);
    Dev d(device);
    CUBO_HIP_TRY(d.init(), "hipStreamCreate");
    const int64_t n = (int64_t)rows * cols;
    const int nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, (n + kThreads - 1) / kThreads));
    double *dmx, *dmy, *dfx, *dfy, *dnx, *dny, *partial;
    for (double **p : {&dmx, &dmy, &dfx, &dfy, &dnx, &dny}) CUBO_HIP_TRY(d.alloc(p, n), "allocating the field");
    CUBO_HIP_TRY(d.alloc(&partial, (size_t)2 * nb * 4), "allocating partial sums");
    CUBO_HIP_TRY(hipMemcpyAsync(dmx, mx, n * 8, hipMemcpyHostToDevice, d.s), "uploading the field");
    CUBO_HIP_TRY(hipMemcpyAsync(dmy, my, n * 8, hipMemcpyHostToDevice, d.s), "uploading the field");
    if (reference == CUBO_HIP_REFERENCE_PLANE) {
        const double *const m[2] = {dmx, dmy};
        double *const out[2] = {dfx, dfy};
        if (int st = planes(d, m, rows, cols, tol, out, partial, nb)) return st;
    } else if (reference == CUBO_HIP_REFERENCE_GIVEN) {    // .given(field): the field is the reference (:45-49)
        CUBO_HIP_TRY(hipMemcpyAsync(dfx, given_row, n * 8, hipMemcpyHostToDevice, d.s), "uploading the reference");
        CUBO_HIP_TRY(hipMemcpyAsync(dfy, given_col, n * 8, hipMemcpyHostToDevice, d.s), "uploading the reference");
    } else {
        double a = 0, b = 0;
        if (reference == CUBO_HIP_REFERENCE_CONSTANT) {      // NumPy means of the raw components (no finite filter)
            std::vector<double> t(mx, mx + n), u(my, my + n);
            auto pw = [](const double *v, int64_t m, auto &&self) -> double {
                if (m < 8) { double s = 0; for (int64_t i = 0; i < m; ++i) s += v[i]; return s; }
                if (m <= 128) {
                    double r[8];
                    for (int j = 0; j < 8; ++j) r[j] = v[j];
                    int64_t i = 8;
                    for (; i < m - (m % 8); i += 8)
                        for (int j = 0; j < 8; ++j) r[j] += v[i + j];
                    double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
                    for (; i < m; ++i) s += v[i];
                    return s;
                }
                int64_t m2 = m / 2;
                m2 -= m2 % 8;
                return self(v, m2, self) + self(v + m2, m - m2, self);
            };
            a = pw(t.data(), n, pw) / (double)n;
            b = pw(u.data(), n, pw) / (double)n;
        }
        std::vector<double> ca(n, a), cb(n, b);
        CUBO_HIP_TRY(hipMemcpyAsync(dfx, ca.data(), n * 8, hipMemcpyHostToDevice, d.s), "uploading the reference");
        CUBO_HIP_TRY(hipMemcpyAsync(dfy, cb.data(), n * 8, hipMemcpyHostToDevice, d.s), "uploading the reference");
        CUBO_HIP_TRY(hipStreamSynchronize(d.s), "waiting");
    }
    hipLaunchKernelGGL(normalize_apply, dim3(blocks(n)), dim3(kThreads), 0, d.s, dmx, dmy, dfx, dfy, n, sampling, dnx,
                       dny);
    CUBO_HIP_TRY(hipGetLastError(), "normalising");
    CUBO_HIP_TRY(hipMemcpyAsync(fx, dfx, n * 8, hipMemcpyDeviceToHost, d.s), "reading the reference");
    CUBO_HIP_TRY(hipMemcpyAsync(fy, dfy, n * 8, hipMemcpyDeviceToHost, d.s), "reading the reference");
    CUBO_HIP_TRY(hipMemcpyAsync(nx, dnx, n * 8, hipMemcpyDeviceToHost, d.s), "reading the field");
    CUBO_HIP_TRY(hipMemcpyAsync(ny, dny, n * 8, hipMemcpyDeviceToHost, d.s), "reading the field");
    CUBO_HIP_TRY(hipStreamSynchronize(d.s), "waiting for the normalisation");
    return CUBO_HIP_OK;
}

int cubo_hip_phase_rotation(int device, const double *nx, const double *ny, int32_t rows, int32_t cols,
                            const double *angles_deg, int32_t nangles, int32_t maximize, const double *force_degrees,
                            int32_t force_transpose, const volatile int32_t *cancel, double *degrees,
                            int32_t *transposed, int32_t *n_curve, int32_t *n_curve_t, double *curve, double *curve_t,
                            double *cx, double *cy) {
    if (!nx || !ny || !degrees || !transposed || !n_curve || !n_curve_t || !cx || !cy)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (force_transpose < -1 || force_transpose > 1)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "force_transpose %d (-1, 0, 1)", force_transpose);
    // requireUsable (RotationSolver.swift:128-135) comes first on every branch
    if (rows < 3 || cols < 3) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the scan needs at least 3 x 3 positions");
    const int64_t n = (int64_t)rows * cols, ni = (int64_t)(rows - 2) * (cols - 2);
    for (int64_t k = 0; k < n; ++k)
        if (!std::isfinite(nx[k]) || !std::isfinite(ny[k]))
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the field must be finite (entry %lld)", (long long)k);
    if (cancel && *cancel) return fail(CUBO_HIP_ERR_CANCELLED, "cancelled");
    // what is scored (RotationSolver.swift:66-90): a forced angle with both orientations, a forced orientation over
    // every angle, both orientations over every angle, or nothing when angle and orientation are both forced
    const double *score_angles = angles_deg;
    int na = nangles, t_first = 0, nt = 2;
    if (force_degrees) {
        if (!std::isfinite(*force_degrees)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalidAngles: the forced angle");
        score_angles = force_degrees;
        na = force_transpose >= 0 ? 0 : 1;
    } else {
        if (!angles_deg || nangles < 1 || nangles > 65535)
            return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalidAngles: %d angles", nangles);
        for (int a = 0; a < nangles; ++a)
            if (!std::isfinite(angles_deg[a]))
                return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "invalidAngles: angle %d is not finite", a);
        if (force_transpose >= 0) { t_first = force_transpose; nt = 1; }
    }
    if (na > 0 && ((t_first == 0 && !curve) || ((t_first == 1 || nt == 2) && !curve_t)))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL curve");
    if (int st = device_ok(device)) return st;
    DeviceGuard g(device);
    Dev d(device);
    CUBO_HIP_TRY(d.init(), "hipStreamCreate");
    double *dnx, *dny, *dcx, *dcy;
    for (double **p : {&dnx, &dny, &dcx, &dcy}) CUBO_HIP_TRY(d.alloc(p, n), "allocating the field");
    CUBO_HIP_TRY(hipMemcpyAsync(dnx, nx, n * 8, hipMemcpyHostToDevice, d.s), "uploading the field");
    CUBO_HIP_TRY(hipMemcpyAsync(dny, ny, n * 8, hipMemcpyHostToDevice, d.s), "uploading the field");
    int t = force_transpose;
    double deg = force_degrees ? *force_degrees : 0.0;
    *n_curve = 0;
    *n_curve_t = 0;
    if (na > 0) {
        std::vector<double> co(na), si(na);
        for (int a = 0; a < na; ++a) {
            const double rad = score_angles[a] * (M_PI / 180.0);   // deg * (pi / 180), RotationSolver.swift:93
            co[a] = std::cos(rad);
            si[a] = std::sin(rad);
        }
        double *dco, *dsi, *partial;
        CUBO_HIP_TRY(d.alloc(&dco, na), "allocating angles");
        CUBO_HIP_TRY(d.alloc(&dsi, na), "allocating angles");
        const size_t np = (size_t)nt * na * kSplit;
        CUBO_HIP_TRY(d.alloc(&partial, np), "allocating partial sums");
        CUBO_HIP_TRY(hipMemcpyAsync(dco, co.data(), na * 8, hipMemcpyHostToDevice, d.s), "uploading angles");
        CUBO_HIP_TRY(hipMemcpyAsync(dsi, si.data(), na * 8, hipMemcpyHostToDevice, d.s), "uploading angles");
        hipLaunchKernelGGL(rotation_sums, dim3(kSplit, na, nt), dim3(kThreads), 0, d.s, dnx, dny, rows, cols, dco, dsi,
                           (int)(maximize != 0), t_first, partial);
        CUBO_HIP_TRY(hipGetLastError(), "launching the rotation curves");
        std::vector<double> h(np), sc((size_t)nt * na);
        CUBO_HIP_TRY(hipMemcpyAsync(h.data(), partial, np * 8, hipMemcpyDeviceToHost, d.s), "reading the curves");
        CUBO_HIP_TRY(hipStreamSynchronize(d.s), "waiting for the rotation curves");
        // parallelScores asks once more after the scores: a cancelled solve returns no curve (ParallelScores.swift:37)
        if (cancel && *cancel) return fail(CUBO_HIP_ERR_CANCELLED, "cancelled");
        for (int z = 0; z < nt; ++z)
            for (int a = 0; a < na; ++a) {
                double s = 0;
                for (int b = 0; b < kSplit; ++b) s += h[((size_t)z * na + a) * kSplit + b];
                sc[(size_t)z * na + a] = s / (double)ni;
            }
        auto better = [&](double a, double b) { return maximize ? a > b : a < b; };
        auto best = [&](const double *c) {
            int k = 0;
            for (int i = 1; i < na; ++i)
                if (better(c[i], c[k])) k = i;
            return k;
        };
        for (int z = 0; z < nt; ++z) {
            double *dst = (t_first + z) ? curve_t : curve;
            std::copy(sc.begin() + (size_t)z * na, sc.begin() + (size_t)(z + 1) * na, dst);
            *((t_first + z) ? n_curve_t : n_curve) = na;
        }
        if (force_degrees) {
            t = better(curve_t[0], curve[0]);                // the plain orientation wins ties
        } else if (nt == 1) {
            t = t_first;
            deg = angles_deg[best(t ? curve_t : curve)];
        } else {
            const int i = best(curve), j = best(curve_t);
            t = better(curve_t[j], curve[i]);
            deg = angles_deg[t ? j : i];
        }
    }
    *transposed = t;
    *degrees = deg;
    const double rad = deg * (M_PI / 180.0);
    hipLaunchKernelGGL(rotate_apply, dim3(blocks(n)), dim3(kThreads), 0, d.s, dnx, dny, n, std::cos(rad), std::sin(rad),
                       t, dcx, dcy);
    CUBO_HIP_TRY(hipGetLastError(), "rotating the field");
    CUBO_HIP_TRY(hipMemcpyAsync(cx, dcx, n * 8, hipMemcpyDeviceToHost, d.s), "reading the field");
    CUBO_HIP_TRY(hipMemcpyAsync(cy, dcy, n * 8, hipMemcpyDeviceToHost, d.s), "reading the field");
    CUBO_HIP_TRY(hipStreamSynchronize(d.s), "waiting for the rotation");
    return CUBO_HIP_OK;
}

}  // extern "C"
