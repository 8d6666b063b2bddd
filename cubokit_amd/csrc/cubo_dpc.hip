// cubo_dpc.hip — C ABI of the iterative-DPC phase reconstruction and its adjoint (include/cubo_hip.h, "DPC").
//
// DPC.reconstruct (Sources/CuboPhase/DPC.swift:26-102): the field is embedded in a grid padded to
// round(R * padding) x round(C * padding) (half-even), and each iteration computes the central differences of
// the phase plus the field, their mean square against the field's, and the update
// Re(IFFT(FFT(dx) i gx + FFT(dy) i gy)) with gx = -k_r / (4 |k|^2) (0 at the origin), halving the step when
// the error grew. On the device: one complex FFT of dx + i dy, the spectrum split by symmetry, one inverse.
// The adjoint of a step (and of a chain of fixed-step iterations from phi = 0) runs the transposed operators:
// one FFT of the incoming phase gradient and one inverse.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/cubo_hip.h"
#include "dpc_kernels.hpp"
#include "errors.hpp"
#include "host_math.hpp"

using cubo::DeviceGuard;
using cubo::fail;
using cubo::hip_fail;

struct cubo_hip_dpc {
    int device;
    hipStream_t stream = nullptr;
    int R, C, Rp, Cp;
    double sr, sc;
    double *kx = nullptr, *ky = nullptr;     // numpy.fft.fftfreq of each side
    double2 *tw_r = nullptr, *tw_c = nullptr;
    cubo::FftPlan *plan_r = nullptr, *plan_c = nullptr;   // grids above kFftMax a side: the general-length plans
    double2 *z = nullptr, *f = nullptr, *tmp = nullptr;
    double *partial = nullptr, *sum = nullptr;
    double *phi = nullptr, *phi2 = nullptr;      // padded grid
    double *fx = nullptr, *fy = nullptr;          // scan region
    double *g_fx = nullptr, *g_fy = nullptr;
    std::vector<double *> chain;                  // stored iterates of a chain (backward)
    double *chain_sums = nullptr;                 // a chain's squared-error sums, one per step
    int chain_sums_cap = 0;
    double *chain_g = nullptr;                    // a chain's dL/dphase (scan region), uploaded before its steps
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double ms = 0;
    int64_t calls = 0;
    double *pin = nullptr;                        // pinned staging of the host arrays (doubles)
    size_t pin_cap = 0;
};

namespace {

void free_dpc(cubo_hip_dpc *d) {
    if (!d) return;
    DeviceGuard g(d->device);
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    void *bufs[] = {d->kx, d->ky, d->tw_r, d->tw_c, d->z, d->f, d->tmp, d->partial, d->sum, d->phi, d->phi2,
                    d->fx, d->fy, d->g_fx, d->g_fy};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (double *b : d->chain) (void)hipFree(b);
    cubo::fft_plan_destroy(d->plan_r);
    cubo::fft_plan_destroy(d->plan_c);
    if (d->chain_sums) (void)hipFree(d->chain_sums);
    if (d->chain_g) (void)hipFree(d->chain_g);
    if (d->t0) (void)hipEventDestroy(d->t0);
    if (d->t1) (void)hipEventDestroy(d->t1);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    if (d->pin) (void)hipHostFree(d->pin);
    delete d;
}

// Host arrays go to and from the device through a pinned staging buffer of the plan, copied on host threads (a DMA
// from or to pageable memory runs at a fraction of the link's rate). A call whose arrays would need more than
// kStageMax doubles, or whose staging cannot be allocated, copies them directly.
constexpr size_t kStageMax = (size_t)32 << 20;
double *stage(cubo_hip_dpc *d, size_t doubles) {
    if (doubles > kStageMax) return nullptr;
    if (d->pin_cap < doubles) {
        if (d->pin) (void)hipHostFree(d->pin);
        d->pin = nullptr;
        d->pin_cap = 0;
        if (hipHostMalloc((void **)&d->pin, doubles * 8, hipHostMallocDefault) != hipSuccess) {
            d->pin = nullptr;
            (void)hipGetLastError();
            return nullptr;
        }
        d->pin_cap = doubles;
    }
    return d->pin;
}

// count doubles on host threads (about 1 MiB each)
void host_copy(double *dst, const double *src, size_t count) {
    const int threads = (int)std::min<size_t>((size_t)cubo::reader_threads(), std::max<size_t>(1, (count * 8) >> 20));
    cubo::parallel_pieces(count * 8, threads, [&](size_t off, size_t len) {
        memcpy((char *)dst + off, (const char *)src + off, len);
        return true;
    });
}

// host -> dev, through the staging region at (the caller keeps the region until the copy has run) or directly
hipError_t upload(cubo_hip_dpc *d, double *dev, const double *host, size_t count, double *at) {
    if (!at) return hipMemcpyAsync(dev, host, count * 8, hipMemcpyHostToDevice, d->stream);
    host_copy(at, host, count);
    return hipMemcpyAsync(dev, at, count * 8, hipMemcpyHostToDevice, d->stream);
}

// dev -> host: the copy into the staging region at (or directly into host), completed by finish_download after
// the stream has been synchronized
hipError_t start_download(cubo_hip_dpc *d, double *host, const double *dev, size_t count, double *at) {
    return hipMemcpyAsync(at ? at : host, dev, count * 8, hipMemcpyDeviceToHost, d->stream);
}
void finish_download(double *host, const double *at, size_t count) {
    if (at) host_copy(host, at, count);
}

// numpy.fft.fftfreq(n, d)
void fftfreq(int n, double d, double *out) {
    const double scale = 1.0 / ((double)n * d);
    const int positive = (n - 1) / 2 + 1;
    for (int k = 0; k < n; ++k) out[k] = (double)(k < positive ? k : k - n) * scale;
}

cubo::DpcGrid grid(const cubo_hip_dpc *d) {
    cubo::DpcGrid g{};
    g.R = d->R; g.C = d->C; g.Rp = d->Rp; g.Cp = d->Cp;
    g.sr = d->sr; g.sc = d->sc;
    g.kx = d->kx; g.ky = d->ky;
    return g;
}

// Field checks and the denominator mean(fx^2 + fy^2) (NumPy's pairwise order, as DPC.swift:49-54).
// The field checked (finite, first offending entry reported), its denominator mean(fx^2 + fy^2) summed pairwise in
// NumPy's order, and fx, fy uploaded (staged in pin[0 .. 2N) when pin is given); on host threads.
int field(cubo_hip_dpc *d, const double *fx, const double *fy, double *denom, double *pin = nullptr) {
    const int64_t N = (int64_t)d->R * d->C;
    std::vector<double> t(N);
    std::atomic<int64_t> bad{N};
    const int threads = (int)std::min<int64_t>(cubo::reader_threads(), std::max<int64_t>(1, (N * 8) >> 20));
    cubo::parallel_pieces((size_t)N * 8, threads, [&](size_t off, size_t len) {
        const int64_t k0 = (int64_t)(off / 8), k1 = k0 + (int64_t)(len / 8);
        for (int64_t k = k0; k < k1; ++k) {
            if (!std::isfinite(fx[k]) || !std::isfinite(fy[k])) {
                int64_t b = bad.load();
                while (k < b && !bad.compare_exchange_weak(b, k)) {}
                return true;
            }
            t[k] = fx[k] * fx[k] + fy[k] * fy[k];
        }
        if (pin) {
            memcpy(pin + k0, fx + k0, len);
            memcpy(pin + N + k0, fy + k0, len);
        }
        return true;
    });
    if (bad.load() < N)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the field must be finite (entry %lld)", (long long)bad.load());
    *denom = cubo::pairwise_sum(t.data(), N) / (double)N;
    if (!(*denom > 0)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the field is zero: DPC has nothing to reconstruct");
    CUBO_HIP_TRY(hipMemcpyAsync(d->fx, pin ? pin : fx, N * 8, hipMemcpyHostToDevice, d->stream), "uploading the field");
    CUBO_HIP_TRY(hipMemcpyAsync(d->fy, pin ? pin + N : fy, N * 8, hipMemcpyHostToDevice, d->stream),
                 "uploading the field");
    return CUBO_HIP_OK;
}

// The differences of phi into z, and their squared-error sum into *sum (default: d->sum).
// On grids the fused passes take, z is not stored: its row transforms go to tmp (the transform's first pass),
// which update() completes.
bool fused(const cubo_hip_dpc *d) { return !d->plan_r && cubo::dpc_fused_supported(d->Rp, d->Cp); }

hipError_t forward(cubo_hip_dpc *d, const double *phi, double *sum = nullptr) {
    int nb = 0;
    hipError_t e = fused(d) ? cubo::launch_dpc_forward_pass(grid(d), phi, d->fx, d->fy, d->tw_c, d->tmp, d->partial, &nb,
                                                            d->stream)
                            : cubo::launch_dpc_forward(grid(d), phi, d->fx, d->fy, d->z, d->partial, &nb, d->stream);
    if (e == hipSuccess) e = cubo::launch_sum_partials(d->partial, nb, sum ? sum : d->sum, d->stream);
    return e;
}

// The 2-D FFT of the padded grid: the LDS transforms when both sides are at most kFftMax, else the plans.
hipError_t fft2(cubo_hip_dpc *d, double2 *data, bool inverse) {
    if (d->plan_r) return cubo::launch_fft2_planned(data, d->tmp, d->Rp, d->Cp, d->plan_r, d->plan_c, inverse, d->stream);
    return cubo::launch_fft2(data, d->tmp, d->Rp, d->Cp, d->tw_r, d->tw_c, inverse, d->stream);
}

// phi_out = phi + step Re(IFFT(i gx FFT(dx) + i gy FFT(dy))), with z = dx + i dy from forward() (on the fused grids,
// the first pass of its transform, in tmp).
hipError_t update(cubo_hip_dpc *d, const double *phi, double step, double *phi_out) {
    if (fused(d)) {
        hipError_t e = cubo::launch_fft2_second_pass(d->tmp, d->z, d->Rp, d->Cp, d->tw_r, d->stream);
        if (e == hipSuccess)
            e = cubo::launch_ifft2_spectrum_update(grid(d), d->z, d->tmp, d->tw_r, d->tw_c, phi, step, phi_out, d->stream);
        return e;
    }
    hipError_t e = fft2(d, d->z, false);
    if (e == hipSuccess) e = cubo::launch_dpc_spectrum(grid(d), d->z, d->f, d->stream);
    if (e == hipSuccess) e = fft2(d, d->f, true);
    if (e == hipSuccess) e = cubo::launch_dpc_update(grid(d), phi, d->f, step, phi_out, d->stream);
    return e;
}

// One step's vector-Jacobian product: g_phi_in (device, padded) -> g_phi_out (device), g_fx/g_fy += (device).
hipError_t vjp(cubo_hip_dpc *d, const double *phi, double step, const double *g_in, double g_err, double denom,
               double msq, double *g_out) {
    const int64_t n = (int64_t)d->Rp * d->Cp;
    hipError_t e = fused(d) ? cubo::launch_fft2_real(g_in, d->tmp, d->z, d->Rp, d->Cp, d->tw_r, d->tw_c, d->stream)
                            : cubo::launch_real_to_complex(g_in, d->z, n, d->stream);
    if (fused(d)) {
        if (e == hipSuccess)
            e = cubo::launch_ifft2_adjoint_spectrum(grid(d), d->z, d->tmp, d->f, d->tw_r, d->tw_c, d->stream);
    } else {
        if (e == hipSuccess) e = fft2(d, d->z, false);
        if (e == hipSuccess) e = cubo::launch_dpc_adjoint_spectrum(grid(d), d->z, d->f, d->stream);
        if (e == hipSuccess) e = fft2(d, d->f, true);
    }
    // error = msq / denom with denom = mean(fx^2 + fy^2): the denominator's share of the field gradient
    const double dd = -g_err * msq / (denom * denom);
    if (e == hipSuccess)
        e = cubo::launch_dpc_vjp(grid(d), phi, d->fx, d->fy, d->f, step, g_err, denom, dd, g_in, g_out, d->g_fx,
                                 d->g_fy, d->stream);
    return e;
}

double read_sum(cubo_hip_dpc *d, hipError_t *e) {
    double s = 0;
    *e = hipMemcpyAsync(&s, d->sum, 8, hipMemcpyDeviceToHost, d->stream);
    if (*e == hipSuccess) *e = hipStreamSynchronize(d->stream);
    return s;
}

}  // namespace

extern "C" {

int cubo_hip_dpc_create(int device, int32_t rows, int32_t cols, double padding, double sampling_row,
                        double sampling_col, cubo_hip_dpc **out) {
    if (!out) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (rows < 1 || cols < 1) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "empty scan %d x %d", rows, cols);
    if (!(sampling_row > 0) || !(sampling_col > 0) || !std::isfinite(sampling_row) || !std::isfinite(sampling_col))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "the sampling must be positive and finite");
    if (!(padding >= 1) || !std::isfinite(padding)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "padding below 1");
    const int Rp = (int)std::nearbyint((double)rows * padding), Cp = (int)std::nearbyint((double)cols * padding);
    if ((int64_t)Rp * Cp > ((int64_t)1 << 28))          // 4 GiB per complex grid buffer
        return fail(CUBO_HIP_ERR_UNSUPPORTED, "padded grid %d x %d: at most 2^28 points", Rp, Cp);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return fail(CUBO_HIP_ERR_NO_DEVICE, "no HIP device %d", device);
    auto *d = new (std::nothrow) cubo_hip_dpc();
    if (!d) return fail(CUBO_HIP_ERR_OUT_OF_MEMORY, "host allocation");
    d->device = device;
    d->R = rows; d->C = cols; d->Rp = Rp; d->Cp = Cp;
    d->sr = sampling_row; d->sc = sampling_col;
    DeviceGuard g(device);
    const size_t n = (size_t)Rp * Cp, N = (size_t)rows * cols;
    // operators (DPC.swift:40-47): k from numpy.fft.fftfreq, kden = 1 / (kx^2 + ky^2) with infinity at 0,
    // gx = -0.25 kx kden, gy = -0.25 ky kden: formed in the kernels from kx and ky
    std::vector<double> kx(Rp), ky(Cp);
    fftfreq(Rp, sampling_row, kx.data());
    fftfreq(Cp, sampling_col, ky.data());
    const bool small = Rp <= cubo::kFftMax && Cp <= cubo::kFftMax;
    const std::vector<double2> twr = cubo::fft_twiddles(small ? Rp : 1), twc = cubo::fft_twiddles(small ? Cp : 1);
    hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    struct { void **p; size_t bytes; } a[] = {
        {(void **)&d->kx, (size_t)Rp * 8}, {(void **)&d->ky, (size_t)Cp * 8}, {(void **)&d->tw_r, twr.size() * 16}, {(void **)&d->tw_c, twc.size() * 16},
        {(void **)&d->z, n * 16}, {(void **)&d->f, n * 16}, {(void **)&d->tmp, n * 16},
        {(void **)&d->partial, ((n + 255) / 256) * 8}, {(void **)&d->sum, 8}, {(void **)&d->phi, n * 8},
        {(void **)&d->phi2, n * 8}, {(void **)&d->fx, N * 8}, {(void **)&d->fy, N * 8}, {(void **)&d->g_fx, N * 8},
        {(void **)&d->g_fy, N * 8}};
    for (auto &x : a)
        if (e == hipSuccess) e = hipMalloc(x.p, x.bytes);
    if (e == hipSuccess) e = hipMemcpy(d->kx, kx.data(), (size_t)Rp * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->ky, ky.data(), (size_t)Cp * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->tw_r, twr.data(), twr.size() * 16, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->tw_c, twc.data(), twc.size() * 16, hipMemcpyHostToDevice);
    // larger grids (DPC.swift:42-44 sets no limit): four-step and Bluestein plans of each side
    if (e == hipSuccess && !small) e = cubo::fft_plan_create(Rp, Cp, d->stream, &d->plan_r);
    if (e == hipSuccess && !small) e = cubo::fft_plan_create(Cp, Rp, d->stream, &d->plan_c);
    if (e == hipSuccess) e = hipEventCreate(&d->t0);
    if (e == hipSuccess) e = hipEventCreate(&d->t1);
    if (e != hipSuccess) { free_dpc(d); return hip_fail(e, "preparing DPC"); }
    *out = d;
    return CUBO_HIP_OK;
}

int cubo_hip_dpc_padded_shape(const cubo_hip_dpc *d, int32_t *rows, int32_t *cols) {
    if (!d || !rows || !cols) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *rows = d->Rp;
    *cols = d->Cp;
    return CUBO_HIP_OK;
}

int cubo_hip_dpc_step(cubo_hip_dpc *d, const double *phi, const double *fx, const double *fy, double step,
                      double *phi_out, double *error) {
    if (!d || !phi || !fx || !fy || !phi_out || !error) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceGuard g(d->device);
    double denom = 0;
    if (int st = field(d, fx, fy, &denom)) return st;
    const int64_t n = (int64_t)d->Rp * d->Cp;
    CUBO_HIP_TRY(hipMemcpyAsync(d->phi, phi, n * 8, hipMemcpyHostToDevice, d->stream), "uploading the phase");
    CUBO_HIP_TRY(hipEventRecord(d->t0, d->stream), "hipEventRecord");
    CUBO_HIP_TRY(forward(d, d->phi), "launching the DPC differences");
    CUBO_HIP_TRY(update(d, d->phi, step, d->phi2), "launching the DPC update");
    CUBO_HIP_TRY(hipEventRecord(d->t1, d->stream), "hipEventRecord");
    CUBO_HIP_TRY(hipMemcpyAsync(phi_out, d->phi2, n * 8, hipMemcpyDeviceToHost, d->stream), "reading the phase");
    hipError_t e;
    const double s = read_sum(d, &e);
    if (e != hipSuccess) return hip_fail(e, "waiting for the DPC step");
    *error = s / (double)((int64_t)d->R * d->C) / denom;
    float ms = 0;
    if (hipEventElapsedTime(&ms, d->t0, d->t1) == hipSuccess) d->ms += ms;
    ++d->calls;
    return CUBO_HIP_OK;
}

int cubo_hip_dpc_step_backward(cubo_hip_dpc *d, const double *phi, const double *fx, const double *fy, double step,
                               const double *g_phi_out, double g_error, double *g_phi, double *g_fx, double *g_fy) {
    if (!d || !phi || !fx || !fy || !g_phi_out || !g_phi || !g_fx || !g_fy)
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceGuard g(d->device);
    double denom = 0;
    if (int st = field(d, fx, fy, &denom)) return st;
    const int64_t n = (int64_t)d->Rp * d->Cp, N = (int64_t)d->R * d->C;
    CUBO_HIP_TRY(hipMemcpyAsync(d->phi, phi, n * 8, hipMemcpyHostToDevice, d->stream), "uploading the phase");
    CUBO_HIP_TRY(hipMemcpyAsync(d->phi2, g_phi_out, n * 8, hipMemcpyHostToDevice, d->stream), "uploading dL/dphi");
    // the buffers are cleared on the device (all-zero bytes are +0.0), with no host copy
    CUBO_HIP_TRY(hipMemsetAsync(d->g_fx, 0, N * 8, d->stream), "clearing");
    CUBO_HIP_TRY(hipMemsetAsync(d->g_fy, 0, N * 8, d->stream), "clearing");
    CUBO_HIP_TRY(forward(d, d->phi), "launching the DPC differences");
    hipError_t e;
    const double msq = read_sum(d, &e) / (double)N;
    if (e != hipSuccess) return hip_fail(e, "waiting for the DPC differences");
    double *gout = nullptr;
    CUBO_HIP_TRY(hipMalloc((void **)&gout, n * 8), "allocating dL/dphi");
    e = vjp(d, d->phi, step, d->phi2, g_error, denom, msq, gout);
    if (e == hipSuccess) e = hipMemcpyAsync(g_phi, gout, n * 8, hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(g_fx, d->g_fx, N * 8, hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(g_fy, d->g_fy, N * 8, hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)hipFree(gout);
    if (e != hipSuccess) return hip_fail(e, "the DPC step adjoint");
    return CUBO_HIP_OK;
}

int cubo_hip_dpc_reconstruct(cubo_hip_dpc *d, const double *fx, const double *fy, int32_t max_iter, double step,
                             double stop, int32_t backtrack, double *phase, double *errors, int32_t *accepted,
                             int32_t *n_iter, double *final_step) {
    if (!d || !fx || !fy || !phase || !n_iter || !final_step) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (max_iter < 0 || !(step > 0) || !std::isfinite(step))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "max_iter must be >= 0 and the step positive");
    if (max_iter > 0 && (!errors || !accepted)) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceGuard g(d->device);
    double denom = 0;
    const int64_t n = (int64_t)d->Rp * d->Cp, N = (int64_t)d->R * d->C;
    double *pin = stage(d, 3 * (size_t)N);            // staging: fx, fy, the phase
    if (int st = field(d, fx, fy, &denom, pin)) return st;
    CUBO_HIP_TRY(hipMemsetAsync(d->phi, 0, n * 8, d->stream), "clearing the phase");
    double error = INFINITY;
    int ne = 0, ni = 0;
    double *cur = d->phi, *nxt = d->phi2;
    CUBO_HIP_TRY(hipEventRecord(d->t0, d->stream), "hipEventRecord");
    for (int it = 0; it < max_iter; ++it) {
        if (step < stop) break;
        CUBO_HIP_TRY(forward(d, cur), "launching the DPC differences");
        hipError_t e;
        const double new_error = read_sum(d, &e) / (double)N / denom;
        if (e != hipSuccess) return hip_fail(e, "waiting for the DPC differences");
        if (new_error > error && backtrack) {               // DPC.swift:83-90
            step /= 2;
            accepted[ni++] = 0;
            continue;
        }
        error = new_error;
        CUBO_HIP_TRY(update(d, cur, step, nxt), "launching the DPC update");
        std::swap(cur, nxt);
        errors[ne++] = error;
        accepted[ni++] = 1;
    }
    CUBO_HIP_TRY(hipEventRecord(d->t1, d->stream), "hipEventRecord");
    CUBO_HIP_TRY(cubo::launch_crop(cur, d->R, d->C, d->Cp, d->g_fx, d->stream), "launching the crop");
    double *at = pin ? pin + 2 * N : nullptr;
    CUBO_HIP_TRY(start_download(d, phase, d->g_fx, N, at), "reading the phase");
    CUBO_HIP_TRY(hipStreamSynchronize(d->stream), "waiting for DPC");
    finish_download(phase, at, N);
    float ms = 0;
    if (hipEventElapsedTime(&ms, d->t0, d->t1) == hipSuccess) d->ms += ms;
    ++d->calls;
    *n_iter = ni;
    *final_step = step;
    return CUBO_HIP_OK;
}

int cubo_hip_dpc_chain_backward(cubo_hip_dpc *d, const double *fx, const double *fy, int32_t nsteps, double step,
                                const double *g_phase, const double *g_errors, double *phase, double *errors,
                                double *g_fx, double *g_fy) {
    if (!d || !fx || !fy || !g_phase || !g_fx || !g_fy) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nsteps < 1 || !(step > 0) || !std::isfinite(step))
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "at least one step, with a positive step size");
    DeviceGuard g(d->device);
    double denom = 0;
    const int64_t n = (int64_t)d->Rp * d->Cp, N = (int64_t)d->R * d->C;
    // staging: fx, fy, dL/dphase, the phase, dL/dfx, dL/dfy
    double *pin = stage(d, 6 * (size_t)N);
    auto at = [&](int i) { return pin ? pin + (size_t)i * N : nullptr; };
    if (int st = field(d, fx, fy, &denom, pin)) return st;
    while ((int)d->chain.size() < nsteps + 1) {
        double *b = nullptr;
        CUBO_HIP_TRY(hipMalloc((void **)&b, n * 8), "allocating the stored iterates");
        d->chain.push_back(b);
    }
    if (d->chain_sums_cap < nsteps) {
        if (d->chain_sums) CUBO_HIP_TRY(hipFree(d->chain_sums), "freeing the error sums");
        d->chain_sums = nullptr;
        d->chain_sums_cap = 0;
        CUBO_HIP_TRY(hipMalloc((void **)&d->chain_sums, (size_t)nsteps * 8), "allocating the error sums");
        d->chain_sums_cap = nsteps;
    }
    if (!d->chain_g) CUBO_HIP_TRY(hipMalloc((void **)&d->chain_g, N * 8), "allocating dL/dphase");
    std::vector<double> msq(nsteps);
    // dL/dphase goes up before the steps (its copy from pageable memory stays out of the chain's timing)
    CUBO_HIP_TRY(upload(d, d->chain_g, g_phase, N, at(2)), "uploading dL/dphase");
    CUBO_HIP_TRY(hipMemsetAsync(d->chain[0], 0, n * 8, d->stream), "clearing");
    CUBO_HIP_TRY(hipEventRecord(d->t0, d->stream), "hipEventRecord");
    // forward: phi_{i+1} = phi_i + step * update(phi_i), keeping every iterate and every step's squared-error
    // sum on the device (the fixed steps need no decision on the host, so the steps run back to back)
    for (int i = 0; i < nsteps; ++i) {
        CUBO_HIP_TRY(forward(d, d->chain[i], d->chain_sums + i), "launching the DPC differences");
        CUBO_HIP_TRY(update(d, d->chain[i], step, d->chain[i + 1]), "launching the DPC update");
    }
    CUBO_HIP_TRY(hipMemcpyAsync(msq.data(), d->chain_sums, (size_t)nsteps * 8, hipMemcpyDeviceToHost, d->stream),
                 "reading the error sums");
    CUBO_HIP_TRY(hipStreamSynchronize(d->stream), "waiting for the DPC steps");
    for (int i = 0; i < nsteps; ++i) {
        msq[i] /= (double)N;
        if (errors) errors[i] = msq[i] / denom;
    }
    if (phase) {
        CUBO_HIP_TRY(cubo::launch_crop(d->chain[nsteps], d->R, d->C, d->Cp, d->g_fx, d->stream), "launching the crop");
        CUBO_HIP_TRY(start_download(d, phase, d->g_fx, N, at(3)), "reading the phase");
        CUBO_HIP_TRY(hipStreamSynchronize(d->stream), "waiting for DPC");
        finish_download(phase, at(3), N);
    }
    // backward: dL/dphase of the cropped phase into the padded grid, then each step's adjoint in reverse
    CUBO_HIP_TRY(hipMemsetAsync(d->g_fx, 0, N * 8, d->stream), "clearing");
    CUBO_HIP_TRY(hipMemsetAsync(d->g_fy, 0, N * 8, d->stream), "clearing");
    CUBO_HIP_TRY(cubo::launch_embed(d->chain_g, d->R, d->C, d->Rp, d->Cp, d->phi, d->stream), "launching the embed");
    double *gin = d->phi, *gout = d->phi2;
    for (int i = nsteps - 1; i >= 0; --i) {
        CUBO_HIP_TRY(vjp(d, d->chain[i], step, gin, g_errors ? g_errors[i] : 0.0, denom, msq[i], gout),
                     "launching the DPC step adjoint");
        std::swap(gin, gout);
    }
    CUBO_HIP_TRY(hipEventRecord(d->t1, d->stream), "hipEventRecord");
    CUBO_HIP_TRY(start_download(d, g_fx, d->g_fx, N, at(4)), "reading dL/dfx");
    CUBO_HIP_TRY(start_download(d, g_fy, d->g_fy, N, at(5)), "reading dL/dfy");
    CUBO_HIP_TRY(hipStreamSynchronize(d->stream), "waiting for the DPC adjoint");
    finish_download(g_fx, at(4), N);
    finish_download(g_fy, at(5), N);
    float ms = 0;
    if (hipEventElapsedTime(&ms, d->t0, d->t1) == hipSuccess) d->ms += ms;
    ++d->calls;
    return CUBO_HIP_OK;
}

int cubo_hip_dpc_time(cubo_hip_dpc *d, double *ms, int64_t *calls) {
    if (!d || !ms || !calls) return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *ms = d->ms;
    *calls = d->calls;
    d->ms = 0;
    d->calls = 0;
    return CUBO_HIP_OK;
}

int cubo_hip_dpc_destroy(cubo_hip_dpc *d) {
    free_dpc(d);
    return CUBO_HIP_OK;
}

}  // extern "C"
