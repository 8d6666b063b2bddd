// synth.hip — seeded synthetic 4D-STEM cubes generated on the device (benchmarks and tests).
//
// Expected counts per pixel: a disk probe (sigmoid edge) at a per-position deflected centre, Gaussian
// rings around the same centre, and a flat background; then Poisson counts drawn from a counter-based
// hash of (seed, element index), so a cube is the same for any launch shape and never needs a host copy.
// Poisson sampling: Knuth's multiplication method below a mean of 12, Hormann's PTRS transformed
// rejection (W. Hormann, "The transformed rejection method for generating Poisson random variables",
// Insurance: Mathematics and Economics 12, 1993) at and above it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/cubo_hip.h"
#include "cube.hpp"


namespace {

struct SynthDev {
    uint64_t seed;
    double cr, cc, radius, edge, counts, background;
    int n_rings;
    double rq[4], ra[4], rs[4];
    int n_waves;
    double wra[8], wca[8], wfr[8], wfc[8], wph[8];
    int poisson;
    int64_t scan_cols, qy, qx, ndet;
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

struct Rng {
    uint64_t key, ctr;
    __device__ double uniform() {   // (0, 1)
        const uint64_t r = splitmix64(key ^ splitmix64(ctr++));
        return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
};

__device__ double poisson(double lam, Rng &g) {
    if (!(lam > 0)) return 0;
    if (lam < 12) {
        const double L = exp(-lam);
        double p = 1;
        int k = 0;
        do { ++k; p *= g.uniform(); } while (p > L && k < 1000);
        return k - 1;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
    for (int it = 0; it < 1000; ++it) {
        const double U = g.uniform() - 0.5, V = g.uniform();
        const double us = 0.5 - fabs(U);
        const double k = floor((2 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1)) return k;
    }
    return floor(lam);
}

template <typename T>
__global__ void synth_kernel(T *out, int64_t first, int64_t count, SynthDev s) {
    const int64_t i = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= first + count) return;
    const int64_t pos = i / s.ndet, px = i % s.ndet;
    const double x = (double)(pos / s.scan_cols), y = (double)(pos % s.scan_cols);
    const double di = (double)(px / s.qx), dj = (double)(px % s.qx);
    double dr = 0, dc = 0;
    for (int k = 0; k < s.n_waves; ++k) {
        const double ph = 2 * M_PI * (s.wfr[k] * x + s.wfc[k] * y) + s.wph[k];
        dr += s.wra[k] * sin(ph);
        dc += s.wca[k] * cos(ph);
    }
    const double u = di - (s.cr + dr), v = dj - (s.cc + dc);
    const double d = sqrt(u * u + v * v);
    double lam = s.background;
    if (s.edge > 0) lam += s.counts / (1 + exp((d - s.radius) / s.edge));
    else if (d < s.radius) lam += s.counts;
    for (int k = 0; k < s.n_rings; ++k) {
        const double t = (d - s.rq[k]) / s.rs[k];
        lam += s.ra[k] * exp(-0.5 * t * t);
    }
    double val = lam;
    if (s.poisson) {
        Rng g{splitmix64(s.seed), (uint64_t)i * 64};
        val = poisson(lam, g);
    }
    if constexpr (sizeof(T) == 2) out[i - first] = (T)fmin(val, 65535.0);
    else out[i - first] = (T)val;
}

}  // namespace

extern "C" int cubo_hip_cube_fill_synthetic(cubo_hip_cube *c, const cubo_hip_synth_params *p) {
    if (!c || !p) return CUBO_HIP_ERR_INVALID_ARGUMENT;
    if (int st = cubo::require_resident(c, "a synthetic fill")) return st;
    if (p->n_rings < 0 || p->n_rings > 4 || p->n_waves < 0 || p->n_waves > 8) return CUBO_HIP_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < p->n_rings; ++k)
        if (!(p->ring_sigma[k] > 0)) return CUBO_HIP_ERR_INVALID_ARGUMENT;
    SynthDev s{};
    s.seed = p->seed;
    s.cr = p->center_row; s.cc = p->center_column; s.radius = p->radius; s.edge = p->edge;
    s.counts = p->counts; s.background = p->background;
    s.n_rings = p->n_rings;
    for (int k = 0; k < 4; ++k) { s.rq[k] = p->ring_q[k]; s.ra[k] = p->ring_amplitude[k]; s.rs[k] = p->ring_sigma[k]; }
    s.n_waves = p->n_waves;
    for (int k = 0; k < 8; ++k) {
        s.wra[k] = p->wave_row_amp[k]; s.wca[k] = p->wave_col_amp[k];
        s.wfr[k] = p->wave_fr[k]; s.wfc[k] = p->wave_fc[k]; s.wph[k] = p->wave_phase[k];
    }
    s.poisson = p->poisson;
    s.scan_cols = c->shape[1]; s.qy = c->shape[2]; s.qx = c->shape[3]; s.ndet = c->ndet;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(c->device);
    ++c->generation;
    const int64_t total = c->npos * c->ndet;
    const int64_t step = (int64_t)1 << 30;     // elements per launch keeps the grid within 2^31 threads
    hipError_t e = hipSuccess;
    for (int64_t first = 0; first < total && e == hipSuccess; first += step) {
        const int64_t n = total - first < step ? total - first : step;
        const unsigned blocks = (unsigned)((n + 255) / 256);
        if (c->dtype == CUBO_HIP_U16)
            hipLaunchKernelGGL(synth_kernel<uint16_t>, dim3(blocks), dim3(256), 0, c->stream,
                               (uint16_t *)c->data + first, first, n, s);
        else
            hipLaunchKernelGGL(synth_kernel<float>, dim3(blocks), dim3(256), 0, c->stream, (float *)c->data + first,
                               first, n, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipSetDevice(prev);
    return e == hipSuccess ? CUBO_HIP_OK : CUBO_HIP_ERR_HIP;
}
