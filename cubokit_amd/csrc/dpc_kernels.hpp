// dpc_kernels.hpp — FP64 FFT and the iterative-DPC phase step on the device (dpc_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <vector>

namespace cubo {

// Largest FFT length per dimension (one row of complex doubles, twice, in LDS).
constexpr int kFftMax = 2048;

// In-place 2-D complex FFT of a rows x cols array of (re, im) pairs, NumPy's convention (forward
// unnormalised, inverse divided by rows * cols), rows then columns. tw_r / tw_c: exp(-2 pi i m / n) for
// m < n of each length; tmp: rows * cols pairs.
hipError_t launch_fft2(double2 *data, double2 *tmp, int rows, int cols, const double2 *tw_r, const double2 *tw_c,
                       bool inverse, hipStream_t s);

// exp(-2 pi i m / n) for m < n: the twiddle table of the FFT kernels, in Double on the host.
inline std::vector<double2> fft_twiddles(int n) {
    std::vector<double2> t(n);
    for (int m = 0; m < n; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)n;
        t[m] = make_double2(std::cos(a), std::sin(a));
    }
    return t;
}

// The same for `batch` arrays of rows x cols stored one after another (batch <= 65535).
hipError_t launch_fft2_batched(double2 *data, double2 *tmp, int batch, int rows, int cols, const double2 *tw_r,
                               const double2 *tw_c, bool inverse, hipStream_t s);

// A 1-D FFT of any length n >= 1 for batches of up to max_batch contiguous rows, unnormalised in both directions:
//  - n <= kFftMax: fft_rows (Stockham stages in LDS for powers of two, the direct transform from the twiddle table
//    otherwise);
//  - larger powers of two: four steps, n = n1 n2 with n1, n2 <= kFftMax powers of two: a transpose of each row seen
//    as n1 x n2, FFTs of length n1, the twiddles W_n^(j2 k1), a transpose, FFTs of length n2, a transpose;
//  - other lengths: Bluestein's chirp convolution through a power-of-two plan of length M >= 2n - 1 (the chirps
//    exp(-+i pi j^2 / n) from j^2 mod 2n, the transform of the convolution kernel made once per plan).
struct FftPlan;
hipError_t fft_plan_create(int n, int max_batch, hipStream_t s, FftPlan **out);
void fft_plan_destroy(FftPlan *p);
hipError_t launch_fft_rows(FftPlan *p, double2 *data, int batch, bool inverse, hipStream_t s);
// In-place 2-D FFT with NumPy's convention (rows then columns, inverse divided by rows * cols) through plans of
// the row length (pc, batches of `rows`) and the column length (pr, batches of `cols`); tmp: rows * cols pairs.
hipError_t launch_fft2_planned(double2 *data, double2 *tmp, int rows, int cols, FftPlan *pr, FftPlan *pc,
                               bool inverse, hipStream_t s);

struct DpcGrid {
    int32_t R, C, Rp, Cp;     // scan region and padded grid
    double sr, sc;            // sampling
    // [Rp], [Cp]: numpy.fft.fftfreq of each side. The kernels form the operators gx = -0.25 kx / (kx^2 + ky^2),
    // gy = -0.25 ky / (kx^2 + ky^2) (0 at the origin; the update operator is i g) from them, in the host's order.
    const double *kx, *ky;
};

// z = dx + i dy on the padded grid (zero outside the scan region; DPC.swift:59-69) and sq = dx^2 + dy^2
// on the scan region; partial[b] = this block's sum of sq (blocks of 256 scan entries).
hipError_t launch_dpc_forward(const DpcGrid &g, const double *phi, const double *fx, const double *fy, double2 *z,
                              double *partial, int *nblocks, hipStream_t s);
// spectrum of the update: F = FFT(dx + i dy) -> i (gx FFT(dx) + gy FFT(dy)) (DPC.swift:71-79), in out.
hipError_t launch_dpc_spectrum(const DpcGrid &g, const double2 *f, double2 *out, hipStream_t s);
// phi_out = phi + step * Re(u).
hipError_t launch_dpc_update(const DpcGrid &g, const double *phi, const double2 *u, double step, double *phi_out,
                             hipStream_t s);
// The step with its element-wise kernels fused into the passes of the 2-D FFT (padded grids whose sides are powers
// of two up to kFftMax, with R Cp a multiple of 256 for the R rows a pass block holds); the same values:
//   - launch_dpc_forward_pass: launch_dpc_forward and the first pass of the forward transform in one kernel: the
//     row transforms of z, stored transposed in tmp, and the partial sums;
//   - launch_fft2_second_pass: the second pass, tmp -> data (the transform of z);
//   - launch_fft2_real: the forward transform of a real array in (real_to_complex and launch_fft2);
//   - launch_ifft2_spectrum_update: launch_dpc_spectrum of F, its inverse transform and launch_dpc_update (the
//     spectrum formed as the inverse's first pass loads its rows, u itself not stored);
//   - launch_ifft2_adjoint_spectrum: launch_dpc_adjoint_spectrum of G and its inverse transform, into out.
bool dpc_fused_supported(int Rp, int Cp);
hipError_t launch_dpc_forward_pass(const DpcGrid &g, const double *phi, const double *fx, const double *fy,
                                   const double2 *tw_c, double2 *tmp, double *partial, int *nblocks, hipStream_t s);
hipError_t launch_fft2_second_pass(const double2 *tmp, double2 *data, int rows, int cols, const double2 *tw_r,
                                   hipStream_t s);
hipError_t launch_fft2_real(const double *in, double2 *tmp, double2 *data, int rows, int cols, const double2 *tw_r,
                            const double2 *tw_c, hipStream_t s);
hipError_t launch_ifft2_spectrum_update(const DpcGrid &g, const double2 *F, double2 *tmp, const double2 *tw_r,
                                        const double2 *tw_c, const double *phi, double step, double *phi_out,
                                        hipStream_t s);
hipError_t launch_ifft2_adjoint_spectrum(const DpcGrid &g, const double2 *G, double2 *tmp, double2 *out,
                                         const double2 *tw_r, const double2 *tw_c, hipStream_t s);
// z = g_phi_out + 0 i (the adjoint's input), padded grid.
hipError_t launch_real_to_complex(const double *in, double2 *z, int64_t n, hipStream_t s);
// spectrum of the adjoint: G = FFT(g) -> G (gy - i gx), whose inverse is A_x^T g + i A_y^T g.
hipError_t launch_dpc_adjoint_spectrum(const DpcGrid &g, const double2 *G, double2 *out, hipStream_t s);
// The rest of the step's vector-Jacobian product (gdx = step Re(w) + 2 g_err dx / (N denom) on the scan region,
// gdy likewise from Im(w)); g_phi = g_phi_out + the stencil adjoint; g_fx += gdx + dd 2 fx / N, g_fy likewise.
hipError_t launch_dpc_vjp(const DpcGrid &g, const double *phi, const double *fx, const double *fy, const double2 *w,
                          double step, double g_err, double denom, double dd, const double *g_phi_out, double *g_phi,
                          double *g_fx, double *g_fy, hipStream_t s);
// sum of partial[0 .. n) into out[0] (one block, fixed order).
hipError_t launch_sum_partials(const double *partial, int n, double *out, hipStream_t s);
// out = in[0 .. R) x [0 .. C) of the padded grid (the phase crop), or the reverse into a zeroed padded grid.
hipError_t launch_crop(const double *padded, int R, int C, int Cp, double *out, hipStream_t s);
hipError_t launch_embed(const double *scan, int R, int C, int Rp, int Cp, double *padded, hipStream_t s);

}  // namespace cubo
