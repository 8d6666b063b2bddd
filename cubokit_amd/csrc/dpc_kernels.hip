// dpc_kernels.hip — FP64 FFT and the iterative-DPC phase step (Sources/CuboPhase/DPC.swift:26-102,
// Sources/CuboCore/DoubleFFT.swift:119-195 for the FFT conventions) and its vector-Jacobian product.
//
// FFT: one block per row, the row in LDS. Lengths that are powers of two run Stockham radix-4 stages (after one
// radix-2 stage for odd powers; no bit reversal: each stage reads one LDS buffer and writes the other in natural
// order); other lengths a direct DFT from the twiddle table. 2-D transforms of powers of two are two passes of
// fft_rows_t (several rows per block, the result stored transposed); others rows, a tiled transpose, rows, a
// transpose back.
// Twiddles exp(-2 pi i m / n) come from the host in Double.
#include "dpc_kernels.hpp"

#include <algorithm>
#include <new>

namespace cubo {
namespace {

enum { kFftPlain = 0, kFftStencil = 1, kFftReal = 2, kFftUpdate = 3, kFftSpectrum = 4, kFftAdjointSpectrum = 5 };
// What a fused pass of fft_rows_t reads and writes besides its rows.
struct FftFuse {
    DpcGrid g;
    const double *phi, *fx, *fy;    // kFftStencil: the phase and the field; kFftUpdate: the phase
    double *partial;                // kFftStencil
    const double *real;             // kFftReal
    double *phi_out;                // kFftUpdate
    double step;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <bool INV>
__global__ __launch_bounds__(256) void fft_rows(double2 *data, int n, const double2 *tw) {
    extern __shared__ double2 sh[];
    double2 *a = sh, *b = sh + n;
    double2 *x = data + (int64_t)blockIdx.x * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = x[i];
    __syncthreads();
    if ((n & (n - 1)) == 0) {
        // Stockham autosort stages: one radix-2 stage when log2 n is odd, then radix-4 stages (half the block
        // barriers of radix-2 throughout). A stage of radix r and span ns takes, for j < n / r and k = j mod ns,
        // the inputs a[j + m n / r] times w^(m k) (w = exp(-+2 pi i / (r ns))), their r-point DFT y_q, and writes
        // b[r (j - k) + k + q ns]; ns grows by r. tw[i] = exp(-2 pi i i / n).
        int ns = 1;
        if (__builtin_ctz((unsigned)n) & 1) {
            const int half = n >> 1;
            for (int j = threadIdx.x; j < half; j += blockDim.x) {
                const double2 u = a[j], v = a[j + half];
                b[2 * j] = make_double2(u.x + v.x, u.y + v.y);
                b[2 * j + 1] = make_double2(u.x - v.x, u.y - v.y);
            }
            __syncthreads();
            double2 *t = a; a = b; b = t;
            ns = 2;
        }
        const int quarter = n >> 2;
        for (; ns < n; ns <<= 2) {
            const int step = quarter / ns;                      // tw index of w^1 = exp(-2 pi i / (4 ns))
            for (int j = threadIdx.x; j < quarter; j += blockDim.x) {
                const int k = j & (ns - 1);
                double2 w1 = tw[k * step], w2 = tw[2 * k * step], w3 = tw[3 * k * step];
                if (INV) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
                const double2 x0 = a[j], x1 = cmul(a[j + quarter], w1), x2 = cmul(a[j + 2 * quarter], w2),
                              x3 = cmul(a[j + 3 * quarter], w3);
                const double2 s02 = make_double2(x0.x + x2.x, x0.y + x2.y), d02 = make_double2(x0.x - x2.x, x0.y - x2.y);
                const double2 s13 = make_double2(x1.x + x3.x, x1.y + x3.y), d13 = make_double2(x1.x - x3.x, x1.y - x3.y);
                // -i d13 (forward) or +i d13 (inverse)
                const double2 r13 = INV ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
                const int o = ((j - k) << 2) + k;
                b[o] = make_double2(s02.x + s13.x, s02.y + s13.y);
                b[o + ns] = make_double2(d02.x + r13.x, d02.y + r13.y);
                b[o + 2 * ns] = make_double2(s02.x - s13.x, s02.y - s13.y);
                b[o + 3 * ns] = make_double2(d02.x - r13.x, d02.y - r13.y);
            }
            __syncthreads();
            double2 *t = a; a = b; b = t;
        }
    } else {
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            double re = 0.0, im = 0.0;
            int m = 0;
            for (int j = 0; j < n; ++j) {
                double2 w = tw[m];
                if (INV) w.y = -w.y;
                re += a[j].x * w.x - a[j].y * w.y;
                im += a[j].x * w.y + a[j].y * w.x;
                m += k;
                if (m >= n) m -= n;
            }
            b[k] = make_double2(re, im);
        }
        __syncthreads();
        double2 *t = a; a = b; b = t;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = a[i];
}

// dx, dy of the scan position (r, c): the central differences of phi plus the field (DPC.swift:59-69).
__device__ __forceinline__ void dpc_d(const DpcGrid &g, const double *phi, const double *fx, const double *fy, int r,
                                      int c, double *dx, double *dy) {
    const int Rp = g.Rp, Cp = g.Cp;
    const double up = phi[(int64_t)((r - 1 + Rp) % Rp) * Cp + c], down = phi[(int64_t)((r + 1) % Rp) * Cp + c];
    const double left = phi[(int64_t)r * Cp + (c - 1 + Cp) % Cp], right = phi[(int64_t)r * Cp + (c + 1) % Cp];
    *dx = (up - down) / (2 * g.sr) + fx[(int64_t)r * g.C + c];
    *dy = (left - right) / (2 * g.sc) + fy[(int64_t)r * g.C + c];
}

// dx^2 + dy^2 and phi + step u with one rounding order in the fused and the separate kernels
__device__ __forceinline__ double dpc_sq(double dx, double dy) { return __builtin_fma(dx, dx, dy * dy); }
__device__ __forceinline__ double dpc_step(double phi, double step, double u) { return __builtin_fma(step, u, phi); }

// The operators at frequency (r, c), rounded as the host's table was (no fused products).
__device__ __forceinline__ void dpc_ops(const DpcGrid &g, int r, int c, double *gx, double *gy) {
#pragma clang fp contract(off)
    const double kx = g.kx[r], ky = g.ky[c];
    double den = kx * kx + ky * ky;
    if (r == 0 && c == 0) den = __builtin_inf();
    const double kden = 1.0 / den;
    *gx = -0.25 * kx * kden;
    *gy = -0.25 * ky * kden;
}

// F = FFT(dx + i dy): FFT(dx)[k] = (F[k] + conj F[-k]) / 2, FFT(dy)[k] = (F[k] - conj F[-k]) / (2i);
// the update's spectrum i (gx FFT(dx) + gy FFT(dy)) at k from fk = F[k], fm = F[-k].
__device__ __forceinline__ double2 dpc_spec(const DpcGrid &g, int r, int c, double2 fk, double2 fm) {
    const double2 X = make_double2(0.5 * (fk.x + fm.x), 0.5 * (fk.y - fm.y));
    const double2 Y = make_double2(0.5 * (fk.y + fm.y), -0.5 * (fk.x - fm.x));
    double gx, gy;
    dpc_ops(g, r, c, &gx, &gy);
    const double sx = gx * X.x + gy * Y.x, sy = gx * X.y + gy * Y.y;
    return make_double2(-sy, sx);
}

// G = FFT(g), g real. The adjoint needs Re(IFFT(P)) and Re(IFFT(Q)) for P = G conj(i gx), Q = G conj(i gy).
// Re(IFFT(P)) = IFFT of P's Hermitian part (P[k] + conj P[-k]) / 2 (P itself is not Hermitian on the Nyquist
// row and column of an even length, where gx[-k] = gx[k]), so out = Herm(P) + i Herm(Q) gives both real parts
// from one inverse transform. At k from vk = G[k], vm = G[-k] and the operators at k and -k:
__device__ __forceinline__ double2 dpc_adj(double2 vk, double2 vm, double gxk, double gxm, double gyk, double gym) {
    // P[k] = -i gx[k] G[k] = (gx G.y, -gx G.x); conj P[-k] = (gx[-k] G[-k].y, gx[-k] G[-k].x); each product at k
    // fused with the one at -k, written out so the rounding does not depend on the compiler's contraction
    const double2 P = make_double2(0.5 * __builtin_fma(gxk, vk.y, gxm * vm.y), 0.5 * __builtin_fma(-gxk, vk.x, gxm * vm.x));
    const double2 Q = make_double2(0.5 * __builtin_fma(gyk, vk.y, gym * vm.y), 0.5 * __builtin_fma(-gyk, vk.x, gym * vm.x));
    return make_double2(P.x - Q.y, P.y + Q.x);
}

// R rows of a power-of-two length n per block (1024 threads: one radix-4 butterfly per thread and stage for 4 rows
// of 1024), each transformed in LDS by the same Stockham stages as fft_rows,
// the results stored transposed: out[j][r0 + r] = scale X_r[j] (out is n x rows), for batch blockIdx.y (in: rows x n,
// both arrays batch-strided by rows * n). Two such passes are a 2-D transform with no transpose kernels. LDS rows are
// padded by 16 / R entries, so the transposed reads (R rows at each of 16 / R consecutive j per 16 lanes) fall on
// distinct banks.
//
// MODE fuses the DPC step's element-wise work into a pass (the values and their rounding are those of the separate
// kernels dpc_forward, real_to_complex and dpc_update):
//   kFftStencil: the rows are z = dx + i dy of the padded grid, computed from phi and the field as they are loaded
//                (zero outside the scan region), and f.partial[k] = dpc_forward's tree sum of dx^2 + dy^2 over the
//                256 grid entries 256 k .. 256 k + 255 (R n a multiple of 256; the tree runs in buffer b);
//   kFftReal:    the rows are (f.real, 0);
//   kFftUpdate:  (second pass of an inverse transform) the result u is not stored: f.phi_out = phi + step Re(u);
//   kFftSpectrum, kFftAdjointSpectrum: (first pass of an inverse transform) the rows are the update's spectrum
//                (dpc_spec) or the adjoint's (dpc_adj) of the forward transform in, formed from its entries at k and
//                -k as they are loaded (the values dpc_spectrum and dpc_adjoint_spectrum write).
// TWL: the radix-4 stages read their twiddles from per-stage tables in LDS behind the two row buffers, copied from
// tw at the start: for the stage of span ns, w^(m k) (m = 1, 2, 3, k < ns) at offset (m - 1) ns + k of the stage's
// table, so the lanes of a stage (consecutive k) read consecutive entries (tw itself is read at a stride of
// n / (4 ns) entries). The values, and so the results, are tw's. The fused modes assume one array (batch 1).
template <bool INV, int R, int MODE, bool TWL>
__global__ __launch_bounds__(1024) void fft_rows_t(const double2 *in, double2 *out, int n, int rows, const double2 *tw,
                                                  double scale, FftFuse f) {
    extern __shared__ double2 sh[];
    const int S = n + 16 / R;
    double2 *a = sh, *b = sh + R * S;
    const int r0 = blockIdx.x * R, lg = __builtin_ctz((unsigned)n);
    const int64_t boff = (int64_t)blockIdx.y * rows * n;
    const int quarter = n >> 2, ns0 = (lg & 1) ? 2 : 1;
    double2 *tws = sh + 2 * R * S;
    if constexpr (TWL) {
        int total = 0;
        for (int m = ns0; m < n; m <<= 2) total += 3 * m;
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            int m = ns0, base = 0;
            while (e >= base + 3 * m) { base += 3 * m; m <<= 2; }
            const int p = (e - base) / m + 1, k = (e - base) - (p - 1) * m;
            tws[e] = tw[p * k * (quarter / m)];
        }
    }
    if constexpr (MODE == kFftStencil) {
        double *q = (double *)b;
        for (int v = threadIdx.x; v < R * n; v += blockDim.x) {
            const int r = r0 + (v >> lg), c = v & (n - 1);
            double dx = 0.0, dy = 0.0, sq = 0.0;
            if (r < f.g.R && c < f.g.C) {
                dpc_d(f.g, f.phi, f.fx, f.fy, r, c, &dx, &dy);
                sq = dpc_sq(dx, dy);
            }
            a[(v >> lg) * S + c] = make_double2(dx, dy);
            q[v] = sq;
        }
        __syncthreads();
        const int chunks = R * n / 256;
        for (int h = 128; h > 0; h >>= 1) {
            for (int v = threadIdx.x; v < chunks * h; v += blockDim.x) {
                const int k = v / h, t = v - k * h;
                q[k * 256 + t] += q[k * 256 + t + h];
            }
            __syncthreads();
        }
        for (int k = threadIdx.x; k < chunks; k += blockDim.x) f.partial[(int64_t)r0 * n / 256 + k] = q[k * 256];
        __syncthreads();                                // b is the first stage's output
    } else if constexpr (MODE == kFftSpectrum || MODE == kFftAdjointSpectrum) {
        for (int v = threadIdx.x; v < R * n; v += blockDim.x) {
            const int r = r0 + (v >> lg), c = v & (n - 1);
            const int rm = (f.g.Rp - r) % f.g.Rp, cm = (f.g.Cp - c) % f.g.Cp;
            const double2 fk = in[(int64_t)r * n + c], fm = in[(int64_t)rm * n + cm];
            double2 sv;
            if constexpr (MODE == kFftSpectrum) {
                sv = dpc_spec(f.g, r, c, fk, fm);
            } else {
                double gxk, gyk, gxm, gym;
                dpc_ops(f.g, r, c, &gxk, &gyk);
                dpc_ops(f.g, rm, cm, &gxm, &gym);
                sv = dpc_adj(fk, fm, gxk, gxm, gyk, gym);
            }
            a[(v >> lg) * S + c] = sv;
        }
        __syncthreads();
    } else if constexpr (MODE == kFftReal) {
        const double *x = f.real + boff + (int64_t)r0 * n;
        for (int v = threadIdx.x; v < R * n; v += blockDim.x) a[(v >> lg) * S + (v & (n - 1))] = make_double2(x[v], 0.0);
        __syncthreads();
    } else {
        const double2 *x = in + boff + (int64_t)r0 * n;
        for (int v = threadIdx.x; v < R * n; v += blockDim.x) a[(v >> lg) * S + (v & (n - 1))] = x[v];
        __syncthreads();
    }
    int ns = 1;
    if (lg & 1) {
        const int half = n >> 1;
        for (int v = threadIdx.x; v < R * half; v += blockDim.x) {
            const int r = v >> (lg - 1), j = v & (half - 1);
            const double2 u = a[r * S + j], w = a[r * S + j + half];
            b[r * S + 2 * j] = make_double2(u.x + w.x, u.y + w.y);
            b[r * S + 2 * j + 1] = make_double2(u.x - w.x, u.y - w.y);
        }
        __syncthreads();
        double2 *t = a; a = b; b = t;
        ns = 2;
    }
    int tb = 0;                                         // the stage's table (TWL)
    for (; ns < n; tb += 3 * ns, ns <<= 2) {
        const int step = quarter / ns;
        for (int v = threadIdx.x; v < R * quarter; v += blockDim.x) {
            const int r = v >> (lg - 2), j = v & (quarter - 1), k = j & (ns - 1);
            const double2 *ar = a + r * S;
            double2 *br = b + r * S;
            double2 w1, w2, w3;
            if constexpr (TWL) {
                w1 = tws[tb + k];
                w2 = tws[tb + ns + k];
                w3 = tws[tb + 2 * ns + k];
            } else {
                w1 = tw[k * step];
                w2 = tw[2 * k * step];
                w3 = tw[3 * k * step];
            }
            if (INV) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
            const double2 x0 = ar[j], x1 = cmul(ar[j + quarter], w1), x2 = cmul(ar[j + 2 * quarter], w2),
                          x3 = cmul(ar[j + 3 * quarter], w3);
            const double2 s02 = make_double2(x0.x + x2.x, x0.y + x2.y), d02 = make_double2(x0.x - x2.x, x0.y - x2.y);
            const double2 s13 = make_double2(x1.x + x3.x, x1.y + x3.y), d13 = make_double2(x1.x - x3.x, x1.y - x3.y);
            const double2 r13 = INV ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
            const int o = ((j - k) << 2) + k;
            br[o] = make_double2(s02.x + s13.x, s02.y + s13.y);
            br[o + ns] = make_double2(d02.x + r13.x, d02.y + r13.y);
            br[o + 2 * ns] = make_double2(s02.x - s13.x, s02.y - s13.y);
            br[o + 3 * ns] = make_double2(d02.x - r13.x, d02.y - r13.y);
        }
        __syncthreads();
        double2 *t = a; a = b; b = t;
    }
    if constexpr (MODE == kFftUpdate) {
        const double *ph = f.phi + boff + r0;
        double *po = f.phi_out + boff + r0;
        for (int v = threadIdx.x; v < R * n; v += blockDim.x) {
            const int r = v % R, j = v / R;
            const int64_t o = (int64_t)j * rows + r;
            po[o] = dpc_step(ph[o], f.step, a[r * S + j].x * scale);
        }
    } else {
        double2 *y = out + boff + r0;
        for (int v = threadIdx.x; v < R * n; v += blockDim.x) {
            const int r = v % R, j = v / R;
            const double2 e = a[r * S + j];
            y[(int64_t)j * rows + r] = make_double2(e.x * scale, e.y * scale);
        }
    }
}

// Rows per block of fft_rows_t for a row length: two LDS buffers of R padded rows within 160 KiB.
int fft_rows_t_rows(int n) { return n <= 1024 ? 4 : 2; }

// One pass of fft_rows_t over batch arrays of rows x n, or false if the shapes do not suit it.
template <int MODE = kFftPlain>
bool launch_fft_rows_t(const double2 *in, double2 *out, int batch, int rows, int n, const double2 *tw, bool inverse,
                       double scale, hipStream_t s, const FftFuse &f = FftFuse{}) {
    if (n < 4 || (n & (n - 1)) || n > kFftMax) return false;
    const int R = fft_rows_t_rows(n);
    if (rows % R) return false;
    // the per-stage twiddle tables when they fit beside the rows (n <= 1024)
    size_t tw_entries = 0;
    for (int m = (__builtin_ctz((unsigned)n) & 1) ? 2 : 1; m < n; m <<= 2) tw_entries += 3 * (size_t)m;
    const size_t rows_lds = 2 * (size_t)R * (n + 16 / R) * sizeof(double2), tw_lds = tw_entries * sizeof(double2);
    const bool twl = rows_lds + tw_lds <= 160 * 1024;
    const size_t lds = rows_lds + (twl ? tw_lds : 0);
    const dim3 grid((unsigned)(rows / R), (unsigned)batch);
#define CUBO_RT(I, RR, T) hipLaunchKernelGGL((fft_rows_t<I, RR, MODE, T>), grid, dim3(1024), lds, s, in, out, n, rows, tw, scale, f)
#define CUBO_RT2(I, RR) { if (twl) CUBO_RT(I, RR, true); else CUBO_RT(I, RR, false); }
    if (R == 4) { if (inverse) CUBO_RT2(true, 4) else CUBO_RT2(false, 4) }
    else { if (inverse) CUBO_RT2(true, 2) else CUBO_RT2(false, 2) }
#undef CUBO_RT2
#undef CUBO_RT
    return true;
}

// out[c][r] = scale * in[r][c], for each of gridDim.z arrays of rows x cols
__global__ __launch_bounds__(256) void transpose(const double2 *in, double2 *out, int rows, int cols, double scale) {
    __shared__ double2 t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    in += (int64_t)blockIdx.z * rows * cols;
    out += (int64_t)blockIdx.z * rows * cols;
    for (int i = ty; i < 32; i += 8) {
        const int r = by + i, c = bx + tx;
        if (r < rows && c < cols) t[i][tx] = in[(int64_t)r * cols + c];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int c = bx + i, r = by + tx;
        if (c < cols && r < rows) {
            const double2 v = t[tx][i];
            out[(int64_t)c * rows + r] = make_double2(v.x * scale, v.y * scale);
        }
    }
}

__global__ __launch_bounds__(256) void dpc_forward(DpcGrid g, const double *phi, const double *fx, const double *fy,
                                                   double2 *z, double *partial) {
    __shared__ double red[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n = (int64_t)g.Rp * g.Cp;
    double sq = 0.0;
    if (i < n) {
        const int r = (int)(i / g.Cp), c = (int)(i % g.Cp);
        double dx = 0.0, dy = 0.0;
        if (r < g.R && c < g.C) {
            dpc_d(g, phi, fx, fy, r, c, &dx, &dy);
            sq = dpc_sq(dx, dy);
        }
        z[i] = make_double2(dx, dy);
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Index m of the frequency -k for k = i, and (r, c) of each; false when m < i (the pair is another thread's).
__device__ __forceinline__ bool dpc_pair(const DpcGrid &g, int64_t i, int *r, int *c, int *rm, int *cm, int64_t *m) {
    *r = (int)(i / g.Cp);
    *c = (int)(i % g.Cp);
    *rm = (g.Rp - *r) % g.Rp;
    *cm = (g.Cp - *c) % g.Cp;
    *m = (int64_t)*rm * g.Cp + *cm;
    return *m >= i;
}

// One thread per pair {k, -k}: both spectra from one read of each.
__global__ __launch_bounds__(256) void dpc_spectrum(DpcGrid g, const double2 *f, double2 *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)g.Rp * g.Cp) return;
    int r, c, rm, cm;
    int64_t m;
    if (!dpc_pair(g, i, &r, &c, &rm, &cm, &m)) return;
    const double2 fk = f[i], fm = f[m];
    out[i] = dpc_spec(g, r, c, fk, fm);
    if (m != i) out[m] = dpc_spec(g, rm, cm, fm, fk);
}

__global__ __launch_bounds__(256) void dpc_update(DpcGrid g, const double *phi, const double2 *u, double step,
                                                  double *phi_out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (int64_t)g.Rp * g.Cp) phi_out[i] = dpc_step(phi[i], step, u[i].x);
}

__global__ __launch_bounds__(256) void real_to_complex(const double *in, double2 *z, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) z[i] = make_double2(in[i], 0.0);
}

__global__ __launch_bounds__(256) void dpc_adjoint_spectrum(DpcGrid g, const double2 *G, double2 *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)g.Rp * g.Cp) return;
    int r, c, rm, cm;
    int64_t m;
    if (!dpc_pair(g, i, &r, &c, &rm, &cm, &m)) return;
    const double2 vk = G[i], vm = G[m];
    double gxk, gyk, gxm, gym;
    dpc_ops(g, r, c, &gxk, &gyk);
    dpc_ops(g, rm, cm, &gxm, &gym);
    out[i] = dpc_adj(vk, vm, gxk, gxm, gyk, gym);
    if (m != i) out[m] = dpc_adj(vm, vk, gxm, gxk, gym, gyk);
}

// gdx, gdy of scan entry (r, c): dL/d(dx), dL/d(dy) of the step (the update's share w step and the error's).
__device__ __forceinline__ void gdxy(const DpcGrid &g, const double *phi, const double *fx, const double *fy,
                                     const double2 *w, double step, double g_err, double denom, int r, int c,
                                     double *gdx, double *gdy) {
    double dx, dy;
    dpc_d(g, phi, fx, fy, r, c, &dx, &dy);
    const double2 wv = w[(int64_t)r * g.Cp + c];
    const double N = (double)g.R * g.C;
    *gdx = wv.x * step + g_err * 2.0 * dx / (N * denom);
    *gdy = wv.y * step + g_err * 2.0 * dy / (N * denom);
}

// The stencil's adjoint over tiles of kVjRows x kVjCols grid entries: the pushes a = gdx / (2 sr), b = gdy / (2 sc)
// of the tile and a one-entry halo (wrapping round the padded grid, zero off the scan region) are formed once each in
// LDS, then g_phi = g_phi_out + a(r + 1, c) - a(r - 1, c) + b(r, c + 1) - b(r, c - 1) (dx at (r + 1, c) pushes +a
// into phi (r, c), dx at (r - 1, c) pushes -a; dy likewise along the row) and, on the scan region, the field's
// gradient g_fx += gdx + dd 2 fx / N, g_fy likewise.
constexpr int kVjRows = 8, kVjCols = 32;
__global__ __launch_bounds__(256) void dpc_vjp(DpcGrid g, const double *phi, const double *fx, const double *fy,
                                               const double2 *w, double step, double g_err, double denom, double dd,
                                               const double *g_phi_out, double *g_phi, double *g_fx, double *g_fy) {
    constexpr int HR = kVjRows + 2, HC = kVjCols + 2;
    __shared__ double sa[HR][HC], sb[HR][HC], sx[kVjRows][kVjCols], sy[kVjRows][kVjCols];
    const int Rp = g.Rp, Cp = g.Cp;
    const int ntx = (Cp + kVjCols - 1) / kVjCols;         // tiles along a row (one-dimensional grid of tiles)
    const int r0 = (int)(blockIdx.x / ntx) * kVjRows, c0 = (int)(blockIdx.x % ntx) * kVjCols;
    for (int p = threadIdx.x; p < HR * HC; p += blockDim.x) {
        const int pr = p / HC, pc = p % HC;
        const int r = (r0 - 1 + pr + Rp) % Rp, c = (c0 - 1 + pc + Cp) % Cp;
        double a = 0.0, b = 0.0, gdx = 0.0, gdy = 0.0;
        if (r < g.R && c < g.C) {
            gdxy(g, phi, fx, fy, w, step, g_err, denom, r, c, &gdx, &gdy);
            a = gdx / (2 * g.sr);
            b = gdy / (2 * g.sc);
        }
        sa[pr][pc] = a;
        sb[pr][pc] = b;
        if (pr >= 1 && pr <= kVjRows && pc >= 1 && pc <= kVjCols) {
            sx[pr - 1][pc - 1] = gdx;
            sy[pr - 1][pc - 1] = gdy;
        }
    }
    __syncthreads();
    const int ty = threadIdx.x / kVjCols, tx = threadIdx.x % kVjCols;
    const int r = r0 + ty, c = c0 + tx;
    if (r >= Rp || c >= Cp) return;
    const int64_t i = (int64_t)r * Cp + c;
    g_phi[i] = g_phi_out[i] + sa[ty + 2][tx + 1] - sa[ty][tx + 1] + sb[ty + 1][tx + 2] - sb[ty + 1][tx];
    if (r < g.R && c < g.C) {
        const double N = (double)g.R * g.C;
        const int64_t k = (int64_t)r * g.C + c;
        g_fx[k] += sx[ty][tx] + dd * 2.0 * fx[k] / N;
        g_fy[k] += sy[ty][tx] + dd * 2.0 * fy[k] / N;
    }
}

__global__ __launch_bounds__(256) void sum_partials(const double *partial, int n, double *out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void crop(const double *padded, int R, int C, int Cp, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)R * C) out[i] = padded[(i / C) * Cp + i % C];
}

__global__ void embed(const double *scan, int R, int C, int Rp, int Cp, double *padded) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)Rp * Cp) return;
    const int r = (int)(i / Cp), c = (int)(i % Cp);
    padded[i] = (r < R && c < C) ? scan[(int64_t)r * C + c] : 0.0;
}

unsigned blocks256(int64_t n) { return (unsigned)((n + 255) / 256); }

__device__ __forceinline__ double2 conj_if(double2 w, bool c) { return c ? make_double2(w.x, -w.y) : w; }

// t[b][j2][k1] *= W_n^(j2 k1) (conjugated for the inverse); j2 k1 < n1 n2 = n, so no reduction is needed.
__global__ __launch_bounds__(256) void fourstep_twiddle(double2 *t, int64_t total, int n1, int n2, const double2 *twn,
                                                       bool inv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const int k1 = (int)(i % n1), j2 = (int)((i / n1) % n2);
    t[i] = cmul(t[i], conj_if(twn[(int64_t)j2 * k1], inv));
}

// a[b][j] = x[b][j] c_j for j < n, 0 up to M (c conjugated for the inverse)
__global__ __launch_bounds__(256) void bluestein_pre(const double2 *x, double2 *a, int64_t total, int n, int M,
                                                    const double2 *chirp, bool inv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const int64_t b = i / M;
    const int j = (int)(i % M);
    a[i] = j < n ? cmul(x[b * n + j], conj_if(chirp[j], inv)) : make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void bluestein_mul(double2 *a, int64_t total, int M, const double2 *bf) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < total) a[i] = cmul(a[i], bf[i % M]);
}

// x[b][k] = c[b][k] / M * c_k for k < n
__global__ __launch_bounds__(256) void bluestein_post(const double2 *c, double2 *x, int64_t total, int n, int M,
                                                     const double2 *chirp, bool inv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const int64_t b = i / n;
    const int k = (int)(i % n);
    const double2 v = c[b * M + k];
    x[i] = cmul(make_double2(v.x / M, v.y / M), conj_if(chirp[k], inv));
}

}  // namespace

struct FftPlan {
    int n = 0, kind = 0;          // 0: fft_rows, 1: four-step, 2: Bluestein
    int max_batch = 0;
    double2 *tw = nullptr;        // kind 0: exp(-2 pi i m / n); kind 1: W_n^m, m < n
    int n1 = 0, n2 = 0, M = 0;
    FftPlan *p1 = nullptr, *p2 = nullptr;   // four-step: lengths n1, n2; Bluestein: length M
    double2 *scratch = nullptr;   // four-step: max_batch * n; Bluestein: max_batch * M
    double2 *chirp = nullptr, *bf = nullptr, *bfi = nullptr;
};

void fft_plan_destroy(FftPlan *p) {
    if (!p) return;
    fft_plan_destroy(p->p1);
    fft_plan_destroy(p->p2);
    for (double2 *b : {p->tw, p->scratch, p->chirp, p->bf, p->bfi})
        if (b) (void)hipFree(b);
    delete p;
}

static hipError_t upload(const std::vector<double2> &v, double2 **dst) {
    hipError_t e = hipMalloc((void **)dst, v.size() * sizeof(double2));
    if (e == hipSuccess) e = hipMemcpy(*dst, v.data(), v.size() * sizeof(double2), hipMemcpyHostToDevice);
    return e;
}

hipError_t fft_plan_create(int n, int max_batch, hipStream_t s, FftPlan **out) {
    *out = nullptr;
    if (n < 1 || max_batch < 1) return hipErrorInvalidValue;
    FftPlan *p = new (std::nothrow) FftPlan();
    if (!p) return hipErrorOutOfMemory;
    p->n = n;
    p->max_batch = max_batch;
    const bool pow2 = (n & (n - 1)) == 0;
    hipError_t e = hipSuccess;
    if (n <= kFftMax) {
        p->kind = 0;
        e = upload(fft_twiddles(n), &p->tw);
    } else if (pow2) {
        p->kind = 1;
        int k = 0;
        while ((1 << k) < n) ++k;
        p->n1 = 1 << (k / 2);
        p->n2 = n / p->n1;
        if (p->n2 > kFftMax) e = hipErrorInvalidValue;                        // n above 2^22
        if (e == hipSuccess) e = upload(fft_twiddles(n), &p->tw);
        if (e == hipSuccess) e = fft_plan_create(p->n1, max_batch * p->n2, s, &p->p1);
        if (e == hipSuccess) e = fft_plan_create(p->n2, max_batch * p->n1, s, &p->p2);
        if (e == hipSuccess) e = hipMalloc((void **)&p->scratch, (size_t)max_batch * n * sizeof(double2));
    } else {
        p->kind = 2;
        int M = 1;
        while (M < 2 * n - 1) M <<= 1;
        p->M = M;
        std::vector<double2> chirp(n), b(M, make_double2(0.0, 0.0)), bc(M, make_double2(0.0, 0.0));
        for (int j = 0; j < n; ++j) {
            const long long q = ((long long)j * j) % (2LL * n);
            const double ang = -M_PI * (double)q / (double)n;
            chirp[j] = make_double2(std::cos(ang), std::sin(ang));
            const double2 cj = make_double2(chirp[j].x, -chirp[j].y);         // conj(c_j): the kernel at +-j
            b[j] = cj;
            bc[j] = chirp[j];
            if (j > 0) { b[M - j] = cj; bc[M - j] = chirp[j]; }
        }
        e = upload(chirp, &p->chirp);
        if (e == hipSuccess) e = upload(b, &p->bf);
        if (e == hipSuccess) e = upload(bc, &p->bfi);
        if (e == hipSuccess) e = fft_plan_create(M, std::max(max_batch, 1), s, &p->p1);
        if (e == hipSuccess) e = hipMalloc((void **)&p->scratch, (size_t)max_batch * M * sizeof(double2));
        // the kernel's transforms, once (forward for both directions: the inverse uses the conjugate kernel)
        if (e == hipSuccess) e = launch_fft_rows(p->p1, p->bf, 1, false, s);
        if (e == hipSuccess) e = launch_fft_rows(p->p1, p->bfi, 1, false, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) {
        fft_plan_destroy(p);
        return e;
    }
    *out = p;
    return hipSuccess;
}

hipError_t launch_fft_rows(FftPlan *p, double2 *data, int batch, bool inverse, hipStream_t s) {
    if (!p || batch < 1 || batch > p->max_batch) return hipErrorInvalidValue;
    const int n = p->n;
    if (p->kind == 0) {
        const size_t lds = 2 * (size_t)n * sizeof(double2);
        if (inverse) hipLaunchKernelGGL(fft_rows<true>, dim3((unsigned)batch), dim3(256), lds, s, data, n, p->tw);
        else hipLaunchKernelGGL(fft_rows<false>, dim3((unsigned)batch), dim3(256), lds, s, data, n, p->tw);
        return hipGetLastError();
    }
    const int64_t total = (int64_t)batch * n;
    if (p->kind == 1) {
        const int n1 = p->n1, n2 = p->n2;
        double2 *t = p->scratch;
        for (int b0 = 0; b0 < batch; b0 += 65535) {                          // grid z of the transposes
            const int nb = std::min(65535, batch - b0);
            hipLaunchKernelGGL(transpose, dim3((n2 + 31) / 32, (n1 + 31) / 32, nb), dim3(256), 0, s,
                               data + (int64_t)b0 * n, t + (int64_t)b0 * n, n1, n2, 1.0);
        }
        if (hipError_t e = launch_fft_rows(p->p1, t, batch * n2, inverse, s)) return e;
        hipLaunchKernelGGL(fourstep_twiddle, dim3(blocks256(total)), dim3(256), 0, s, t, total, n1, n2, p->tw, inverse);
        for (int b0 = 0; b0 < batch; b0 += 65535) {
            const int nb = std::min(65535, batch - b0);
            hipLaunchKernelGGL(transpose, dim3((n1 + 31) / 32, (n2 + 31) / 32, nb), dim3(256), 0, s,
                               t + (int64_t)b0 * n, data + (int64_t)b0 * n, n2, n1, 1.0);
        }
        if (hipError_t e = launch_fft_rows(p->p2, data, batch * n1, inverse, s)) return e;
        for (int b0 = 0; b0 < batch; b0 += 65535) {
            const int nb = std::min(65535, batch - b0);
            hipLaunchKernelGGL(transpose, dim3((n2 + 31) / 32, (n1 + 31) / 32, nb), dim3(256), 0, s,
                               data + (int64_t)b0 * n, t + (int64_t)b0 * n, n1, n2, 1.0);
        }
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(data, t, (size_t)total * sizeof(double2), hipMemcpyDeviceToDevice, s);
        return e;
    }
    const int M = p->M;
    const int64_t tm = (int64_t)batch * M;
    double2 *a = p->scratch;
    hipLaunchKernelGGL(bluestein_pre, dim3(blocks256(tm)), dim3(256), 0, s, data, a, tm, n, M, p->chirp, inverse);
    if (hipError_t e = launch_fft_rows(p->p1, a, batch, false, s)) return e;
    hipLaunchKernelGGL(bluestein_mul, dim3(blocks256(tm)), dim3(256), 0, s, a, tm, M, inverse ? p->bfi : p->bf);
    if (hipError_t e = launch_fft_rows(p->p1, a, batch, true, s)) return e;
    hipLaunchKernelGGL(bluestein_post, dim3(blocks256(total)), dim3(256), 0, s, a, data, total, n, M, p->chirp, inverse);
    return hipGetLastError();
}

hipError_t launch_fft2_planned(double2 *data, double2 *tmp, int rows, int cols, FftPlan *pr, FftPlan *pc,
                               bool inverse, hipStream_t s) {
    if (!pr || !pc || pr->n != rows || pc->n != cols) return hipErrorInvalidValue;
    if (hipError_t e = launch_fft_rows(pc, data, rows, inverse, s)) return e;
    hipLaunchKernelGGL(transpose, dim3((cols + 31) / 32, (rows + 31) / 32, 1), dim3(256), 0, s, data, tmp, rows, cols,
                       1.0);
    if (hipError_t e = launch_fft_rows(pr, tmp, cols, inverse, s)) return e;
    const double scale = inverse ? 1.0 / ((double)rows * cols) : 1.0;
    hipLaunchKernelGGL(transpose, dim3((rows + 31) / 32, (cols + 31) / 32, 1), dim3(256), 0, s, tmp, data, cols, rows,
                       scale);
    return hipGetLastError();
}

hipError_t launch_fft2_batched(double2 *data, double2 *tmp, int batch, int rows, int cols, const double2 *tw_r,
                               const double2 *tw_c, bool inverse, hipStream_t s) {
    if (rows > kFftMax || cols > kFftMax || rows < 1 || cols < 1 || batch < 1 || batch > 65535)
        return hipErrorInvalidValue;
    const double scale = inverse ? 1.0 / ((double)rows * cols) : 1.0;
    // powers of two whose row counts divide into the blocks: two passes of rows stored transposed
    const auto fits = [](int n, int m) { return n >= 4 && !(n & (n - 1)) && m % fft_rows_t_rows(n) == 0; };
    if (fits(cols, rows) && fits(rows, cols)) {
        launch_fft_rows_t(data, tmp, batch, rows, cols, tw_c, inverse, 1.0, s);
        launch_fft_rows_t(tmp, data, batch, cols, rows, tw_r, inverse, scale, s);
        return hipGetLastError();
    }
    const size_t lds_c = 2 * (size_t)cols * sizeof(double2), lds_r = 2 * (size_t)rows * sizeof(double2);
    const unsigned nr = (unsigned)((int64_t)batch * rows), nc = (unsigned)((int64_t)batch * cols);
    if (inverse) hipLaunchKernelGGL(fft_rows<true>, dim3(nr), dim3(256), lds_c, s, data, cols, tw_c);
    else hipLaunchKernelGGL(fft_rows<false>, dim3(nr), dim3(256), lds_c, s, data, cols, tw_c);
    hipLaunchKernelGGL(transpose, dim3((cols + 31) / 32, (rows + 31) / 32, batch), dim3(256), 0, s, data, tmp, rows, cols,
                       1.0);
    if (inverse) hipLaunchKernelGGL(fft_rows<true>, dim3(nc), dim3(256), lds_r, s, tmp, rows, tw_r);
    else hipLaunchKernelGGL(fft_rows<false>, dim3(nc), dim3(256), lds_r, s, tmp, rows, tw_r);
    hipLaunchKernelGGL(transpose, dim3((rows + 31) / 32, (cols + 31) / 32, batch), dim3(256), 0, s, tmp, data, cols, rows,
                       scale);
    return hipGetLastError();
}

hipError_t launch_fft2(double2 *data, double2 *tmp, int rows, int cols, const double2 *tw_r, const double2 *tw_c,
                       bool inverse, hipStream_t s) {
    return launch_fft2_batched(data, tmp, 1, rows, cols, tw_r, tw_c, inverse, s);
}

bool dpc_fused_supported(int Rp, int Cp) {
    const auto fits = [](int n, int m) { return n >= 4 && !(n & (n - 1)) && n <= kFftMax && m % fft_rows_t_rows(n) == 0; };
    return fits(Cp, Rp) && fits(Rp, Cp) && (fft_rows_t_rows(Cp) * Cp) % 256 == 0;
}

hipError_t launch_dpc_forward_pass(const DpcGrid &g, const double *phi, const double *fx, const double *fy,
                                   const double2 *tw_c, double2 *tmp, double *partial, int *nblocks, hipStream_t s) {
    if (!dpc_fused_supported(g.Rp, g.Cp)) return hipErrorInvalidValue;
    FftFuse f{};
    f.g = g;
    f.phi = phi;
    f.fx = fx;
    f.fy = fy;
    f.partial = partial;
    *nblocks = (int)((int64_t)g.Rp * g.Cp / 256);
    launch_fft_rows_t<kFftStencil>(nullptr, tmp, 1, g.Rp, g.Cp, tw_c, false, 1.0, s, f);
    return hipGetLastError();
}

hipError_t launch_fft2_second_pass(const double2 *tmp, double2 *data, int rows, int cols, const double2 *tw_r,
                                   hipStream_t s) {
    if (!dpc_fused_supported(rows, cols)) return hipErrorInvalidValue;
    launch_fft_rows_t(tmp, data, 1, cols, rows, tw_r, false, 1.0, s);
    return hipGetLastError();
}

hipError_t launch_fft2_real(const double *in, double2 *tmp, double2 *data, int rows, int cols, const double2 *tw_r,
                            const double2 *tw_c, hipStream_t s) {
    if (!dpc_fused_supported(rows, cols)) return hipErrorInvalidValue;
    FftFuse f{};
    f.real = in;
    launch_fft_rows_t<kFftReal>(nullptr, tmp, 1, rows, cols, tw_c, false, 1.0, s, f);
    launch_fft_rows_t(tmp, data, 1, cols, rows, tw_r, false, 1.0, s);
    return hipGetLastError();
}

hipError_t launch_ifft2_spectrum_update(const DpcGrid &g, const double2 *F, double2 *tmp, const double2 *tw_r,
                                        const double2 *tw_c, const double *phi, double step, double *phi_out,
                                        hipStream_t s) {
    if (!dpc_fused_supported(g.Rp, g.Cp)) return hipErrorInvalidValue;
    FftFuse f{};
    f.g = g;
    f.phi = phi;
    f.phi_out = phi_out;
    f.step = step;
    launch_fft_rows_t<kFftSpectrum>(F, tmp, 1, g.Rp, g.Cp, tw_c, true, 1.0, s, f);
    launch_fft_rows_t<kFftUpdate>(tmp, nullptr, 1, g.Cp, g.Rp, tw_r, true, 1.0 / ((double)g.Rp * g.Cp), s, f);
    return hipGetLastError();
}

hipError_t launch_ifft2_adjoint_spectrum(const DpcGrid &g, const double2 *G, double2 *tmp, double2 *out,
                                         const double2 *tw_r, const double2 *tw_c, hipStream_t s) {
    if (!dpc_fused_supported(g.Rp, g.Cp)) return hipErrorInvalidValue;
    FftFuse f{};
    f.g = g;
    launch_fft_rows_t<kFftAdjointSpectrum>(G, tmp, 1, g.Rp, g.Cp, tw_c, true, 1.0, s, f);
    launch_fft_rows_t(tmp, out, 1, g.Cp, g.Rp, tw_r, true, 1.0 / ((double)g.Rp * g.Cp), s);
    return hipGetLastError();
}

hipError_t launch_dpc_forward(const DpcGrid &g, const double *phi, const double *fx, const double *fy, double2 *z,
                              double *partial, int *nblocks, hipStream_t s) {
    const int64_t n = (int64_t)g.Rp * g.Cp;
    *nblocks = (int)blocks256(n);
    hipLaunchKernelGGL(dpc_forward, dim3(*nblocks), dim3(256), 0, s, g, phi, fx, fy, z, partial);
    return hipGetLastError();
}

hipError_t launch_dpc_spectrum(const DpcGrid &g, const double2 *f, double2 *out, hipStream_t s) {
    hipLaunchKernelGGL(dpc_spectrum, dim3(blocks256((int64_t)g.Rp * g.Cp)), dim3(256), 0, s, g, f, out);
    return hipGetLastError();
}

hipError_t launch_dpc_update(const DpcGrid &g, const double *phi, const double2 *u, double step, double *phi_out,
                             hipStream_t s) {
    hipLaunchKernelGGL(dpc_update, dim3(blocks256((int64_t)g.Rp * g.Cp)), dim3(256), 0, s, g, phi, u, step, phi_out);
    return hipGetLastError();
}

hipError_t launch_real_to_complex(const double *in, double2 *z, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(real_to_complex, dim3(blocks256(n)), dim3(256), 0, s, in, z, n);
    return hipGetLastError();
}

hipError_t launch_dpc_adjoint_spectrum(const DpcGrid &g, const double2 *G, double2 *out, hipStream_t s) {
    hipLaunchKernelGGL(dpc_adjoint_spectrum, dim3(blocks256((int64_t)g.Rp * g.Cp)), dim3(256), 0, s, g, G, out);
    return hipGetLastError();
}

hipError_t launch_dpc_vjp(const DpcGrid &g, const double *phi, const double *fx, const double *fy, const double2 *w,
                          double step, double g_err, double denom, double dd, const double *g_phi_out, double *g_phi,
                          double *g_fx, double *g_fy, hipStream_t s) {
    const int64_t tiles = (int64_t)((g.Cp + kVjCols - 1) / kVjCols) * ((g.Rp + kVjRows - 1) / kVjRows);
    hipLaunchKernelGGL(dpc_vjp, dim3((unsigned)tiles), dim3(kVjRows * kVjCols), 0, s, g, phi, fx, fy, w, step, g_err,
                       denom, dd, g_phi_out, g_phi, g_fx, g_fy);
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double *partial, int n, double *out, hipStream_t s) {
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, s, partial, n, out);
    return hipGetLastError();
}

hipError_t launch_crop(const double *padded, int R, int C, int Cp, double *out, hipStream_t s) {
    hipLaunchKernelGGL(crop, dim3(blocks256((int64_t)R * C)), dim3(256), 0, s, padded, R, C, Cp, out);
    return hipGetLastError();
}

hipError_t launch_embed(const double *scan, int R, int C, int Rp, int Cp, double *padded, hipStream_t s) {
    hipLaunchKernelGGL(embed, dim3(blocks256((int64_t)Rp * Cp)), dim3(256), 0, s, scan, R, C, Rp, Cp, padded);
    return hipGetLastError();
}

}  // namespace cubo
