// origin_kernels.hip — per-position origins for gfx950: py4DSTEM's get_origin in Double and the iterated-window
// estimator in Float32 (the launch interface and the algorithms are summarised in origin_kernels.hpp).
//
// get_origin: one block per position at a time (blocks loop over the tile's positions). The axis-0 Gaussian pass
// goes to the block's scratch pattern in HBM (L2-resident), the axis-1 pass is fused with the argmax, and the
// centre of mass within r * rscale of it is a compensated (TwoSum) sum over the block. Every filtered value is the
// reference's operation sequence (the centre weight times the pixel, then (p[i - t] + p[i + t]) * w for t = R down
// to 1, no multiply-add contraction), so the argmax is the reference's.
//
// Iterated window: one wave per position. The coarse step gives each lane whole bin x bin blocks, summed in the
// reference's order (so the block sums and the first-brightest block are its bits); each refinement pass gives each
// lane whole window rows, and lane 0 adds the row sums in row order.
#include "origin_kernels.hpp"

#include <algorithm>
#include <climits>
#include <type_traits>

#pragma clang fp contract(off)

namespace cubo {

namespace {

#define CUBO_GLOBAL __attribute__((address_space(1)))
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kPyBlock = 256;
constexpr int kIWWaves = 4;

template <bool U16>
__device__ __forceinline__ double pix(const void *pat, int64_t i) {
    if constexpr (U16) return (double)((const uint16_t *)pat)[i];
    else return (double)((const float *)pat)[i];
}

// (s, c) += x, Knuth's TwoSum.
__device__ __forceinline__ void two_sum(double &s, double &c, double x) {
    const double t = s + x, bb = t - s;
    c += (s - (t - bb)) + (x - bb);
    s = t;
}

template <bool U16>
__global__ __launch_bounds__(kPyBlock) void py4_origin(Py4Launch a) {
    const int tid = threadIdx.x;
    const int qy = a.qy, qx = a.qx, R = a.R;
    const int64_t ndet = (int64_t)qy * qx;
    const double *w = a.weights;
    const double centre = w[R];
    double *sa = a.scratch + (int64_t)blockIdx.x * ndet;
    __shared__ double red_v[kPyBlock];
    __shared__ int64_t red_i[kPyBlock], red_n[kPyBlock];
    __shared__ double red_s[6][kPyBlock];
    __shared__ int64_t best_at;
    const int64_t count = a.list ? (int64_t)*a.nlist : a.npos;
    for (int64_t k = blockIdx.x; k < count; k += gridDim.x) {
        const int64_t pos = a.list ? (int64_t)a.list[k] : k;
        if (a.take && !a.take[pos]) {
            if (tid == 0) { a.out_row[pos] = __builtin_nan(""); a.out_col[pos] = __builtin_nan(""); }
            continue;   // uniform across the block
        }
        const void *pat = (const char *)a.cube + pos * ndet * (U16 ? 2 : 4);
        // axis 0 (rows), mode nearest
        for (int64_t idx = tid; idx < ndet; idx += kPyBlock) {
            const int i = (int)(idx / qx), j = (int)(idx % qx);
            double acc = pix<U16>(pat, idx) * centre;
            for (int t = R; t >= 1; --t) {
                const int ia = i - t < 0 ? 0 : i - t, ib = i + t >= qy ? qy - 1 : i + t;
                acc += (pix<U16>(pat, (int64_t)ia * qx + j) + pix<U16>(pat, (int64_t)ib * qx + j)) * w[R - t];
            }
            sa[idx] = acc;
        }
        __syncthreads();
        // axis 1 (columns) and the argmax: the first NaN if any, else the first maximum
        double bv = -__builtin_inf();
        int64_t bi = INT64_MAX, bn = INT64_MAX;
        for (int64_t idx = tid; idx < ndet; idx += kPyBlock) {
            const int i = (int)(idx / qx), j = (int)(idx % qx);
            const double *row = sa + (int64_t)i * qx;
            double acc = row[j] * centre;
            for (int t = R; t >= 1; --t) {
                const int ja = j - t < 0 ? 0 : j - t, jb = j + t >= qx ? qx - 1 : j + t;
                acc += (row[ja] + row[jb]) * w[R - t];
            }
            if (a.filtered && pos == 0) a.filtered[idx] = acc;
            if (__builtin_isnan(acc)) {
                if (bn == INT64_MAX) bn = idx;
            } else if (bi == INT64_MAX || acc > bv) {
                bv = acc;
                bi = idx;
            }
        }
        red_v[tid] = bv; red_i[tid] = bi; red_n[tid] = bn;
        __syncthreads();
        if (tid == 0) {
            double v = red_v[0];
            int64_t b = red_i[0], n = red_n[0];
            for (int k = 1; k < kPyBlock; ++k) {
                n = red_n[k] < n ? red_n[k] : n;
                if (red_i[k] != INT64_MAX && (b == INT64_MAX || red_v[k] > v || (red_v[k] == v && red_i[k] < b))) {
                    v = red_v[k];
                    b = red_i[k];
                }
            }
            best_at = n != INT64_MAX ? n : (b != INT64_MAX ? b : 0);
        }
        __syncthreads();
        const int64_t at = best_at;
        const double i0 = (double)(at / qx), j0 = (double)(at % qx);
        if (a.fast_center) {
            if (tid == 0) { a.out_row[pos] = i0; a.out_col[pos] = j0; }
            __syncthreads();
            continue;
        }
        // get_CoM(dp * mask): every pixel times its 0/1 weight, so a NaN anywhere gives NaN
        double s = 0, sc = 0, r = 0, rc = 0, c = 0, cc = 0;
        for (int64_t idx = tid; idx < ndet; idx += kPyBlock) {
            const double fi = (double)(idx / qx), fj = (double)(idx % qx);
            const double di = fi - i0, dj = fj - j0;
            const double v = pix<U16>(pat, idx) * (__builtin_sqrt(di * di + dj * dj) < a.limit ? 1.0 : 0.0);
            two_sum(s, sc, v);
            two_sum(r, rc, fi * v);
            two_sum(c, cc, fj * v);
        }
        red_s[0][tid] = s; red_s[1][tid] = sc; red_s[2][tid] = r; red_s[3][tid] = rc; red_s[4][tid] = c; red_s[5][tid] = cc;
        __syncthreads();
        if (tid == 0) {
            double S = 0, SC = 0, MR = 0, MRC = 0, MC = 0, MCC = 0;
            for (int k = 0; k < kPyBlock; ++k) {
                two_sum(S, SC, red_s[0][k]); SC += red_s[1][k];
                two_sum(MR, MRC, red_s[2][k]); MRC += red_s[3][k];
                two_sum(MC, MCC, red_s[4][k]); MCC += red_s[5][k];
            }
            const double tot = __builtin_isfinite(S) ? S + SC : S;
            const double mr = __builtin_isfinite(MR) ? MR + MRC : MR;
            const double mc = __builtin_isfinite(MC) ? MC + MCC : MC;
            a.out_row[pos] = mr / tot;
            a.out_col[pos] = mc / tot;
        }
        __syncthreads();
    }
}

// ---- get_origin with the pattern in LDS -------------------------------------------------------------------
//
// For patterns whose elements fit in LDS beside a strip of the axis-0 pass: the raw pattern is staged once, the
// axis-0 pass of S rows goes to a transposed strip T[j][i] (row stride S + 1 doubles, so lanes of consecutive
// columns write different banks), and the axis-1 pass reads it along lanes of consecutive rows. Each lane
// computes kPyV consecutive outputs along the filtered axis from two register windows of the line (the values
// left and right of the outputs), which slide kPyV taps at a time: two loads per kPyV outputs and taps instead of
// two per output and tap. The operation sequence of every output is the reference's.
constexpr int kPyV = 16;

// min(max(k, 0), hi) for hi >= 0 in one instruction (the compiler emits a max and a min: it cannot know 0 <= hi).
__device__ __forceinline__ int clamp_index(int k, int hi) {
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(k), "v"(hi));
    return r;
}

// acc[v] = the filtered values at o .. o + kPyV - 1 of a line x(0 .. n - 1) with nearest-edge extension: x(o) times
// the centre weight, then (x(o - t) + x(o + t)) * w[R - t] for t = R down to 1.
template <typename X>
__device__ __forceinline__ void filter_run(double (&acc)[kPyV], int o, int n, int R, const double *w, double centre,
                                           X x) {
    auto at = [&](int k) { return x(clamp_index(k, n - 1)); };
#pragma unroll
    for (int v = 0; v < kPyV; ++v) acc[v] = at(o + v) * centre;
    int t = R;
    if (t < 1) return;
    // the taps in groups t .. t - kPyV + 1 (the last group: t .. 1): left value of (output v, tap t - s) is
    // lb[v + s] = x(o - t + v + s), right rb[v - s + kPyV - 1] = x(o + t + v - s)
    double lb[2 * kPyV - 1], rb[2 * kPyV - 1];
#pragma unroll
    for (int k = 0; k < 2 * kPyV - 1; ++k) {
        lb[k] = at(o - t + k);
        rb[k] = at(o + t - kPyV + 1 + k);
    }
    for (;;) {
        const int count = t < kPyV ? t : kPyV;             // uniform
#pragma unroll
        for (int sidx = 0; sidx < kPyV; ++sidx) {
            if (sidx < count) {
                const double wt = w[R - t + sidx];
                // the sums, products and accumulations of the kPyV outputs as three independent batches, so
                // consecutive FP64 operations never wait on each other's results
                double pr[kPyV];
#pragma unroll
                for (int v = 0; v < kPyV; ++v) pr[v] = lb[v + sidx] + rb[v - sidx + kPyV - 1];
#pragma unroll
                for (int v = 0; v < kPyV; ++v) pr[v] = pr[v] * wt;
#pragma unroll
                for (int v = 0; v < kPyV; ++v) acc[v] += pr[v];
            }
        }
        t -= kPyV;
        if (t < 1) break;
#pragma unroll
        for (int k = 0; k < kPyV - 1; ++k) lb[k] = lb[k + kPyV];
#pragma unroll
        for (int k = 0; k < kPyV; ++k) lb[kPyV - 1 + k] = at(o - t + kPyV - 1 + k);
#pragma unroll
        for (int k = 2 * kPyV - 2; k >= kPyV; --k) rb[k] = rb[k - kPyV];
#pragma unroll
        for (int k = 0; k < kPyV; ++k) rb[k] = at(o + t - kPyV + 1 + k);
    }
}

// (v, b, n) <- the merge of two argmax states: the smallest NaN index n; of the non-NaN maxima the larger value, on a
// tie the smaller index b (INT_MAX: none).
__device__ __forceinline__ void argmax_merge(double &v, int &b, int &n, double v2, int b2, int n2) {
    n = n2 < n ? n2 : n;
    if (b2 != INT_MAX && (b == INT_MAX || v2 > v || (v2 == v && b2 < b))) {
        v = v2;
        b = b2;
    }
}

__device__ __forceinline__ void argmax_merge_wave(double &v, int &b, int &n) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const double v2 = __shfl_xor(v, m);
        const int b2 = __shfl_xor(b, m), n2 = __shfl_xor(n, m);
        argmax_merge(v, b, n, v2, b2, n2);
    }
}

// (s, c) summed over the wave: s by TwoSum, the corrections and the rounding errors into c.
__device__ __forceinline__ void two_sum_wave(double &s, double &c) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const double s2 = __shfl_xor(s, m), c2 = __shfl_xor(c, m);
        two_sum(s, c, s2);
        c += c2;
    }
}

// POW2: qx = 1 << lqx (row offsets by shifts); S = 1 << lS rows per strip.
template <bool U16, bool POW2>
__global__ __launch_bounds__(kPyBlock) void py4_origin_lds(Py4Launch a, int lS, int lqx) {
    using T = typename std::conditional<U16, uint16_t, float>::type;
    extern __shared__ double py_lds[];
    const int tid = threadIdx.x;
    const int qy = a.qy, qx = a.qx, R = a.R, ndet = qy * qx, S = 1 << lS, SP = S + 1;
    auto col_at = [&](int k) { return (k << lS) + k; };                  // k * (S + 1)
    constexpr int lT = sizeof(T) == 2 ? 1 : 2;
    const unsigned qxb = (unsigned)qx * sizeof(T), sp8 = (unsigned)SP * 8;   // row strides in bytes (< 2^24)
    T *raw = (T *)py_lds;                                  // [ndet], padded to 16 bytes
    double *tr = py_lds + ((ndet * (int)sizeof(T) + 15) & ~15) / 8;   // [qx][S + 1]: axis 0, transposed
    double *wl = tr + qx * SP;                             // [2R + 1]
    constexpr int kWaves = kPyBlock / 64;
    __shared__ double red_v[kWaves];
    __shared__ int red_i[kWaves], red_n[kWaves];
    __shared__ double red_s[6][kWaves];
    __shared__ int best_at;
    const int lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < 2 * R + 1; i += kPyBlock) wl[i] = a.weights[i];
    const double centre = a.weights[R];
    const bool vec = (ndet * (int)sizeof(T)) % 16 == 0;
    const int64_t count = a.list ? (int64_t)*a.nlist : a.npos;
    for (int64_t k = blockIdx.x; k < count; k += gridDim.x) {
        const int64_t pos = a.list ? (int64_t)a.list[k] : k;
        if (a.take && !a.take[pos]) {
            if (tid == 0) { a.out_row[pos] = __builtin_nan(""); a.out_col[pos] = __builtin_nan(""); }
            continue;   // uniform across the block
        }
        const T *pat = (const T *)((const char *)a.cube + pos * (int64_t)ndet * (int64_t)sizeof(T));
        if (vec) {
            const int nv = ndet * (int)sizeof(T) / 16;
            for (int v = tid; v < nv; v += kPyBlock)
                ((u32x4 *)raw)[v] = __builtin_nontemporal_load((const CUBO_GLOBAL u32x4 *)pat + v);
        } else {
            for (int v = tid; v < ndet; v += kPyBlock) raw[v] = pat[v];
        }
        __syncthreads();
        double bv = -__builtin_inf();
        int bi = INT_MAX, bn = INT_MAX;
        for (int s0 = 0; s0 < qy; s0 += S) {
            const int rows = qy - s0 < S ? qy - s0 : S;
            // axis 0: item (row group g, column j), lanes along j
            const int groups = (rows + kPyV - 1) / kPyV;
            for (int it = tid; it < groups * qx; it += kPyBlock) {
                const int g = it / qx, j = it - g * qx;
                double acc[kPyV];
                // byte addresses in one shift-add (POW2) or multiply-add per load
                const char *rj = (const char *)(raw + j);
                filter_run(acc, s0 + g * kPyV, qy, R, wl, centre, [&](int k) {
                    return (double)*(const T *)(rj + (POW2 ? k << (lqx + lT) : (int)__umul24(k, qxb)));
                });
#pragma unroll
                for (int v = 0; v < kPyV; ++v)
                    if (g * kPyV + v < rows) tr[col_at(j) + g * kPyV + v] = acc[v];
            }
            __syncthreads();
            // axis 1 and the argmax: item (column group c, strip row il), lanes along il
            const int cgroups = (qx + kPyV - 1) / kPyV;
            for (int it = tid; it < cgroups * rows; it += kPyBlock) {
                const int c = it / rows, il = it - c * rows;
                double acc[kPyV];
                const char *ti = (const char *)(tr + il);
                filter_run(acc, c * kPyV, qx, R, wl, centre,
                           [&](int k) { return *(const double *)(ti + __umul24(k, sp8)); });
#pragma unroll
                for (int v = 0; v < kPyV; ++v) {
                    const int j = c * kPyV + v;
                    if (j < qx) {
                        const int idx = (s0 + il) * qx + j;
                        if (a.filtered && pos == 0) a.filtered[idx] = acc[v];
                        if (__builtin_isnan(acc[v])) {
                            bn = idx < bn ? idx : bn;
                        } else if (bi == INT_MAX || acc[v] > bv || (acc[v] == bv && idx < bi)) {
                            bv = acc[v];
                            bi = idx;
                        }
                    }
                }
            }
            __syncthreads();
        }
        // the first NaN, else the largest value with the smallest index: the same result in any merge order
        argmax_merge_wave(bv, bi, bn);
        if (lane == 0) { red_v[wave] = bv; red_i[wave] = bi; red_n[wave] = bn; }
        __syncthreads();
        if (tid == 0) {
            double v = red_v[0];
            int b = red_i[0], n = red_n[0];
            for (int k = 1; k < kWaves; ++k) argmax_merge(v, b, n, red_v[k], red_i[k], red_n[k]);
            best_at = n != INT_MAX ? n : (b != INT_MAX ? b : 0);
        }
        __syncthreads();
        const int at = best_at;
        const double i0 = (double)(at / qx), j0 = (double)(at % qx);
        if (a.fast_center) {
            if (tid == 0) { a.out_row[pos] = i0; a.out_col[pos] = j0; }
            __syncthreads();
            continue;
        }
        // get_CoM(dp * mask), pixels tid, tid + block, ... per thread as in py4_origin
        double s = 0, sc = 0, r = 0, rc = 0, c = 0, cc = 0;
        for (int idx = tid; idx < ndet; idx += kPyBlock) {
            const int ii = idx / qx;
            const double fi = (double)ii, fj = (double)(idx - ii * qx);
            const double di = fi - i0, dj = fj - j0;
            const double v = (double)raw[idx] * (__builtin_sqrt(di * di + dj * dj) < a.limit ? 1.0 : 0.0);
            two_sum(s, sc, v);
            two_sum(r, rc, fi * v);
            two_sum(c, cc, fj * v);
        }
        // compensated pairs merged across the wave (TwoSum of the sums, corrections added), then the waves in order
        two_sum_wave(s, sc);
        two_sum_wave(r, rc);
        two_sum_wave(c, cc);
        if (lane == 0) {
            red_s[0][wave] = s; red_s[1][wave] = sc; red_s[2][wave] = r;
            red_s[3][wave] = rc; red_s[4][wave] = c; red_s[5][wave] = cc;
        }
        __syncthreads();
        if (tid == 0) {
            double S0 = 0, SC = 0, MR = 0, MRC = 0, MC = 0, MCC = 0;
            for (int k = 0; k < kWaves; ++k) {
                two_sum(S0, SC, red_s[0][k]); SC += red_s[1][k];
                two_sum(MR, MRC, red_s[2][k]); MRC += red_s[3][k];
                two_sum(MC, MCC, red_s[4][k]); MCC += red_s[5][k];
            }
            const double tot = __builtin_isfinite(S0) ? S0 + SC : S0;
            const double mr = __builtin_isfinite(MR) ? MR + MRC : MR;
            const double mc = __builtin_isfinite(MC) ? MC + MCC : MC;
            a.out_row[pos] = mr / tot;
            a.out_col[pos] = mc / tot;
        }
        __syncthreads();
    }
}

// Strip rows and dynamic LDS bytes of py4_origin_lds for a detector (the largest strip that leaves room for two
// blocks per CU, else for one), or 0 rows when the pattern does not fit.
int py4_lds_strip(int qy, int qx, int R, bool u16, int *bytes) {
    const int64_t raw = ((int64_t)qy * qx * (u16 ? 2 : 4) + 15) & ~(int64_t)15;
    for (const int64_t budget : {(int64_t)78 * 1024, (int64_t)156 * 1024})
        for (int S = 32; S >= 8; S /= 2) {
            const int64_t b = (int64_t)qx * (S + 1) * 8 + (int64_t)(2 * R + 1) * 8 + raw;
            if (b <= budget) {
                *bytes = (int)b;
                return S;
            }
        }
    return 0;
}

// ---- get_origin on the matrix cores: the filter approximated, the argmax's candidates recomputed exactly --------
//
// uint16 patterns of 128 x 128 pixels. The separable filter is two matrix products, g = G P G^T, with G the
// 128 x 128 matrix of the weights (mode nearest folds the taps past an edge onto the edge pixel). On the matrix
// cores: the pattern in its exact FP16 split (x & 0xF800) + (x & 0x07FF), G scaled by a power of two and split into
// FP16 hi + lo (lo parts below the FP16 normal range zeroed on the host), FP32 sums. The first product
// B1 = P G^T leaves the accumulators in the layout the second product's B operand needs (the second product's K
// order is permuted to match, in G's second stored copy); they are scaled by a power of two and split into FP16
// hi + lo again. Every term is non-negative, so each approximate value g~ is within a relative e of its exact value
// g, plus an absolute slack for second-pass inputs below the FP16 normal range: the two FP32 sums of at most 512
// terms each (2 x 512 x 2^-24), G's representation (measured on the host), B1's split and the dropped lo x lo
// product. The pixel the reference picks has g = max g >= M~ / (1 + e), so its g~ >= M~ (1 - e) / (1 + e) - slack
// > M~ (1 - 2e) - slack: the pixels at or above M~ (1 - margin) - slack, with margin >= 2e set on the host, contain
// it. They (a few pixels around the peak of the smooth filtered pattern) are recomputed in Double in the reference's
// operation order, the axis-0 values of their rows and then the axis-1 sum, and the first largest wins: the argmax
// is the reference's. More than kPyCand candidates, or no positive value (an empty pattern), leave the pattern to
// the exact kernel (the undecided list).
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));
typedef unsigned short us2 __attribute__((ext_vector_type(2)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
constexpr int kPyN = kPyMfmaSide;             // detector side
constexpr int kPyNB = kPyN / 32;              // 32-pixel blocks per side (and waves per block)
constexpr int kPyKS = kPyN / 16;              // k-steps of one product
constexpr int kPyCand = 32;                   // candidates recomputed exactly
constexpr int kPyRows = 4;                    // distinct candidate rows whose axis-0 values are held at once
constexpr int kPyTaps = 511;                  // weights held in LDS (2R + 1 at most: R <= 255)

__device__ __forceinline__ uint32_t py_cvt_pair(uint32_t w) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(__builtin_bit_cast(us2, w), h2));
}

__global__ __launch_bounds__(64 * kPyNB) __attribute__((amdgpu_waves_per_eu(1, 1))) void py4_origin_mfma(Py4Launch a) {
    constexpr int N = kPyN, NB = kPyNB, KS = kPyKS;
    // pattern rows padded by 16 bytes: the first product's A operands (16 bytes of 32 consecutive rows at one column)
    // then fall on distinct banks
    constexpr int NP = N + 8;
    __shared__ half8 g2s[2][NB][KS][64];                       // the second product's A operands (hi, lo)
    __shared__ __align__(16) uint16_t pat[2][N * NP];          // the pattern, double-buffered
    __shared__ double arow[kPyRows][N];                        // exact axis-0 values of candidate rows
    __shared__ float wmax[NB];
    __shared__ int cand[kPyCand], cand_n, rows_of[kPyRows];
    __shared__ double cand_g[kPyCand];
    __shared__ double red_s[3][NB];
    __shared__ double w[kPyTaps];                              // the weights, for the exact recomputation
    const int tid = threadIdx.x, lane = tid & 63, n = lane & 31, h = lane >> 5;
    const int J = __builtin_amdgcn_readfirstlane(tid >> 6);    // this wave's 32-column block
    const int R = a.R;
    for (int t = tid; t < 2 * R + 1; t += 64 * NB) w[t] = a.weights[t];
    const double centre = a.weights[R];

    // the first product's B operands (rows 32 J + n of G, columns 16 k + 8 h ..), kept in registers
    half8 gb[KS][2];
#pragma unroll
    for (int k = 0; k < KS; ++k)
#pragma unroll
        for (int p = 0; p < 2; ++p)
            gb[k][p] = *(const half8 *)(a.gmat + (size_t)p * N * N + (size_t)(32 * J + n) * N + 16 * k + 8 * h);
    for (int e = tid; e < 2 * NB * KS * 64; e += 64 * NB) {
        const int l = e & 63, k = (e >> 6) % KS, I = (e / (64 * KS)) % NB, p = e / (64 * KS * NB);
        g2s[p][I][k][l] = *(const half8 *)(a.gmat + (size_t)(2 + p) * N * N + (size_t)(32 * I + (l & 31)) * N +
                                           16 * k + 8 * (l >> 5));
    }

    // positions blockIdx.x, + gridDim.x, ... that are taken
    auto next_taken = [&](int64_t k) {
        while (k < a.npos && a.take && !a.take[k]) {
            a.out_row[k] = __builtin_nan("");                   // every thread writes the same value
            a.out_col[k] = __builtin_nan("");
            k += gridDim.x;
        }
        return k;
    };
    constexpr int PV = N * N / 8 / (64 * NB);                  // 16-byte pieces per thread
    u32x4 pf[PV];
    auto fetch = [&](int64_t pos) {
        const u32x4 *src = (const u32x4 *)((const uint16_t *)a.cube + pos * (int64_t)(N * N));
#pragma unroll
        for (int v = 0; v < PV; ++v) pf[v] = __builtin_nontemporal_load((const CUBO_GLOBAL u32x4 *)src + tid + v * 64 * NB);
    };
    // the pattern from registers into LDS buffer b; whether any count needs the FP16 high part (the block's OR)
    auto stage = [&](int b) {
        uint32_t any = 0;
#pragma unroll
        for (int v = 0; v < PV; ++v) {
            const int e = tid + v * 64 * NB;                   // 16-byte piece e % 16 of row e / 16
            *(u32x4 *)&pat[b][(e >> 4) * NP + 8 * (e & 15)] = pf[v];
            any |= (pf[v].x | pf[v].y | pf[v].z | pf[v].w) & 0xf800f800u;
        }
        return __syncthreads_or(any != 0) != 0;
    };
    // B1 = P G^T for this wave's 32 columns: acc1[I] rows 32 I + 8 (r / 4) + 4 h + r % 4, column 32 J + n. The row
    // blocks' accumulations interleave (k outer), so consecutive matrix instructions do not depend on each other.
    auto pass1 = [&](const uint16_t *P, float16v (&acc1)[NB], auto hi_tag) {
        constexpr bool HI = decltype(hi_tag)::value;
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc1[I][r] = 0.f;
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            u32x4 x[NB];
#pragma unroll
            for (int I = 0; I < NB; ++I) x[I] = *(const u32x4 *)&P[(32 * I + n) * NP + 16 * k + 8 * h];
#pragma unroll
            for (int I = 0; I < NB; ++I) {
                u32x4 lo;
                lo.x = py_cvt_pair(x[I].x & 0x07ff07ffu); lo.y = py_cvt_pair(x[I].y & 0x07ff07ffu);
                lo.z = py_cvt_pair(x[I].z & 0x07ff07ffu); lo.w = py_cvt_pair(x[I].w & 0x07ff07ffu);
                const half8 alo = __builtin_bit_cast(half8, lo);
                acc1[I] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, gb[k][0], acc1[I], 0, 0, 0);
                acc1[I] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, gb[k][1], acc1[I], 0, 0, 0);
                if constexpr (HI) {
                    u32x4 hi;
                    hi.x = py_cvt_pair(x[I].x & 0xf800f800u); hi.y = py_cvt_pair(x[I].y & 0xf800f800u);
                    hi.z = py_cvt_pair(x[I].z & 0xf800f800u); hi.w = py_cvt_pair(x[I].w & 0xf800f800u);
                    const half8 ahi = __builtin_bit_cast(half8, hi);
                    acc1[I] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, gb[k][0], acc1[I], 0, 0, 0);
                    acc1[I] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, gb[k][1], acc1[I], 0, 0, 0);
                }
            }
        }
    };
    auto undecided = [&](int64_t pos) {
        if (tid == 0) a.undecided[atomicAdd(a.n_undecided, 1)] = (int32_t)pos;
    };

    int64_t pos = next_taken(blockIdx.x);
    int cur = 0;
    bool phi = false;                                          // the current pattern has counts of 2048 or more
    if (pos < a.npos) {
        fetch(pos);
        phi = stage(0);
    }
    while (pos < a.npos) {
        const int64_t nxt = next_taken(pos + gridDim.x);
        if (nxt < a.npos) fetch(nxt);
        const uint16_t *P = pat[cur];

        float16v acc1[NB];
        if (phi) pass1(P, acc1, std::true_type{});
        else pass1(P, acc1, std::false_type{});
        // g~ = G B1: k-step k takes rows 32 (k / 2) + 16 (k % 2) + 8 (e / 4) + 4 h + e % 4 of B1 (registers
        // 8 (k % 2) + e of acc1[k / 2]); G's second copy has its columns in that order
        float16v acc2[NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc2[I][r] = 0.f;
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            half8 bh, bl;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float v = acc1[k / 2][8 * (k % 2) + e] * a.pass2_scale;
                const _Float16 hv = (_Float16)v;
                bh[e] = hv;
                bl[e] = (_Float16)(v - (float)hv);
            }
#pragma unroll
            for (int I = 0; I < NB; ++I) {
                const half8 ah = g2s[0][I][k][lane], al = g2s[1][I][k][lane];
                acc2[I] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc2[I], 0, 0, 0);
                acc2[I] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc2[I], 0, 0, 0);
                acc2[I] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc2[I], 0, 0, 0);
            }
        }
        // the largest approximate value, then the candidates at or above the threshold (only waves holding some
        // look at their values one by one)
        float m = 0.f;
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int r = 0; r < 16; ++r) m = __builtin_fmaxf(m, acc2[I][r]);
        float wm = m;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) wm = __builtin_fmaxf(wm, __shfl_xor(wm, off, 64));
        if (lane == 0) wmax[J] = wm;
        if (tid == 0) cand_n = 0;
        __syncthreads();
        float mx = wmax[0];
#pragma unroll
        for (int k = 1; k < NB; ++k) mx = __builtin_fmaxf(mx, wmax[k]);
        const float th = mx * (1.0f - a.candidate_margin) - a.slack;
        if (mx > 0.f && th > 0.f && __builtin_amdgcn_ballot_w64(m >= th) != 0) {
            uint64_t bits = 0;                                 // bit 16 I + r: acc2[I][r] >= th
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int r = 0; r < 16; ++r) bits |= acc2[I][r] >= th ? (uint64_t)1 << (16 * I + r) : 0;
            while (bits) {
                const int b = __builtin_ctzll(bits), I = b >> 4, r = b & 15;
                bits &= bits - 1;
                const int i = 32 * I + 8 * (r >> 2) + 4 * h + (r & 3), j = 32 * J + n;
                const int slot = atomicAdd(&cand_n, 1);
                if (slot < kPyCand) cand[slot] = i * N + j;
            }
        }
        __syncthreads();
        const int nc = cand_n;
        // the candidates' distinct rows (in candidate order; past kPyRows a repeat may count again: the pattern goes
        // to the exact kernel either way), by every thread from LDS
        int nr = 0;
        if (nc >= 1 && nc <= kPyCand) {
            for (int c = 0; c < nc && nr <= kPyRows; ++c) {
                const int i = cand[c] / N;
                bool seen = false;
                for (int q = 0; q < min(nr, kPyRows); ++q) seen = seen || rows_of[q] == i;
                if (!seen) {
                    if (nr < kPyRows && tid == 0) rows_of[nr] = i;
                    ++nr;
                }
                if (!seen && nr <= kPyRows) __syncthreads();   // every thread sees the row it just added
            }
        }
        if (nc == 0 || nc > kPyCand || nr > kPyRows) {
            // no positive value, or too many near the top: the exact kernel measures this pattern
            undecided(pos);
        } else {
            // axis 0 in the reference's order: a(i, j') = p(i, j') c + sum_{t = R..1} (p(i - t, j') + p(i + t, j')) w[R - t]
            // (a wave's 64 lanes share one row q: the taps' rows and weights are uniform)
            const uint16_t *Pc = pat[cur];
            for (int e = tid; e < nr * N; e += 64 * NB) {
                const int q = __builtin_amdgcn_readfirstlane(e / N), jj = e - q * N;
                const int i = __builtin_amdgcn_readfirstlane(rows_of[q]);
                // (the two counts added as integers, exact: Double(p_a) + Double(p_b) to the bit; eight taps' terms
                // formed together, then added one by one in order)
                double acc = (double)Pc[i * NP + jj] * centre;
                int t = R;
                for (; t >= 8; t -= 8) {
                    double pr[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int tt = t - u, ia = i - tt < 0 ? 0 : i - tt, ib = i + tt >= N ? N - 1 : i + tt;
                        pr[u] = (double)((uint32_t)Pc[ia * NP + jj] + (uint32_t)Pc[ib * NP + jj]) * w[R - tt];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) acc += pr[u];
                }
                for (; t >= 1; --t) {
                    const int ia = i - t < 0 ? 0 : i - t, ib = i + t >= N ? N - 1 : i + t;
                    acc += (double)((uint32_t)Pc[ia * NP + jj] + (uint32_t)Pc[ib * NP + jj]) * w[R - t];
                }
                arow[q][jj] = acc;
            }
            __syncthreads();
            // axis 1: one thread per candidate
            if (tid < nc) {
                const int i = cand[tid] / N, j = cand[tid] - i * N;
                int q = 0;
                while (rows_of[q] != i) ++q;
                const double *ar = arow[q];
                double acc = ar[j] * centre;
                int t = R;
                for (; t >= 8; t -= 8) {
                    double pr[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int tt = t - u, ja = j - tt < 0 ? 0 : j - tt, jb = j + tt >= N ? N - 1 : j + tt;
                        pr[u] = (ar[ja] + ar[jb]) * w[R - tt];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) acc += pr[u];
                }
                for (; t >= 1; --t) {
                    const int ja = j - t < 0 ? 0 : j - t, jb = j + t >= N ? N - 1 : j + t;
                    acc += (ar[ja] + ar[jb]) * w[R - t];
                }
                cand_g[tid] = acc;
            }
            __syncthreads();
            // the largest value, the smallest index among equal ones (uint16 patterns: no NaN), by every thread
            int at = cand[0];
            double v = cand_g[0];
            for (int c = 1; c < nc; ++c)
                if (cand_g[c] > v || (cand_g[c] == v && cand[c] < at)) { v = cand_g[c]; at = cand[c]; }
            const int ai = at / N, aj = at - ai * N;
            const double i0 = (double)ai, j0 = (double)aj;
            if (a.fast_center) {
                if (tid == 0) { a.out_row[pos] = i0; a.out_col[pos] = j0; }
            } else {
                // get_CoM(dp * mask). The pixels outside the disk add exact zeros, so only its bounding box is
                // read; for integer counts every partial sum is an integer below 2^53, so plain Double adds in any
                // order give the exact sums, as the reference's compensated sums do.
                const int ext = (int)__builtin_ceil(a.limit) + 1;
                const int r0 = max(0, ai - ext), r1 = min(N - 1, ai + ext), c0 = max(0, aj - ext), c1 = min(N - 1, aj + ext);
                const int bw = c1 - c0 + 1, cells = (r1 - r0 + 1) * bw;
                double s0 = 0, s1 = 0, s2 = 0;
                for (int e = tid; e < cells; e += 64 * NB) {
                    const int ii = r0 + e / bw, jj = c0 + e % bw;
                    const double fi = (double)ii, fj = (double)jj;
                    const double di = fi - i0, dj = fj - j0;
                    const double v = (double)P[ii * NP + jj] * (__builtin_sqrt(di * di + dj * dj) < a.limit ? 1.0 : 0.0);
                    s0 += v;
                    s1 += fi * v;
                    s2 += fj * v;
                }
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    s0 += __shfl_xor(s0, off, 64);
                    s1 += __shfl_xor(s1, off, 64);
                    s2 += __shfl_xor(s2, off, 64);
                }
                if (lane == 0) { red_s[0][J] = s0; red_s[1][J] = s1; red_s[2][J] = s2; }
                __syncthreads();
                if (tid == 0) {
                    double S0 = 0, MR = 0, MC = 0;
                    for (int k = 0; k < NB; ++k) { S0 += red_s[0][k]; MR += red_s[1][k]; MC += red_s[2][k]; }
                    a.out_row[pos] = MR / S0;
                    a.out_col[pos] = MC / S0;
                }
            }
        }
        // the next pattern into the other buffer (its last readers finished before the barriers above)
        if (nxt < a.npos) phi = stage(cur ^ 1);
        cur ^= 1;
        pos = nxt;
    }
}

// ---- iterated window (Float32) ----

__device__ __forceinline__ bool non_finite(float v) {
    return (__builtin_bit_cast(uint32_t, v) & 0x7F800000u) == 0x7F800000u;
}

template <bool U16>
__device__ __forceinline__ float raw(const void *pat, int64_t i) {
    if constexpr (U16) return (float)((const uint16_t *)pat)[i];
    else return ((const float *)pat)[i];
}

// The pixel, or for a non-finite one the median of its finite 8 neighbours of the original pattern (0 when none),
// with the reference's insertion order (IteratedWindowOrigin.swift:48-70).
template <bool U16>
__device__ float filled(const void *pat, int row, int col, int qy, int qx) {
    const float v = raw<U16>(pat, (int64_t)row * qx + col);
    if constexpr (U16) {
        return v;
    } else {
        if (!non_finite(v)) return v;
        float nb[8];
        int k = 0;
        for (int dr = -1; dr <= 1; ++dr)
            for (int dc = -1; dc <= 1; ++dc) {
                if (dr == 0 && dc == 0) continue;
                const int rr = row + dr, cc = col + dc;
                if (rr < 0 || rr >= qy || cc < 0 || cc >= qx) continue;
                const float w = raw<U16>(pat, (int64_t)rr * qx + cc);
                if (non_finite(w)) continue;
                int i = k++;
                while (i > 0 && nb[i - 1] > w) { nb[i] = nb[i - 1]; --i; }
                nb[i] = w;
            }
        if (k == 0) return 0.0f;
        return (k % 2 == 1) ? nb[k / 2] : 0.5f * (nb[k / 2 - 1] + nb[k / 2]);
    }
}

// One position of the iterated window (IteratedWindowOrigin.swift:136-241) by one wave: the pattern at pat (global
// memory or the wave's copy in LDS), at most rows_per_wave window rows; returns the origin.
template <bool U16>
__device__ __forceinline__ void iw_position(const IWLaunch &a, const void *pat, int rows_per_wave, int lane,
                                            float *out_r, float *out_c) {
    const int qy = a.qy, qx = a.qx;
    // coarse: the brightest bin x bin block (the first in raster order), its centre
    const float r = a.radius;
    const int bin = a.binned ? max(1, (int)__builtin_roundf(r)) : 1;
    const int nby = (qy + bin - 1) / bin, nbx = (qx + bin - 1) / bin;
    float best = -3.402823466e38f;
    int bestb = INT32_MAX;
    for (int b = lane; b < nby * nbx; b += 64) {
        const int by = (b / nbx) * bin, bx = (b % nbx) * bin;
        const int yEnd = min(by + bin, qy), xEnd = min(bx + bin, qx);
        float s = 0.0f;
        for (int y = by; y < yEnd; ++y) {
            // four pixels' reads issued together, added in raster order
            int x = bx;
            for (; x + 4 <= xEnd; x += 4) {
                const float f0 = filled<U16>(pat, y, x, qy, qx), f1 = filled<U16>(pat, y, x + 1, qy, qx),
                            f2 = filled<U16>(pat, y, x + 2, qy, qx), f3 = filled<U16>(pat, y, x + 3, qy, qx);
                s += f0;
                s += f1;
                s += f2;
                s += f3;
            }
            for (; x < xEnd; ++x) s += filled<U16>(pat, y, x, qy, qx);
        }
        if (s > best) { best = s; bestb = b; }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const float ov = __shfl_xor(best, off, 64);
        const int ob = __shfl_xor(bestb, off, 64);
        if (ob != INT32_MAX && (bestb == INT32_MAX || ov > best || (ov == best && ob < bestb))) { best = ov; bestb = ob; }
    }
    float cr = 0.5f * (float)(qy - 1), cc = 0.5f * (float)(qx - 1);
    if (bestb != INT32_MAX) {
        const int by = (bestb / nbx) * bin, bx = (bestb % nbx) * bin;
        const int yEnd = min(by + bin, qy), xEnd = min(bx + bin, qx);
        cc = 0.5f * (float)(bx + xEnd - 1);
        cr = 0.5f * (float)(by + yEnd - 1);
    }
    // refinement passes
    const float rs = r * a.radius_scale, rm = r + a.margin;
    const float win = rm >= rs ? rm : rs;
    const float win2 = win * win;
    for (int pass = 0; pass < a.passes; ++pass) {
        const int y0 = max(0, (int)__builtin_floorf(cr - win));
        const int y1 = min(qy - 1, (int)__builtin_ceilf(cr + win));
        const int x0 = max(0, (int)__builtin_floorf(cc - win));
        const int x1 = min(qx - 1, (int)__builtin_ceilf(cc + win));
        const int nrows = min(y1 >= y0 && x1 >= x0 ? y1 - y0 + 1 : 0, rows_per_wave);
        float tI = 0.0f, tX = 0.0f, tY = 0.0f;        // this lane's rows k, k + 64, ..., added in that order
        for (int k = lane; k < nrows; k += 64) {
            const int y = y0 + k;
            float sI = 0.0f, sX = 0.0f, sY = 0.0f;
            const float dy = (float)y - cr;
            const float dy2 = dy * dy;
            if (dy2 <= win2) {
                // pixels inside the window, in column order; four reads issued together
                auto add = [&](int x, float f) {
                    const float dx = (float)x - cc;
                    if (__builtin_fmaf(dx, dx, dy2) > win2) return;
                    const float v = 0.0f >= f ? 0.0f : f;
                    sI += v;
                    sX = __builtin_fmaf(v, (float)x, sX);
                    sY = __builtin_fmaf(v, (float)y, sY);
                };
                int x = x0;
                for (; x + 4 <= x1 + 1; x += 4) {
                    const float f0 = filled<U16>(pat, y, x, qy, qx), f1 = filled<U16>(pat, y, x + 1, qy, qx),
                                f2 = filled<U16>(pat, y, x + 2, qy, qx), f3 = filled<U16>(pat, y, x + 3, qy, qx);
                    add(x, f0);
                    add(x + 1, f1);
                    add(x + 2, f2);
                    add(x + 3, f3);
                }
                for (; x <= x1; ++x) add(x, filled<U16>(pat, y, x, qy, qx));
            }
            tI += sI; tX += sX; tY += sY;
        }
        // the row sums added across the wave in a fixed tree (lane 0's result, broadcast): the reference adds all
        // window pixels in one running sum, so the totals differ from it by Float32 rounding only
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            tI += __shfl_down(tI, off, 64);
            tX += __shfl_down(tX, off, 64);
            tY += __shfl_down(tY, off, 64);
        }
        const float sumI = __shfl(tI, 0, 64), sumIX = __shfl(tX, 0, 64), sumIY = __shfl(tY, 0, 64);
        if (!(sumI > 0.0f)) break;
        const float nc = sumIX / sumI, nr = sumIY / sumI;
        const bool settled = __builtin_fabsf(nc - cc) < 1e-4f && __builtin_fabsf(nr - cr) < 1e-4f;
        cc = nc;
        cr = nr;
        if (settled) break;
    }
    *out_r = cr;
    *out_c = cc;
}

template <bool U16>
__global__ __launch_bounds__(64 * kIWWaves) void iterated_window(IWLaunch a, int rows_per_wave, int waves) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t pos = (int64_t)blockIdx.x * waves + wave;
    if (pos >= a.npos) return;                 // whole waves; no block barrier below
    const int64_t ndet = (int64_t)a.qy * a.qx;
    const void *pat = (const char *)a.cube + pos * ndet * (U16 ? 2 : 4);
    float cr, cc;
    iw_position<U16>(a, pat, rows_per_wave, lane, &cr, &cc);
    if (lane == 0) {
        a.out_row[pos] = (double)cr;
        a.out_col[pos] = (double)cc;
    }
}

}  // namespace

int py4_blocks(int cus, int64_t npos) {
    const int64_t b = (int64_t)cus * 4;
    return (int)(npos < b ? (npos > 0 ? npos : 1) : b);
}

hipError_t launch_py4_origin(const Py4Launch &a, int blocks, hipStream_t s) {
    if (a.npos <= 0) return hipSuccess;
    int bytes = 0;
    const int S = py4_lds_strip(a.qy, a.qx, a.R, a.u16, &bytes);
    if (S > 0) {
        const int lS = __builtin_ctz((unsigned)S);
        const bool pow2 = (a.qx & (a.qx - 1)) == 0;
        const int lqx = pow2 ? __builtin_ctz((unsigned)a.qx) : 0;
        auto launch = [&](auto kernel) {
            const hipError_t e = hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kPyBlock), bytes, s, a, lS, lqx);
            return hipGetLastError();
        };
        if (a.u16) return pow2 ? launch(py4_origin_lds<true, true>) : launch(py4_origin_lds<true, false>);
        return pow2 ? launch(py4_origin_lds<false, true>) : launch(py4_origin_lds<false, false>);
    }
    if (a.u16) hipLaunchKernelGGL(py4_origin<true>, dim3(blocks), dim3(kPyBlock), 0, s, a);
    else hipLaunchKernelGGL(py4_origin<false>, dim3(blocks), dim3(kPyBlock), 0, s, a);
    return hipGetLastError();
}

bool py4_mfma_supported(const Py4Launch &a) {
    return a.u16 && a.qy == kPyN && a.qx == kPyN && !a.filtered && 2 * a.R + 1 <= kPyTaps;
}

hipError_t launch_py4_origin_mfma(const Py4Launch &a, int blocks, hipStream_t s) {
    if (!py4_mfma_supported(a) || !a.gmat || !a.undecided || !a.n_undecided || a.list) return hipErrorInvalidValue;
    if (a.npos <= 0) return hipSuccess;
    hipLaunchKernelGGL(py4_origin_mfma, dim3(blocks), dim3(64 * kPyNB), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_iterated_window(const IWLaunch &a, hipStream_t s) {
    if (a.npos <= 0) return hipSuccess;
    // window rows: at most the detector's rows
    const float rs = a.radius * a.radius_scale, rm = a.radius + a.margin;
    const double win = rm >= rs ? rm : rs;
    const int64_t span = (int64_t)(2.0 * win) + 3;
    const int rows = (int)(span < a.qy ? span : a.qy);
    const int waves = kIWWaves;
    const dim3 grid((unsigned)((a.npos + waves - 1) / waves));
    if (a.u16)
        hipLaunchKernelGGL(iterated_window<true>, grid, dim3(64 * waves), 0, s, a, rows, waves);
    else
        hipLaunchKernelGGL(iterated_window<false>, grid, dim3(64 * waves), 0, s, a, rows, waves);
    return hipGetLastError();
}

}  // namespace cubo
