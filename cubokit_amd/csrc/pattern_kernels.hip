// pattern_kernels.hip — selected-area, weighted and centre-shifted mean and maximum patterns for gfx950.
//
// Replaces the Metal kernel cubo_weighted_pattern (Sources/CuboVirtual/MetalReducers.swift:268-302: one thread
// per pixel striding the positions of a tile) and the .cpu loops it mirrors (SelectedDiffraction.swift:110-179,
// 225-271; ShiftedImaging.swift:125-167).
//
// Layout. Block b.x owns 64 * VEC consecutive pixels (64 lanes of VEC pixels, one 16-byte load per lane and
// position, so a wave reads 1 KiB contiguous per position); block b.y is a position group. The four waves of a
// block take the same pixels and consecutive quarters of the group's entries, in scan order, and are combined in
// LDS in that order before one partial per (group, pixel) is written. A second kernel combines the groups in
// order. Every entry's pattern is read once; positions outside the selection are not read.
//
// A rolled pattern (PatternShift.rounded): the output pixel (r, c) takes the pattern's pixel
// ((r - sr) mod qy, (c - sc) mod qx). With 16-byte vectors the lane loads the two aligned vectors that hold its
// VEC source pixels and picks them with the position's alignment offset, which is the same for every lane.
#include "pattern_kernels.hpp"

#include <algorithm>

// TwoSum and the Double products round as written: no multiply-add contraction in this file.
#pragma clang fp contract(off)

namespace cubo {

namespace {

#define CUBO_GLOBAL __attribute__((address_space(1)))
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short us2 __attribute__((ext_vector_type(2)));

constexpr int kPatWaves = 4;
constexpr int kPatBlock = 64 * kPatWaves;

__device__ __forceinline__ u32x4 load16(const char *p) {
    return __builtin_nontemporal_load((const CUBO_GLOBAL u32x4 *)p);
}

template <typename T>
__device__ __forceinline__ uint32_t load1(const char *p) {
    if constexpr (sizeof(T) == 2)
        return (uint32_t)__builtin_nontemporal_load((const CUBO_GLOBAL uint16_t *)p);
    else
        return __builtin_nontemporal_load((const CUBO_GLOBAL uint32_t *)p);
}

__device__ __forceinline__ uint32_t pk_max_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(us2, a), __builtin_bit_cast(us2, b)));
}

// Elements O .. O + VEC - 1 of the 32 bytes A:B, as a 16-byte vector (uint16: 8 elements, float: 4).
template <int O, bool U16>
__device__ __forceinline__ u32x4 take(const u32x4 &A, const u32x4 &B) {
    const uint32_t S[8] = {A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w};
    u32x4 r;
    if constexpr (U16) {
        constexpr int t = O / 2;
        if constexpr (O % 2 == 0) {
            r.x = S[t]; r.y = S[t + 1]; r.z = S[t + 2]; r.w = S[t + 3];
        } else {
            r.x = __builtin_amdgcn_alignbit(S[t + 1], S[t], 16);
            r.y = __builtin_amdgcn_alignbit(S[t + 2], S[t + 1], 16);
            r.z = __builtin_amdgcn_alignbit(S[t + 3], S[t + 2], 16);
            r.w = __builtin_amdgcn_alignbit(S[t + 4], S[t + 3], 16);
        }
    } else {
        r.x = S[O]; r.y = S[O + 1]; r.z = S[O + 2]; r.w = S[O + 3];
    }
    return r;
}

// o is uniform across the wave (it depends on the position's column shift only), so this is a scalar branch.
template <bool U16>
__device__ __forceinline__ u32x4 pick(const u32x4 &A, const u32x4 &B, int o) {
    switch (o) {
    case 0: return A;
    case 1: return take<1, U16>(A, B);
    case 2: return take<2, U16>(A, B);
    case 3: return take<3, U16>(A, B);
    case 4: if constexpr (U16) return take<4, U16>(A, B); else return A;
    case 5: if constexpr (U16) return take<5, U16>(A, B); else return A;
    case 6: if constexpr (U16) return take<6, U16>(A, B); else return A;
    default: if constexpr (U16) return take<7, U16>(A, B); else return A;
    }
}

// Element k of a loaded vector (VEC > 1) or the loaded element (VEC == 1), as Double.
template <typename T, int VEC>
__device__ __forceinline__ double elem(const u32x4 &v, int k) {
    uint32_t bits;
    if constexpr (VEC == 1) bits = v.x;
    else if constexpr (sizeof(T) == 2) bits = (v[k >> 1] >> (16 * (k & 1))) & 0xffffu;
    else bits = v[k];
    if constexpr (sizeof(T) == 2) return (double)bits;
    else return (double)__builtin_bit_cast(float, bits);
}

// The maximum step: NaN wins; of equal values the first (m) or the later (x) is kept.
template <bool FIRST, bool NAN_POSSIBLE>
__device__ __forceinline__ double max_step(double m, double x) {
    if constexpr (NAN_POSSIBLE) {
        if constexpr (FIRST) return (x > m || __builtin_isnan(x)) ? x : m;
        else return (x >= m || __builtin_isnan(x)) ? x : m;
    } else {
        if constexpr (FIRST) return x > m ? x : m;
        else return x >= m ? x : m;
    }
}

// (s, c) += x with Knuth's TwoSum: s the running sum, c the sum of the exact rounding errors.
__device__ __forceinline__ void two_sum(double &s, double &c, double x) {
    const double t = s + x;
    const double bb = t - s;
    c += (s - (t - bb)) + (x - bb);
    s = t;
}

template <typename T, int VEC, bool INT, int STAT, bool SHIFT, bool FIRST>
__global__ __launch_bounds__(kPatBlock) void pattern_kernel(PatternLaunch a) {
    constexpr bool U16 = sizeof(T) == 2;
    constexpr bool MEAN = STAT == kPatternMean, MAX = STAT == kPatternMax;
    constexpr int U = SHIFT ? 4 : 8;                  // positions in flight per lane
    constexpr int ES = (int)sizeof(T);
    constexpr bool NAN_POSSIBLE = !U16;               // uint16 times a finite weight is never NaN
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t ndet = a.ndet;
    const int64_t pix = ((int64_t)blockIdx.x * 64 + lane) * VEC;
    const bool live = pix < ndet;
    const int64_t lp = live ? pix : 0;                // lanes past the detector load pixel 0 and write nothing
    const int qy = a.qy, qx = a.qx;
    const int orow = SHIFT ? (int)(lp / qx) : 0, ocol = SHIFT ? (int)(lp % qx) : 0;

    const int group = blockIdx.y;
    const int64_t g0 = (int64_t)group * a.ppg, g1 = min(a.count, g0 + a.ppg);
    const int64_t len = g1 > g0 ? g1 - g0 : 0;
    const int64_t q = (len + kPatWaves - 1) / kPatWaves;
    const int64_t j0 = g0 + min(len, (int64_t)wave * q), j1 = g0 + min(len, (int64_t)(wave + 1) * q);

    const char *cube = (const char *)a.cube;
    const int32_t *list = a.list;
    const double *wts = a.w;
    const uint32_t *shp = a.shift;
    const int64_t pbytes = ndet * ES;

    // per-lane state
    uint32_t isum[VEC], imax[VEC];
    double s[VEC], c[VEC], m[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        isum[k] = 0u; imax[k] = 0u;
        s[k] = 0.0; c[k] = 0.0; m[k] = -__builtin_inf();
    }

    // Entry j: its position, weight (1 without weights: x * 1 is x, bit for bit) and the lane's vector; with a
    // roll the two aligned vectors that hold it and its alignment offset, picked only when consumed so that the
    // loads of all U entries are in flight first.
    struct In {
        u32x4 A, B;
        double w;
        int o;
    };
    auto fetch = [&](int64_t j, In &in) {
        const int64_t p = list ? (int64_t)list[a.first + j] : a.first + j;
        in.w = wts ? wts[p] : 1.0;
        in.o = 0;
        const char *base = cube + p * pbytes;
        if constexpr (!SHIFT) {
            if constexpr (VEC == 1) in.A.x = load1<T>(base + lp * ES);
            else in.A = load16(base + lp * ES);
        } else {
            const uint32_t sh = shp[p];
            const int sr = (int)(sh >> 16), sc = (int)(sh & 0xffffu);
            int r = orow - sr;
            r += r < 0 ? qy : 0;
            int u = ocol - sc;
            u += u < 0 ? qx : 0;
            if constexpr (VEC == 1) {
                in.A.x = load1<T>(base + ((int64_t)r * qx + u) * ES);
            } else {
                in.o = (VEC - sc % VEC) % VEC;            // uniform: ocol and qx are multiples of VEC
                const int c0 = u - in.o;
                int c1 = c0 + VEC;
                c1 = c1 == qx ? 0 : c1;
                in.A = load16(base + ((int64_t)r * qx + c0) * ES);
                in.B = load16(base + ((int64_t)r * qx + c1) * ES);
            }
        }
    };
    auto consume = [&](const In &in) {
        u32x4 v;
        if constexpr (SHIFT && VEC > 1) v = pick<U16>(in.A, in.B, in.o);
        else v = in.A;
        if constexpr (INT) {
            if constexpr (VEC == 1) {
                if constexpr (MEAN) isum[0] += v.x;
                if constexpr (MAX) imax[0] = max(imax[0], v.x);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    // whole words accumulate (the low halves with the high ones carried in) and the high halves
                    // apart; the low sums are recovered at the end
                    if constexpr (MEAN) { isum[2 * k] += v[k]; isum[2 * k + 1] += v[k] >> 16; }
                    if constexpr (MAX) imax[k] = pk_max_u16(imax[k], v[k]);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                const double x = elem<T, VEC>(v, k) * in.w;
                if constexpr (MEAN) two_sum(s[k], c[k], x);
                if constexpr (MAX) m[k] = max_step<FIRST, NAN_POSSIBLE>(m[k], x);
            }
        }
    };

    int64_t j = j0;
    for (; j + U <= j1; j += U) {
        In in[U];
#pragma unroll
        for (int uu = 0; uu < U; ++uu) fetch(j + uu, in[uu]);
#pragma unroll
        for (int uu = 0; uu < U; ++uu) consume(in[uu]);
    }
    for (; j < j1; ++j) {
        In in;
        fetch(j, in);
        consume(in);
    }
    if constexpr (INT && VEC > 1) {
        if constexpr (MEAN) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t whole = isum[2 * k], hi = isum[2 * k + 1];
                isum[2 * k] = whole - (hi << 16);         // exact: the low sum stays below 2^32
            }
        }
        if constexpr (MAX) {
#pragma unroll
            for (int k = 3; k >= 0; --k) {
                const uint32_t pk = imax[k];
                imax[2 * k] = pk & 0xffffu;
                imax[2 * k + 1] = pk >> 16;
            }
        }
    }

    // the four waves in order: waves 1..3 hand their state to wave 0 through LDS
    constexpr int PER = INT ? ((MEAN ? VEC : 0) + (MAX ? VEC : 0)) : ((MEAN ? 2 * VEC : 0) + (MAX ? VEC : 0));
    __shared__ uint64_t lds[(kPatWaves - 1) * PER * 64];
    if (wave > 0) {
        uint64_t *dst = lds + (size_t)(wave - 1) * PER * 64 + lane;
        int e = 0;
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            if constexpr (INT) {
                if constexpr (MEAN) dst[64 * e++] = isum[k];
                if constexpr (MAX) dst[64 * e++] = imax[k];
            } else {
                if constexpr (MEAN) { dst[64 * e++] = __builtin_bit_cast(uint64_t, s[k]); dst[64 * e++] = __builtin_bit_cast(uint64_t, c[k]); }
                if constexpr (MAX) dst[64 * e++] = __builtin_bit_cast(uint64_t, m[k]);
            }
        }
    }
    __syncthreads();
    if (wave != 0) return;
    for (int w = 1; w < kPatWaves; ++w) {
        const uint64_t *src = lds + (size_t)(w - 1) * PER * 64 + lane;
        int e = 0;
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            if constexpr (INT) {
                if constexpr (MEAN) isum[k] += (uint32_t)src[64 * e++];
                if constexpr (MAX) imax[k] = max(imax[k], (uint32_t)src[64 * e++]);
            } else {
                if constexpr (MEAN) {
                    const double ps = __builtin_bit_cast(double, src[64 * e++]);
                    const double pc = __builtin_bit_cast(double, src[64 * e++]);
                    two_sum(s[k], c[k], ps);
                    c[k] += pc;
                }
                if constexpr (MAX) m[k] = max_step<FIRST, NAN_POSSIBLE>(m[k], __builtin_bit_cast(double, src[64 * e++]));
            }
        }
    }
    if (!live) return;
    const int64_t at = (int64_t)group * ndet + pix;
#pragma unroll
    for (int k = 0; k < VEC; ++k) {
        if constexpr (INT) {
            if constexpr (MEAN) ((uint32_t *)a.sum)[at + k] = isum[k];
            if constexpr (MAX) ((uint32_t *)a.max)[at + k] = imax[k];
        } else {
            if constexpr (MEAN) ((double2 *)a.sum)[at + k] = make_double2(s[k], c[k]);
            if constexpr (MAX) ((double *)a.max)[at + k] = m[k];
        }
    }
}

template <bool INT, int STAT, bool FIRST>
__global__ void pattern_accumulate(const void *part_sum, const void *part_max, int ngroups, int64_t ndet, void *run_sum,
                                   void *run_max, bool first, bool finish, double count, double *out) {
    constexpr bool MEAN = STAT == kPatternMean, MAX = STAT == kPatternMax;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndet) return;
    if constexpr (INT) {
        if constexpr (MEAN) {
            uint64_t sum = first ? 0u : ((const uint64_t *)run_sum)[i];
            for (int g = 0; g < ngroups; ++g) sum += ((const uint32_t *)part_sum)[(int64_t)g * ndet + i];
            if (finish) out[i] = (double)sum / count;
            else ((uint64_t *)run_sum)[i] = sum;
        }
        if constexpr (MAX) {
            uint32_t mx = first ? 0u : ((const uint32_t *)run_max)[i];
            for (int g = 0; g < ngroups; ++g) mx = max(mx, ((const uint32_t *)part_max)[(int64_t)g * ndet + i]);
            if (finish) out[i] = (double)mx;
            else ((uint32_t *)run_max)[i] = mx;
        }
    } else {
        if constexpr (MEAN) {
            double2 sc = first ? make_double2(0.0, 0.0) : ((const double2 *)run_sum)[i];
            for (int g = 0; g < ngroups; ++g) {
                const double2 p = ((const double2 *)part_sum)[(int64_t)g * ndet + i];
                two_sum(sc.x, sc.y, p.x);
                sc.y += p.y;
            }
            if (finish) {
                // CompensatedSum.value (CompensatedSum.swift:157-161): a non-finite sum as it is, -0 kept
                double v;
                if (!__builtin_isfinite(sc.x)) v = sc.x;
                else if (sc.x == 0.0 && __builtin_signbit(sc.x) && sc.y == 0.0) v = sc.x;
                else v = sc.x + sc.y;
                out[i] = v / count;
            } else {
                ((double2 *)run_sum)[i] = sc;
            }
        }
        if constexpr (MAX) {
            double mx = first ? -__builtin_inf() : ((const double *)run_max)[i];
            for (int g = 0; g < ngroups; ++g) mx = max_step<FIRST, true>(mx, ((const double *)part_max)[(int64_t)g * ndet + i]);
            if (finish) out[i] = __builtin_isnan(mx) ? __builtin_nan("") : mx;
            else ((double *)run_max)[i] = mx;
        }
    }
}

__device__ __forceinline__ int64_t entry_position(const FourierBatch &b, int64_t entry) {
    return b.list ? (int64_t)b.list[entry] : entry;
}

template <bool U16>
__global__ void fourier_expand(FourierBatch b) {
    const int64_t n = (int64_t)b.batch * b.ndet, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t k = i / b.ndet, u = i % b.ndet;
        const int64_t p = entry_position(b, b.first + k);
        double x;
        if constexpr (U16) x = (double)((const uint16_t *)b.cube)[p * b.ndet + u];
        else x = (double)((const float *)b.cube)[p * b.ndet + u];
        b.data[i] = make_double2(x, 0.0);
    }
}

// The reference's phase factor, operation for operation (SelectedDiffraction.swift:278-281).
__global__ void fourier_phase(FourierBatch b) {
    const int64_t n = (int64_t)b.batch * b.ndet, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t k = i / b.ndet, u = i % b.ndet;
        const int r = (int)(u / b.qx), c = (int)(u % b.qx);
        const int64_t p = entry_position(b, b.first + k);
        const double angle = -2 * M_PI * (b.sc[p] * b.fc[c] + b.sr[p] * b.fr[r]);
        const double co = cos(angle), si = sin(angle);
        const double2 v = b.data[i];
        b.data[i] = make_double2(v.x * co - v.y * si, v.x * si + v.y * co);
    }
}

// One group partial over the batch, entries in order: the real part (times the weight), TwoSum and the maximum.
template <int STAT, bool FIRST>
__global__ void fourier_reduce(FourierBatch b) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= b.ndet) return;
    double s = 0.0, c = 0.0, m = -__builtin_inf();
    for (int k = 0; k < b.batch; ++k) {
        double v = b.data[(int64_t)k * b.ndet + u].x;
        if (b.w) v = v * b.w[entry_position(b, b.first + k)];
        if constexpr (STAT == kPatternMean) two_sum(s, c, v);
        else m = max_step<FIRST, true>(m, v);
    }
    if constexpr (STAT == kPatternMean) b.part_sum[u] = make_double2(s, c);
    else b.part_max[u] = m;
}

// ---- PatternShift.fourier in one kernel -------------------------------------------------------------------------
//
// LDS (doubles): the pattern region [qy / 2 pairs][2 (qx / 2 + 1) complex], then the ramp factors A [qy] and
// B [qx / 2 + 1], the per-stage twiddle tables, the doubly-Nyquist factor and, for the mean, the sums of the
// Nyquist column's ramped bins [qy]. The mean transforms each pattern forward only and adds its ramped spectrum
// into per-lane sums; the block inverts the summed spectrum once. The maximum transforms each pattern forward and
// back. Pair q (rows 2q, 2q + 1) owns doubles
// q * 2 (qx + 2) .. (q + 1) * 2 (qx + 2) - 1 in two layouts:
//   - split, for the row transforms: the real parts of the packed row z = row 2q + i row 2q+1 at offset c, the
//     imaginary parts at offset qx + 2 + c;
//   - interleaved, for the column transforms: the half spectrum of row r, bin k, as the complex number
//     (r (qx / 2 + 1) + k) of the region.
// Wave w owns the pairs w, w + 8, ... (their row transforms, the pack and unpack between the layouts and the
// pixels it adds up) and the columns w, w + 8, ...; two block barriers per pattern separate the row and column
// phases. Within a wave, LDS operations complete in issue order, so a wave-level fence orders its stages.
constexpr int kFfWaves = 8;
constexpr int kFfBlock = 64 * kFfWaves;
constexpr int kFfPix = 32;        // pixels per lane: qy * qx <= 64 * kFfWaves * kFfPix

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double2 cmul(double2 x, double2 y) {
    return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
}

__device__ __forceinline__ double2 cmul_conj(double2 x, double2 y) {   // x * conj(y)
    return make_double2(x.x * y.x + x.y * y.y, x.y * y.x - x.x * y.y);
}

__device__ __forceinline__ int bit_reverse(int x, int bits) {
    return (int)(__builtin_bitreverse32((unsigned)x) >> (32 - bits));
}

__device__ __forceinline__ double2 cadd(double2 x, double2 y) { return make_double2(x.x + y.x, x.y + y.y); }
__device__ __forceinline__ double2 csub(double2 x, double2 y) { return make_double2(x.x - y.x, x.y - y.y); }

// One radix-4 stage of a decimation-in-frequency transform (exp(-2 pi i ...)): the elements j, j + h, j + 2h, j + 3h
// of a block of 4h, w1..w3 = W^j, W^2j, W^3j with W = exp(-2 pi i / 4h). The results land where the two radix-2
// stages of spans 2h and h would put them, so the output order stays bit-reversed.
__device__ __forceinline__ void dif4(double2 &x0, double2 &x1, double2 &x2, double2 &x3, double2 w1, double2 w2,
                                     double2 w3, bool twiddle) {
    const double2 a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3), d = csub(x1, x3);
    const double2 a3 = make_double2(d.y, -d.x);       // (x1 - x3) * -i
    double2 y0 = cadd(a0, a2), y2 = csub(a0, a2), y1 = cadd(a1, a3), y3 = csub(a1, a3);
    if (twiddle) {
        y1 = cmul(y1, w1);
        y2 = cmul(y2, w2);
        y3 = cmul(y3, w3);
    }
    x0 = y0;   // position j
    x1 = y2;   // j + h
    x2 = y1;   // j + 2h
    x3 = y3;   // j + 3h
}

// One radix-4 stage of a decimation-in-time inverse (exp(+2 pi i ...)): the radix-2 stages of spans h and 2h, with
// c1 = W^j and c2 = W^2j (conjugated here).
__device__ __forceinline__ void dit4(double2 &x0, double2 &x1, double2 &x2, double2 &x3, double2 c1, double2 c2,
                                     bool twiddle) {
    double2 t1 = x1, t3 = x3;
    if (twiddle) {
        t1 = cmul_conj(x1, c2);
        t3 = cmul_conj(x3, c2);
    }
    const double2 p0 = cadd(x0, t1), p1 = csub(x0, t1), p2 = cadd(x2, t3), p3 = csub(x2, t3);
    double2 q2 = p2, q3 = p3;
    if (twiddle) {
        q2 = cmul_conj(p2, c1);
        q3 = cmul_conj(p3, c1);
    }
    const double2 q3i = make_double2(-q3.y, q3.x);    // * i
    x0 = cadd(p0, q2);    // position j
    x2 = csub(p0, q2);    // j + 2h
    x1 = cadd(p1, q3i);   // j + h
    x3 = csub(p1, q3i);   // j + 3h
}

// Bank-spreading permutations of the fused kernel's LDS (each an XOR of higher bits into lower ones, so a
// bijection within aligned blocks): logical row r of the half spectrum is stored as row swz_row(r), and element i of
// a packed row in the split layout at swz_el(i). With them the lanes of every radix-4 stage of a 128x128 pattern
// reach distinct banks (rows 4b + m, 16 (b >> 2) + (b & 3) + 4m, ... of one column or packed row).
__device__ __forceinline__ int swz_row(int r) {
    const int g = (r >> 4) & 3;
    return r ^ (g | (g << 2));
}
__device__ __forceinline__ int swz_el(int i) {
    const int g = (i >> 5) & 3;
    return i ^ (g | (g << 2)) ^ (((i >> 6) & 1) << 4);
}

// Entries of the per-stage twiddle tables for transforms of length up to l: 3 h for each radix-4 stage of quarter
// h = 1, 2, .., l / 4 and h for each radix-2 stage of span h = 1, .., l / 2.
__host__ __device__ constexpr int ff_twiddle_entries(int l) { return 3 * (l / 2 - 1) + (l - 1); }

int fourier_fused_lds(int qy, int qx) {
    const int h = qx / 2 + 1, l = qy > qx ? qy : qx;
    return 8 * (qy * (qx + 2) + 2 * qy + 2 * h + 2 * ff_twiddle_entries(l) + 2 + 2 * qy);
}

template <bool U16, int STAT, bool FIRST, int LM>
__global__ __launch_bounds__(kFfBlock) void fourier_fused(FourierFused a) {
    constexpr bool MEAN = STAT == kPatternMean;
    constexpr int VE = U16 ? 8 : 4;                      // elements per 16-byte load
    constexpr int CH = U16 ? 4 : 8;                      // loads per lane and pattern
    constexpr int ES = U16 ? 2 : 4;
    extern __shared__ double ff_lds[];
    constexpr int M = 1 << LM, H = M / 2 + 1;
    constexpr int R2 = M + 2;                            // split layout: real plane to imaginary plane
    const int N = a.qy;
    const int LN = __builtin_ctz((unsigned)N);
    const int L = N > M ? N : M, LL = __builtin_ctz((unsigned)L);
    double *S = ff_lds;
    double2 *C = (double2 *)ff_lds;
    double2 *A = (double2 *)(ff_lds + N * R2);
    double2 *B = A + N;
    // Per-stage twiddle tables (the entries of a.tw, exp(-2 pi i m / L), that each stage uses, in the order its lanes
    // read them, so consecutive lanes read consecutive entries instead of a.tw at a stride of L / (4 h)):
    //   radix-4 stage of quarter h = 2^lq: TWS[3 (h - 1) + (m - 1) h + j] = a.tw[(m j) << (LL - lq - 2)], m = 1, 2, 3;
    //   radix-2 stage of span h = 2^lh:    TWS[R4T + h - 1 + j] = a.tw[j << (LL - 1 - lh)], R4T = 3 (L / 2 - 1).
    double2 *TWS = B + H;
    const int R4T = 3 * (L / 2 - 1);
    double2 *NYQ = TWS + ff_twiddle_entries(L);
    double2 *NQ = NYQ + 1;                               // the mean: sums of the Nyquist column's ramped bins
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int NP = N / 16;                               // pairs per wave
    const int NC = wave < H ? (H - wave + kFfWaves - 1) / kFfWaves : 0;   // columns of this wave
    const int nchunks = NP * 2 * M / VE;
    const int J = NP * 2 * M / 64;                       // pixels per lane (qx >= 32: whole 64-pixel slices)
    // packed pair q (logical rows 2q, 2q + 1) lives in the region of stored rows swz_row(2q) and its neighbour;
    // flip: logical row 2q is the region's second stored row
    auto region = [&](int q) { return (swz_row(2 * q) >> 1) * 2 * R2; };
    auto flip = [&](int q) { return swz_row(2 * q) & 1; };
    const int64_t e_begin = a.first + (int64_t)blockIdx.x * a.ppg;
    const int64_t e_end = min(a.first + a.count, e_begin + a.ppg);
    for (int e = tid; e < ff_twiddle_entries(L); e += kFfBlock) {
        int idx;
        if (e < R4T) {
            int lq = 0;
            while (e >= 3 * ((2 << lq) - 1)) ++lq;
            const int h = 1 << lq, off = e - 3 * (h - 1), m = off / h + 1, j = off - (m - 1) * h;
            idx = (m * j) << (LL - lq - 2);
        } else {
            const int e2 = e - R4T;
            int lh = 0;
            while (e2 >= (2 << lh) - 1) ++lh;
            idx = (e2 - ((1 << lh) - 1)) << (LL - 1 - lh);
        }
        TWS[e] = a.tw[idx];
    }
    auto tw2 = [&](int lh, int j) { return TWS[R4T + (1 << lh) - 1 + j]; };
    auto tw4 = [&](int lq, int m, int j) { return TWS[3 * ((1 << lq) - 1) + (m - 1) * (1 << lq) + j]; };

    // per slot: the running maximum, or, for the mean, the block's sum once its summed spectrum is inverted
    double acc_s[kFfPix];
#pragma unroll
    for (int j = 0; j < kFfPix; ++j) acc_s[j] = MEAN ? 0.0 : -__builtin_inf();
    u32x4 pre[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) pre[j] = u32x4{0u, 0u, 0u, 0u};
    auto position = [&](int64_t e) { return a.list ? (int64_t)a.list[e] : e; };
    // The next pattern's loads, all issued back to back: a uniform base and 32-bit offsets (a pattern is at most
    // 64 KiB), a lane without a chunk re-reading the first one, so no branch and no 64-bit addresses to keep.
    auto prefetch = [&](int64_t e) {
        const char *base = (const char *)a.cube + position(e) * a.ndet * ES;
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int u = lane + 64 * j;
            const int per = 2 * M / VE, t = u / per, cc = u - t * per;
            const int q = wave + kFfWaves * t;
            const uint32_t off = u < nchunks ? (uint32_t)((2 * q * M + cc * VE) * ES) : 0u;
            pre[j] = load16(base + off);
        }
    };
    // Stages over this wave's packed pairs (split layout). Radix-2 span h = 2^lh, radix-4 quarter h = 2^lq.
    auto row_r2_dif = [&](int lh) {
        const int h = 1 << lh;
        for (int u = lane; u < NP * (M / 2); u += 64) {
            const int t = u >> (LM - 1), b = u & (M / 2 - 1), j = b & (h - 1);
            const int i0 = swz_el(((b >> lh) << (lh + 1)) + j), i1 = swz_el(((b >> lh) << (lh + 1)) + j + h);
            double *re = S + region(wave + kFfWaves * t), *im = re + R2;
            const double2 x = make_double2(re[i0], im[i0]), y = make_double2(re[i1], im[i1]);
            double2 d = csub(x, y);
            if (h > 1) d = cmul(d, tw2(lh, j));
            re[i0] = x.x + y.x;
            im[i0] = x.y + y.y;
            re[i1] = d.x;
            im[i1] = d.y;
        }
        wave_sync();
    };
    auto row_r2_dit = [&](int lh) {
        const int h = 1 << lh;
        for (int u = lane; u < NP * (M / 2); u += 64) {
            const int t = u >> (LM - 1), b = u & (M / 2 - 1), j = b & (h - 1);
            const int i0 = swz_el(((b >> lh) << (lh + 1)) + j), i1 = swz_el(((b >> lh) << (lh + 1)) + j + h);
            double *re = S + region(wave + kFfWaves * t), *im = re + R2;
            const double2 x = make_double2(re[i0], im[i0]);
            double2 y = make_double2(re[i1], im[i1]);
            if (h > 1) y = cmul_conj(y, tw2(lh, j));
            re[i0] = x.x + y.x;
            im[i0] = x.y + y.y;
            re[i1] = x.x - y.x;
            im[i1] = x.y - y.y;
        }
        wave_sync();
    };
    auto row_r4 = [&](int lq, bool inverse) {
        const int h = 1 << lq;
        for (int u = lane; u < NP * (M / 4); u += 64) {
            const int t = u >> (LM - 2), b = u & (M / 4 - 1), j = b & (h - 1);
            const int i0 = ((b >> lq) << (lq + 2)) + j;
            const int e0 = swz_el(i0), e1 = swz_el(i0 + h), e2 = swz_el(i0 + 2 * h), e3 = swz_el(i0 + 3 * h);
            double *re = S + region(wave + kFfWaves * t), *im = re + R2;
            double2 x0 = make_double2(re[e0], im[e0]), x1 = make_double2(re[e1], im[e1]);
            double2 x2 = make_double2(re[e2], im[e2]), x3 = make_double2(re[e3], im[e3]);
            if (inverse) dit4(x0, x1, x2, x3, tw4(lq, 1, j), tw4(lq, 2, j), h > 1);
            else dif4(x0, x1, x2, x3, tw4(lq, 1, j), tw4(lq, 2, j), tw4(lq, 3, j), h > 1);
            re[e0] = x0.x; im[e0] = x0.y;
            re[e1] = x1.x; im[e1] = x1.y;
            re[e2] = x2.x; im[e2] = x2.y;
            re[e3] = x3.x; im[e3] = x3.y;
        }
        wave_sync();
    };
    auto row_r4_dif = [&](int lq) { row_r4(lq, false); };
    auto row_r4_dit = [&](int lq) { row_r4(lq, true); };
    // Stages over this wave's columns (interleaved layout: bin k of row r is C[r H + k]).
    auto col_r2_dif = [&](int lh) {            // never the last forward stage (that one is radix-4)
        const int h = 1 << lh;
        for (int u = lane; u < NC * (N / 2); u += 64) {
            const int i = u >> (LN - 1), b = u & (N / 2 - 1), j = b & (h - 1);
            const int k = wave + kFfWaves * i;
            const int r0 = swz_row(((b >> lh) << (lh + 1)) + j), r1 = swz_row(((b >> lh) << (lh + 1)) + j + h);
            const double2 x = C[r0 * H + k], y = C[r1 * H + k];
            double2 d = csub(x, y);
            if (h > 1) d = cmul(d, tw2(lh, j));
            C[r0 * H + k] = cadd(x, y);
            C[r1 * H + k] = d;
        }
        wave_sync();
    };
    auto col_r2_dit = [&](int lh) {
        const int h = 1 << lh;
        for (int u = lane; u < NC * (N / 2); u += 64) {
            const int i = u >> (LN - 1), b = u & (N / 2 - 1), j = b & (h - 1);
            const int k = wave + kFfWaves * i;
            const int r0 = swz_row(((b >> lh) << (lh + 1)) + j), r1 = swz_row(((b >> lh) << (lh + 1)) + j + h);
            const double2 x = C[r0 * H + k];
            double2 y = C[r1 * H + k];
            if (h > 1) y = cmul_conj(y, tw2(lh, j));
            C[r0 * H + k] = cadd(x, y);
            C[r1 * H + k] = csub(x, y);
        }
        wave_sync();
    };
    auto ramp = [&](int f, int k) { return (k == M / 2 && f == N / 2) ? NYQ[0] : cmul(A[f], B[k]); };
    auto col_r4 = [&](int lq, bool inverse) {
        const int h = 1 << lq;
        for (int u = lane; u < NC * (N / 4); u += 64) {
            const int i = u >> (LN - 2), b = u & (N / 4 - 1), j = b & (h - 1);
            const int k = wave + kFfWaves * i;
            const int i0 = ((b >> lq) << (lq + 2)) + j;
            double2 *c0 = C + swz_row(i0) * H + k, *c1 = C + swz_row(i0 + h) * H + k;
            double2 *c2 = C + swz_row(i0 + 2 * h) * H + k, *c3 = C + swz_row(i0 + 3 * h) * H + k;
            double2 x0 = *c0, x1 = *c1, x2 = *c2, x3 = *c3;
            if (inverse) {
                dit4(x0, x1, x2, x3, tw4(lq, 1, j), tw4(lq, 2, j), h > 1);
            } else {
                dif4(x0, x1, x2, x3, tw4(lq, 1, j), tw4(lq, 2, j), tw4(lq, 3, j), h > 1);
                if (lq == 0) {                         // the last forward stage: frequencies bit-reversed
                    x0 = cmul(x0, ramp(bit_reverse(i0, LN), k));
                    x1 = cmul(x1, ramp(bit_reverse(i0 + 1, LN), k));
                    x2 = cmul(x2, ramp(bit_reverse(i0 + 2, LN), k));
                    x3 = cmul(x3, ramp(bit_reverse(i0 + 3, LN), k));
                }
            }
            *c0 = x0;
            *c1 = x1;
            *c2 = x2;
            *c3 = x3;
        }
        wave_sync();
    };
    auto col_r4_dif = [&](int lq) { col_r4(lq, false); };
    auto col_r4_dit = [&](int lq) { col_r4(lq, true); };
    // The mean: the sum of the shifted patterns is the inverse transform of the sum of their ramped spectra, so each
    // pattern only goes forward. The last forward column stage (radix-4 of quarter 1) adds its ramped outputs,
    // times the position's weight, into sums instead of storing them, and the block inverts its summed spectrum
    // once. Columns below M / 2 (M / 16 of them per wave) are summed in registers, spec[it][m] holding bin
    // (bit-reversed row 4 b + m, column k) of task u = lane + 64 it, the same bins for every pattern (at most 4
    // tasks: 64 registers); the Nyquist column, wave 0's extra one, in LDS (NQ, each entry always the same lane's).
    constexpr int KB = MEAN ? (M / 16) * (64 * kFfWaves * kFfPix / M / 4) / 64 : 1;   // tasks per lane
    static_assert(!MEAN || (M >= 16 && KB >= 1), "whole columns per wave below the Nyquist column");
    const int NCL = M / 16;                              // this wave's columns below M / 2
    double2 spec[KB][4];
#pragma unroll
    for (int it = 0; it < KB; ++it)
#pragma unroll
        for (int m = 0; m < 4; ++m) spec[it][m] = make_double2(0.0, 0.0);
    auto col_task = [&](int u, int &i0, int &k) {
        const int i = u >> (LN - 2), b = u & (N / 4 - 1);
        k = wave + kFfWaves * i;
        i0 = 4 * b;
    };
    // the ramped outputs of the last stage's butterfly of rows i0 .. i0 + 3 of column k
    auto last_bins = [&](int i0, int k, double wgt, double2 (&x)[4]) {
#pragma unroll
        for (int m = 0; m < 4; ++m) x[m] = C[swz_row(i0 + m) * H + k];
        dif4(x[0], x[1], x[2], x[3], tw4(0, 1, 0), tw4(0, 2, 0), tw4(0, 3, 0), false);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int f = bit_reverse(i0 + m, LN);
            const double2 r = (k == M / 2 && f == N / 2) ? NYQ[0] : cmul(A[f], B[k]);
            x[m] = cmul(x[m], r);
            if (a.w) x[m] = make_double2(x[m].x * wgt, x[m].y * wgt);
        }
    };
    auto col_sum_last = [&](double wgt) {
#pragma unroll
        for (int it = 0; it < KB; ++it) {
            const int u = lane + 64 * it;
            if (u < NCL * (N / 4)) {
                int i0, k;
                col_task(u, i0, k);
                double2 x[4];
                last_bins(i0, k, wgt, x);
#pragma unroll
                for (int m = 0; m < 4; ++m) spec[it][m] = cadd(spec[it][m], x[m]);
            }
            __builtin_amdgcn_sched_barrier(0);           // one task's loads in flight at a time (registers)
        }
        if (wave == 0) {
            for (int b = lane; b < N / 4; b += 64) {
                double2 x[4];
                last_bins(4 * b, M / 2, wgt, x);
#pragma unroll
                for (int m = 0; m < 4; ++m) NQ[4 * b + m] = cadd(NQ[4 * b + m], x[m]);
            }
        }
    };
    // The block's summed spectrum into the bins the last forward stage stores.
    auto col_store_sums = [&]() {
#pragma unroll
        for (int it = 0; it < KB; ++it) {
            const int u = lane + 64 * it;
            if (u < NCL * (N / 4)) {
                int i0, k;
                col_task(u, i0, k);
#pragma unroll
                for (int m = 0; m < 4; ++m) C[swz_row(i0 + m) * H + k] = spec[it][m];
            }
        }
        if (wave == 0)
            for (int i = lane; i < N; i += 64) C[swz_row(i) * H + M / 2] = NQ[i];
        wave_sync();
    };
    // The inverse: the columns (decimation in time: bit-reversed in, natural out); then pack W = X0 + i X1 over the
    // full row (W[-k] = conj X0[k] + i conj X1[k]), bit-reversed, and the rows (decimation in time), after which
    // rows 2q and 2q + 1 are real.
    auto to_pixels = [&]() {
        {
            int lq = 0;
            for (; lq + 2 <= LN; lq += 2) col_r4_dit(lq);
            if (lq < LN) col_r2_dit(lq);
        }
        __syncthreads();
        // (the bins below qx / 2 pair by pair, then the Nyquist bin of every pair at once, a pair per lane: its
        // interleaved slot lies in the split rows' padding and its split element is written by no other bin)
        constexpr int PI = (M / 2 + 63) / 64;
        for (int t = 0; t < NP; ++t) {
            const int q = wave + kFfWaves * t;
            double *re = S + region(q), *im = re + R2;
            double2 z0[PI], z1[PI];
#pragma unroll
            for (int i = 0; i < PI; ++i) {
                const int k = lane + 64 * i;
                if (k < M / 2) {
                    z0[i] = ((const double2 *)(re + flip(q) * R2))[k];
                    z1[i] = ((const double2 *)(re + (flip(q) ^ 1) * R2))[k];
                }
            }
            wave_sync();
#pragma unroll
            for (int i = 0; i < PI; ++i) {
                const int k = lane + 64 * i;
                if (k < M / 2) {
                    const int pk = swz_el(bit_reverse(k, LM));
                    re[pk] = z0[i].x - z1[i].y;
                    im[pk] = z0[i].y + z1[i].x;
                    if (k > 0) {
                        const int pm = swz_el(bit_reverse(M - k, LM));
                        re[pm] = z0[i].x + z1[i].y;
                        im[pm] = z1[i].x - z0[i].y;
                    }
                }
            }
            wave_sync();
        }
        if (lane < NP) {
            const int qn = wave + kFfWaves * lane;
            double *re = S + region(qn), *im = re + R2;
            const double2 z0 = ((const double2 *)(re + flip(qn) * R2))[M / 2];
            const double2 z1 = ((const double2 *)(re + (flip(qn) ^ 1) * R2))[M / 2];
            const int pn = swz_el(bit_reverse(M / 2, LM));
            re[pn] = z0.x - z1.y;
            im[pn] = z0.y + z1.x;
        }
        wave_sync();
        // inverse row transforms (decimation in time: bit-reversed in, natural out): rows 2q and 2q+1 real
        {
            int lq = 0;
            for (; lq + 2 <= LM; lq += 2) row_r4_dit(lq);
            if (lq < LM) row_r2_dit(lq);
        }
    };

    if constexpr (MEAN)
        for (int i = tid; i < N; i += kFfBlock) NQ[i] = make_double2(0.0, 0.0);
    __syncthreads();                                     // the twiddles (and the zeroed Nyquist sums)
    if (e_begin < e_end) prefetch(e_begin);
    for (int64_t e = e_begin; e < e_end; ++e) {
        // the raw pattern into the split layout of this wave's pairs
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int u = lane + 64 * j;
            if (u < nchunks) {
                const int per = 2 * M / VE, t = u / per, cc = u - t * per;
                const int q = wave + kFfWaves * t, off = cc * VE;
                const int odd = off >= M ? 1 : 0;
                const int c0 = off - odd * M, x = swz_el(c0) ^ c0;   // the chunk's XOR (bits 5, 6 of c0)
                double *dst = S + region(q) + odd * R2;
#pragma unroll
                for (int i = 0; i < VE; i += 2) {
                    double x0, x1;
                    if constexpr (U16) {
                        x0 = (double)(pre[j][i >> 1] & 0xffffu);
                        x1 = (double)(pre[j][i >> 1] >> 16);
                    } else {
                        // through scalars: __builtin_bit_cast of a vector element subscript reads element 0
                        const uint32_t b0 = pre[j][i], b1 = pre[j][i + 1];
                        x0 = (double)__builtin_bit_cast(float, b0);
                        x1 = (double)__builtin_bit_cast(float, b1);
                    }
                    // elements c0 + i, c0 + i + 1 land on (c0 + i) ^ x and its neighbour: swapped when x is odd
                    *(double2 *)(dst + ((c0 + i) ^ (x & ~1))) = (x & 1) ? make_double2(x1, x0) : make_double2(x0, x1);
                }
            }
        }
        const int64_t pe = position(e);
        if (e + 1 < e_end) prefetch(e + 1);
        // the ramp: A[f] B[k] is exp(-2 pi i (sr fr + sc fc)) / (qy qx) with NumPy's fftfreq, the angles taken
        // in half turns (sincospi); at the Nyquist row and column the ramp averaged with its conjugate mirror
        // (the Hermitian part, whose inverse is the real part of the reference's inverse transform)
        {
            const double sr = a.sr[pe], sc = a.sc[pe];
            const double scale = 1.0 / ((double)N * (double)M);
            for (int i = tid; i < N + H; i += kFfBlock) {
                if (i < N) {
                    if (i == N / 2) {
                        A[i] = make_double2(cospi(sr) * scale, 0.0);
                    } else {
                        const double fr = (double)(i < N / 2 ? i : i - N) / (double)N;
                        double sn, cs;
                        sincospi(-2.0 * (sr * fr), &sn, &cs);
                        A[i] = make_double2(cs * scale, sn * scale);
                    }
                } else {
                    const int k = i - N;
                    if (k == M / 2) {
                        B[k] = make_double2(cospi(sc), 0.0);
                    } else {
                        double sn, cs;
                        sincospi(-2.0 * (sc * ((double)k / (double)M)), &sn, &cs);
                        B[k] = make_double2(cs, sn);
                    }
                }
            }
            if (tid == 0) NYQ[0] = make_double2(cospi(sr + sc) * scale, 0.0);
        }
        wave_sync();
        // forward row transforms of the packed pairs (decimation in frequency: natural in, bit-reversed out): a
        // radix-2 stage first when log2 qx is odd, then radix-4
        {
            int top = LM - 1;
            if (LM & 1) {
                row_r2_dif(top);
                top -= 1;
            }
            for (; top >= 1; top -= 2) row_r4_dif(top - 1);
        }
        // unpack: Z = FFT(row 2q + i row 2q+1) gives both half spectra, X0[k] = (Z[k] + conj Z[-k]) / 2 and
        // X1[k] = -i (Z[k] - conj Z[-k]) / 2, written interleaved
        // (the Nyquist bin k = qx / 2, its own mirror, of every pair at once, a pair per lane: read before the pairs'
        // writes and written after them, into the interleaved slot that lies in the split rows' padding, which no
        // other write touches; the pairs' loop then takes the qx / 2 bins below it, whole waves at 128 columns)
        constexpr int UI = (M / 2 + 63) / 64;
        const int qn = wave + kFfWaves * lane;
        double2 zn = make_double2(0.0, 0.0);
        if (lane < NP) {
            const int pn = swz_el(bit_reverse(M / 2, LM));
            zn = make_double2(S[region(qn) + pn], S[region(qn) + R2 + pn]);
        }
        wave_sync();
        for (int t = 0; t < NP; ++t) {
            const int q = wave + kFfWaves * t;
            double *re = S + region(q), *im = re + R2;
            double2 z0[UI], z1[UI];
#pragma unroll
            for (int i = 0; i < UI; ++i) {
                const int k = lane + 64 * i;
                if (k < M / 2) {
                    const int pk = swz_el(bit_reverse(k, LM)), pm = swz_el(bit_reverse((M - k) & (M - 1), LM));
                    z0[i] = make_double2(re[pk], im[pk]);
                    z1[i] = make_double2(re[pm], im[pm]);
                }
            }
            wave_sync();
#pragma unroll
            for (int i = 0; i < UI; ++i) {
                const int k = lane + 64 * i;
                if (k < M / 2) {
                    // logical rows 2q, 2q + 1 are the region's stored rows swz_row(2q), swz_row(2q) ^ 1
                    ((double2 *)(re + flip(q) * R2))[k] = make_double2(0.5 * (z0[i].x + z1[i].x), 0.5 * (z0[i].y - z1[i].y));
                    ((double2 *)(re + (flip(q) ^ 1) * R2))[k] =
                        make_double2(0.5 * (z0[i].y + z1[i].y), -0.5 * (z0[i].x - z1[i].x));
                }
            }
            wave_sync();
        }
        if (lane < NP) {
            double *re = S + region(qn);
            ((double2 *)(re + flip(qn) * R2))[M / 2] = make_double2(0.5 * (zn.x + zn.x), 0.5 * (zn.y - zn.y));
            ((double2 *)(re + (flip(qn) ^ 1) * R2))[M / 2] = make_double2(0.5 * (zn.y + zn.y), -0.5 * (zn.x - zn.x));
        }
        wave_sync();
        __syncthreads();
        // the columns: forward (decimation in frequency; the last stage, radix-4, applies the ramp), then, for the
        // maximum, inverse
        const double wgt = a.w ? a.w[pe] : 1.0;
        {
            int top = LN - 1;
            if (LN & 1) {
                col_r2_dif(top);
                top -= 1;
            }
            for (; top >= 3; top -= 2) col_r4_dif(top - 1);
            if constexpr (MEAN) col_sum_last(wgt);
            else col_r4_dif(0);
        }
        if constexpr (MEAN) {
            __syncthreads();                             // every wave's columns read before the next pattern's load
            continue;
        }
        to_pixels();
        // this lane's slots: stored element rem & (M - 1) of plane rem >= M (the same logical pixel for every
        // entry: swz_el is its own inverse); the shifted value (times the weight) and the maximum
#pragma unroll
        for (int j = 0; j < kFfPix; ++j) {
            const int u = lane + 64 * j;
            if (j < J) {
                const int t = u >> (LM + 1), rem = u & (2 * M - 1);
                double v = S[region(wave + kFfWaves * t) + (rem >= M ? R2 : 0) + (rem & (M - 1))];
                if (a.w) v = v * wgt;
                acc_s[j] = max_step<FIRST, true>(acc_s[j], v);
            }
            if (j % 8 == 7) __builtin_amdgcn_sched_barrier(0);   // at most 8 reads in flight
        }
        wave_sync();
    }
    if constexpr (MEAN) {
        // the block's sum of shifted patterns from its summed spectrum; the slots as for the maximum
        col_store_sums();
        to_pixels();
#pragma unroll
        for (int j = 0; j < kFfPix; ++j) {
            const int u = lane + 64 * j;
            if (j < J) {
                const int t = u >> (LM + 1), rem = u & (2 * M - 1);
                acc_s[j] = S[region(wave + kFfWaves * t) + (rem >= M ? R2 : 0) + (rem & (M - 1))];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kFfPix; ++j) {
        const int u = lane + 64 * j;
        if (j < J) {
            const int t = u >> (LM + 1), rem = u & (2 * M - 1);
            const int64_t pix = (int64_t)(wave + kFfWaves * t) * 2 * M + (rem & M) + swz_el(rem & (M - 1));
            const int64_t at = (int64_t)blockIdx.x * a.ndet + pix;
            if constexpr (MEAN) a.part_sum[at] = make_double2(acc_s[j], 0.0);   // (sum, compensation)
            else a.part_max[at] = acc_s[j];
        }
    }
}

template <bool U16, int STAT, bool FIRST, int LM>
hipError_t launch_ff(const FourierFused &a, int lds, hipStream_t s) {
    const void *fn = (const void *)fourier_fused<U16, STAT, FIRST, LM>;
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((fourier_fused<U16, STAT, FIRST, LM>), dim3((unsigned)a.ngroups), dim3(kFfBlock), lds, s, a);
    return hipGetLastError();
}

template <bool U16, int STAT, bool FIRST>
hipError_t launch_ff_m(const FourierFused &a, int lds, hipStream_t s) {
    switch (a.qx) {
    case 32: return launch_ff<U16, STAT, FIRST, 5>(a, lds, s);
    case 64: return launch_ff<U16, STAT, FIRST, 6>(a, lds, s);
    case 128: return launch_ff<U16, STAT, FIRST, 7>(a, lds, s);
    case 256: return launch_ff<U16, STAT, FIRST, 8>(a, lds, s);
    default: return hipErrorInvalidValue;
    }
}

template <bool U16>
hipError_t launch_ff_u(const FourierFused &a, int lds, hipStream_t s) {
    if (a.stat == kPatternMean) return launch_ff_m<U16, kPatternMean, true>(a, lds, s);
    if (a.first_wins) return launch_ff_m<U16, kPatternMax, true>(a, lds, s);
    return launch_ff_m<U16, kPatternMax, false>(a, lds, s);
}

template <typename T, int VEC, bool INT, int STAT>
hipError_t launch_t(const PatternLaunch &a, const PatternKind &k, dim3 grid, hipStream_t s) {
    if (k.shift) {
        if (k.first_wins) hipLaunchKernelGGL((pattern_kernel<T, VEC, INT, STAT, true, true>), grid, dim3(kPatBlock), 0, s, a);
        else hipLaunchKernelGGL((pattern_kernel<T, VEC, INT, STAT, true, false>), grid, dim3(kPatBlock), 0, s, a);
    } else {
        if (k.first_wins) hipLaunchKernelGGL((pattern_kernel<T, VEC, INT, STAT, false, true>), grid, dim3(kPatBlock), 0, s, a);
        else hipLaunchKernelGGL((pattern_kernel<T, VEC, INT, STAT, false, false>), grid, dim3(kPatBlock), 0, s, a);
    }
    return hipGetLastError();
}

template <typename T, int VEC>
hipError_t launch_v(const PatternLaunch &a, const PatternKind &k, dim3 grid, hipStream_t s) {
    if (k.integer) {
        if constexpr (sizeof(T) == 2)
            return k.stat == kPatternMean ? launch_t<T, VEC, true, kPatternMean>(a, k, grid, s)
                                          : launch_t<T, VEC, true, kPatternMax>(a, k, grid, s);
        else
            return hipErrorInvalidValue;
    }
    return k.stat == kPatternMean ? launch_t<T, VEC, false, kPatternMean>(a, k, grid, s)
                                  : launch_t<T, VEC, false, kPatternMax>(a, k, grid, s);
}

}  // namespace

int pattern_vec(bool u16, int64_t ndet, int qx, bool shift) {
    const int v = u16 ? 8 : 4;
    if ((ndet * (u16 ? 2 : 4)) % 16 != 0) return 1;
    if (shift && qx % v != 0) return 1;
    return v;
}

int pattern_block_pixels(int vec) { return 64 * vec; }

hipError_t launch_pattern(const PatternLaunch &a, const PatternKind &k, hipStream_t s) {
    if (a.count <= 0 || a.ngroups <= 0) return hipSuccess;
    const int bp = pattern_block_pixels(k.vec);
    const dim3 grid((unsigned)((a.ndet + bp - 1) / bp), (unsigned)a.ngroups);
    if (k.u16) return k.vec == 8 ? launch_v<uint16_t, 8>(a, k, grid, s) : launch_v<uint16_t, 1>(a, k, grid, s);
    return k.vec == 4 ? launch_v<float, 4>(a, k, grid, s) : launch_v<float, 1>(a, k, grid, s);
}

hipError_t launch_fourier_expand(const FourierBatch &b, hipStream_t s) {
    const int64_t n = (int64_t)b.batch * b.ndet;
    const dim3 grid((unsigned)std::min<int64_t>((n + 255) / 256, 16384));
    if (b.u16) hipLaunchKernelGGL(fourier_expand<true>, grid, dim3(256), 0, s, b);
    else hipLaunchKernelGGL(fourier_expand<false>, grid, dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_fourier_phase(const FourierBatch &b, hipStream_t s) {
    const int64_t n = (int64_t)b.batch * b.ndet;
    hipLaunchKernelGGL(fourier_phase, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 16384)), dim3(256), 0, s, b);
    return hipGetLastError();
}

bool fourier_fused_supported(int qy, int qx) {
    auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
    return pow2(qy) && pow2(qx) && qy >= 16 && qx >= 32 && qx <= 256 &&
           (int64_t)qy * qx <= (int64_t)64 * kFfWaves * kFfPix && fourier_fused_lds(qy, qx) <= 160 * 1024;
}

hipError_t launch_fourier_fused(const FourierFused &a, hipStream_t s) {
    if (!fourier_fused_supported(a.qy, a.qx) || a.ngroups <= 0) return hipErrorInvalidValue;
    if (a.count <= 0) return hipSuccess;
    const int lds = fourier_fused_lds(a.qy, a.qx);
    if (a.u16) return launch_ff_u<true>(a, lds, s);
    return launch_ff_u<false>(a, lds, s);
}

hipError_t launch_fourier_reduce(const FourierBatch &b, hipStream_t s) {
    const dim3 grid((unsigned)((b.ndet + 255) / 256));
    if (b.stat == kPatternMean) hipLaunchKernelGGL((fourier_reduce<kPatternMean, true>), grid, dim3(256), 0, s, b);
    else if (b.first_wins) hipLaunchKernelGGL((fourier_reduce<kPatternMax, true>), grid, dim3(256), 0, s, b);
    else hipLaunchKernelGGL((fourier_reduce<kPatternMax, false>), grid, dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_pattern_accumulate(const PatternKind &k, const void *part_sum, const void *part_max, int ngroups,
                                     int64_t ndet, void *run_sum, void *run_max, bool first, bool finish, double count,
                                     double *out, hipStream_t s) {
    const dim3 grid((unsigned)((ndet + 255) / 256));
#define CUBO_ACC(INT, STAT, FIRST)                                                                                     \
    hipLaunchKernelGGL((pattern_accumulate<INT, STAT, FIRST>), grid, dim3(256), 0, s, part_sum, part_max, ngroups,    \
                       ndet, run_sum, run_max, first, finish, count, out)
    if (k.integer) {
        if (k.stat == kPatternMean) CUBO_ACC(true, kPatternMean, true);
        else CUBO_ACC(true, kPatternMax, true);
    } else if (k.stat == kPatternMean) {
        CUBO_ACC(false, kPatternMean, true);
    } else if (k.first_wins) {
        CUBO_ACC(false, kPatternMax, true);
    } else {
        CUBO_ACC(false, kPatternMax, false);
    }
#undef CUBO_ACC
    return hipGetLastError();
}

}  // namespace cubo
