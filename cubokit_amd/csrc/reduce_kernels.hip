// reduce_kernels.hip — the fused, cube-streaming reduction for gfx950.
//
// Replaces the three per-tile Metal kernels cubo_virtual_image, cubo_sum_and_max and cubo_center_of_mass
// (Sources/CuboVirtual/MetalReducers.swift:86-176) and the .cpu loops they mirror (CPUReducers.swift:13-97,
// CenterOfMass.swift:51-79) with one pass over the cube.
//
// Layout. The cube is [position][pixel], positions contiguous. A pattern's pixels are cut into slices; a
// wave owns one slice and walks a contiguous run of positions. Inside a slice, lane l of chunk c holds
// pixels s0 + (c*64 + l)*VEC ... + VEC-1, so every load instruction of a wave reads one contiguous span
// (1 KiB for 16-byte loads). Because a lane keeps the same pixels for every position, the column
// statistics (sum and maximum over positions) stay in registers for the whole run. The row statistics
// (pattern . weight row) are per-lane partials that a wave reduces through LDS in batches: each lane
// writes its partials, then each lane sums a row of the transposed batch — one add per value instead of
// a shuffle tree per position.
//
// Column partials of the four waves of a block are combined in LDS and written once per block (one
// partial per position group); row partials are written once per (slice, position). Two small kernels
// then sum groups (columns) and slices (rows) in a fixed order, so every result is deterministic.
// (Device-scope atomics for the column partials measured slower: their cost grew with the block count.)
#include "reduce_kernels.hpp"
#include "device_sync.hpp"

namespace cubo {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t udot2(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_udot2(__builtin_bit_cast(us2, a), __builtin_bit_cast(us2, b), c, false);
}

__device__ __forceinline__ uint32_t pk_max_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(us2, a), __builtin_bit_cast(us2, b)));
}

// The maximum step of the reference: NaN wins, and among equal values the later one is kept
// (Swift.max(a, b) = b >= a ? b : a; MetalReducers.swift:482, CPUReducers.swift:79).
__device__ __forceinline__ float max_step(float m, float x) {
    return (__builtin_isnan(m) || __builtin_isnan(x)) ? __builtin_nanf("") : (x >= m ? x : m);
}

// ------------------------------------------------------------------------------------------------
// Row reduction through LDS. ROWS rows of 64 lane values per batch; lane l sums row l % ROWS over the
// columns of its group l / ROWS, then the groups are combined with shuffles. B = ROWS / NV positions per
// batch.
// ------------------------------------------------------------------------------------------------
constexpr int kLdsStride = kWave + 1;   // one pad word per row: a row's 64 values fall in distinct banks

template <typename Acc, int NV, int ROWS>
struct RowBatch {
    static constexpr int B = ROWS / NV;
    static constexpr int G = kWave / ROWS;       // lane groups per row
    static constexpr int PER = kWave / G;         // columns each lane sums
    static_assert(B >= 1, "too many rows per position for one batch");
    Acc *buf;                                     // this wave's ROWS x kLdsStride region

    __device__ __forceinline__ void put(int b, const Acc (&acc)[NV], int lane) {
#pragma unroll
        for (int v = 0; v < NV; ++v) buf[(v * B + b) * kLdsStride + lane] = acc[v];
    }

    // The wave total of row r = lane % ROWS (v = r / B, b = r % B) in the lanes of group 0 (lane < ROWS);
    // ends with the batch region free for the next batch.
    template <typename Out>
    __device__ __forceinline__ Out total(int lane) {
        wave_sync();
        const int r = lane % ROWS, g = lane / ROWS;
        const Acc *row = buf + r * kLdsStride + g * PER;
        Out s = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) s += (Out)row[k];
#pragma unroll
        for (int off = ROWS; off < kWave; off <<= 1) s += shfl_down(s, off);
        wave_sync();
        return s;
    }

    template <typename Out>
    __device__ __forceinline__ void flush(int nb, int64_t first, Out *rowpart_slice, int lane) {
        wave_sync();
        const int r = lane % ROWS, g = lane / ROWS;
        const Acc *row = buf + r * kLdsStride + g * PER;
        Out s = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) s += (Out)row[k];
#pragma unroll
        for (int off = ROWS; off < kWave; off <<= 1) s += shfl_down(s, off);
        const int v = r / B, b = r % B;
        if (g == 0 && v < NV && b < nb) rowpart_slice[(first + b) * NV + v] = s;
        wave_sync();
    }

    __device__ __forceinline__ static uint64_t shfl_down(uint64_t x, int off) {
        uint32_t lo = __shfl_down((uint32_t)x, off, kWave), hi = __shfl_down((uint32_t)(x >> 32), off, kWave);
        return ((uint64_t)hi << 32) | lo;
    }
    __device__ __forceinline__ static double shfl_down(double x, int off) { return __shfl_down(x, off, kWave); }
};

// ------------------------------------------------------------------------------------------------
// Exact integer kernel: uint16 cube, NW weight rows of uint16 (packed in pairs), optional arithmetic
// centre-of-mass rows (w = 1: S, sum i*x, sum j*x) when the detector rows are multiples of 8 pixels.
// ------------------------------------------------------------------------------------------------
// Row totals of positions [p_begin, p_end) of a launch: slices summed in slice order, then the centre of
// mass from the totals. Threads tid, tid + nthreads, ... take one position each.
template <typename Acc>
__device__ __forceinline__ void finalize_positions(const Acc *rowpart, int nslices, int64_t npos, const RowsOut &o,
                                                   int64_t p_begin, int64_t p_end, int tid, int nthreads) {
    for (int64_t p = p_begin + tid; p < p_end; p += nthreads) {
        double com3[3] = {0.0, 0.0, 0.0};
        for (int v = 0; v < o.nv; ++v) {
            Acc acc = 0;
            for (int sl = 0; sl < nslices; ++sl) acc += rowpart[((int64_t)sl * npos + p) * o.nv + v];
            const double d = (double)acc;
            o.rows[(int64_t)v * o.total_pos + o.pos0 + p] = d;
            const int k = v - o.com_base;
            if (o.com_base >= 0 && k >= 0 && k < 3) com3[k] = d;
        }
        if (o.com_base >= 0) {
            o.com[o.pos0 + p] = com3[1] / com3[0];                    // NaN where S = 0 (CenterOfMass.swift:44-45)
            o.com[o.total_pos + o.pos0 + p] = com3[2] / com3[0];
        }
    }
}

// Position buffers per wave of fused_u16_int (NB - 1 positions in flight).
#ifndef CUBO_FUSED_NB
#define CUBO_FUSED_NB(CH) (8 / (CH) + 1)
#endif

// Two block layouts:
//  - FULL = false: a block is (slice, position group); its 4 waves cover the same slice over consecutive
//    position runs; per-slice row partials go to rowpart and a second kernel adds the slices.
//  - FULL = true (at most kFullSlices slices): a block is a position group and has one wave per slice; all
//    waves walk the same positions, so each batch of row partials is added across the waves in LDS and
//    the block writes the row totals and the centre of mass itself (to mapped host memory: they reach the
//    host while the pass still streams). One block barrier per batch.
template <int CH, int NW, bool COM1, bool FULL>
__global__ __launch_bounds__(FULL ? kWave * kFullSlices : kBlock) void fused_u16_int(IntLaunch a) {
    constexpr int NV0 = NW + (COM1 ? 3 : 0);
    constexpr int NV = NV0 > 0 ? NV0 : 1;       // row outputs (one unused row when only columns are asked)
    constexpr int ROWS = 32;
    constexpr int SLICE = kWave * CH * 8;
    constexpr int WAVES = FULL ? kFullSlices : kWavesPerBlock;
    constexpr int LDS_ROWS = WAVES * ROWS * kLdsStride;
    constexpr int LDS_COLS = FULL ? 0 : (kWavesPerBlock - 1) * CH * 12 * kWave;
    __shared__ uint32_t lds[LDS_ROWS > LDS_COLS ? LDS_ROWS : LDS_COLS];
    __shared__ uint64_t part[FULL ? 2 * kFullSlices * ROWS : 1];    // FULL: [batch parity][wave][b * NV + v]

    // the wave index is uniform; readfirstlane lets the compiler keep the position run in scalar registers
    const int lane = threadIdx.x & (kWave - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int slice = FULL ? wave : (int)(blockIdx.x % a.nslices);
    const int group = FULL ? (int)blockIdx.x : (int)(blockIdx.x / a.nslices);
    const int s0 = slice * SLICE;
    const int pairs = a.ndet >> 1;

    // Pixels past the pattern (a partial last slice) are loaded from the row start instead, so that every
    // load is unconditional: loads under a lane mask make the compiler wait for all loads in flight, which
    // drains the prefetch. Their weights are zero and their column sums are never written; only the
    // arithmetic centre-of-mass rows mask them out (vmask).
    int px[CH], pxl[CH];
    bool valid[CH];
    uint32_t vmask[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        px[c] = s0 + (c * kWave + lane) * 8;
        valid[c] = px[c] < a.ndet;
        pxl[c] = valid[c] ? px[c] : 0;
        vmask[c] = valid[c] ? ~0u : 0u;
    }
    uint32_t w[NW > 0 ? NW : 1][CH][4];
#pragma unroll
    for (int v = 0; v < NW; ++v)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            uint4 t = valid[c] ? *reinterpret_cast<const uint4 *>(a.wpk + (size_t)v * pairs + (px[c] >> 1))
                               : make_uint4(0, 0, 0, 0);
            w[v][c][0] = t.x; w[v][c][1] = t.y; w[v][c][2] = t.z; w[v][c][3] = t.w;
        }
    uint32_t ic[CH], jc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        ic[c] = COM1 ? (uint32_t)(px[c] / a.qx) : 0u;
        jc[c] = COM1 ? (uint32_t)(px[c] % a.qx) : 0u;
    }

    uint32_t csum[CH][8], cmax[CH][4];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
#pragma unroll
        for (int e = 0; e < 8; ++e) csum[c][e] = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) cmax[c][e] = 0;
    }

    RowBatch<uint32_t, NV, ROWS> rb{lds + wave * ROWS * kLdsStride};
    uint64_t *rowslice = FULL ? nullptr : a.rowpart + (size_t)slice * a.npos * NV;

    const int64_t gbeg = (int64_t)group * a.ppg;
    const int64_t gend = min(a.npos, gbeg + a.ppg);
    const int64_t pbeg = FULL ? gbeg : min(gend, gbeg + (int64_t)wave * a.ppw);
    const int64_t pend = FULL ? gend : min(gend, pbeg + a.ppw);

    // FULL: the wave totals of a batch of nb positions go to part[parity]; after the barrier the first nb
    // threads add the slices of one position each and write its totals. The parity alternates, so the
    // next batch never overwrites a buffer that is still being read (its readers pass the next barrier
    // first).
    int parity = 0;
    auto combine = [&](int nb, int64_t first_pos) {
        const uint64_t t = rb.template total<uint64_t>(lane);
        const int r = lane % ROWS, v = r / rb.B, bb = r % rb.B;
        if (lane < ROWS && v < NV && bb < nb) part[(parity * kFullSlices + wave) * ROWS + bb * NV + v] = t;
        lds_barrier();                                   // LDS only: the prefetched positions stay in flight
        if ((int)threadIdx.x < nb) {
            const int pb = threadIdx.x;
            const RowsOut &o = a.out;
            double com3[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int vv = 0; vv < NV; ++vv) {
                uint64_t sum = 0;
                for (int ww = 0; ww < a.nslices; ++ww) sum += part[(parity * kFullSlices + ww) * ROWS + pb * NV + vv];
                const double d = (double)sum;
                o.rows[(int64_t)vv * o.total_pos + o.pos0 + first_pos + pb] = d;
                const int k = vv - o.com_base;
                if (o.com_base >= 0 && k >= 0 && k < 3) com3[k] = d;
            }
            if (o.com_base >= 0) {
                o.com[o.pos0 + first_pos + pb] = com3[1] / com3[0];      // NaN where S = 0
                o.com[o.total_pos + o.pos0 + first_pos + pb] = com3[2] / com3[0];
            }
        }
        parity ^= 1;
    };

    auto load = [&](u32x4 (&d)[CH], int64_t p) {
        const uint16_t *row = a.cube + p * a.ndet;
#pragma unroll
        for (int c = 0; c < CH; ++c)
            d[c] = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(row + pxl[c]));
    };

    int b = 0;
    int64_t first = pbeg;
    auto process = [&](const u32x4 (&d)[CH]) {
        uint32_t acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const uint32_t x[4] = {d[c].x, d[c].y, d[c].z, d[c].w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                csum[c][2 * k] += x[k] & 0xffffu;
                csum[c][2 * k + 1] += x[k] >> 16;
                cmax[c][k] = pk_max_u16(cmax[c][k], x[k]);
#pragma unroll
                for (int v = 0; v < NW; ++v) acc[v] = udot2(x[k], w[v][c][k], acc[v]);
            }
            if constexpr (COM1) {
                uint32_t cs = udot2(x[0], 0x00010001u, 0u);
                cs = udot2(x[1], 0x00010001u, cs);
                cs = udot2(x[2], 0x00010001u, cs);
                cs = udot2(x[3], 0x00010001u, cs);
                uint32_t jj = udot2(x[0], 0x00010000u, 0u);       // pixel offsets (0, 1)
                jj = udot2(x[1], 0x00030002u, jj);                 // (2, 3)
                jj = udot2(x[2], 0x00050004u, jj);                 // (4, 5)
                jj = udot2(x[3], 0x00070006u, jj);                 // (6, 7)
                cs &= vmask[c];
                jj &= vmask[c];
                acc[NW] += cs;
                acc[NW + 1] += ic[c] * cs;
                acc[NW + 2] += jc[c] * cs + jj;
            }
        }
        if constexpr (NV0 > 0) {
            rb.put(b, acc, lane);
            if (++b == rb.B) {
                if constexpr (FULL) combine(b, first);
                else rb.flush(b, first, rowslice, lane);
                first += b;
                b = 0;
            }
        }
    };

    // NB position buffers in rotation: NB - 1 positions in flight while one is reduced. The prefetches are
    // unconditional
    // (past the run they repeat its last position, whose copy is never reduced), so the compiler can count
    // them and wait only for the position it reduces.
    constexpr int NB = CUBO_FUSED_NB(CH);
    u32x4 buf[NB][CH];
    int64_t p = pbeg;
    if (pbeg < pend) {
        const int64_t last = pend - 1;
#pragma unroll
        for (int j = 0; j < NB - 1; ++j) load(buf[j], min(p + j, last));
        for (; p + NB <= pend; p += NB) {
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                load(buf[(j + NB - 1) % NB], min(p + j + NB - 1, last));
                process(buf[j]);
            }
        }
        const int64_t rest = pend - p;                      // < NB; held in buf[0 .. rest - 1]
#pragma unroll
        for (int j = 0; j < NB - 1; ++j)
            if (j < rest) process(buf[j]);
    }
    if constexpr (NV0 > 0) {
        if (b > 0) {
            if constexpr (FULL) combine(b, first);
            else rb.flush(b, first, rowslice, lane);
        }
    }

    if constexpr (FULL) {
        // one wave per slice: the column partials need no combining
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (valid[c]) {
                uint32_t *cs = a.colsum + (size_t)group * a.ndet + px[c];
                uint32_t *cm = a.colmax + ((size_t)group * a.ndet + px[c]) / 2;    // the maxima stay packed
                reinterpret_cast<uint4 *>(cs)[0] = make_uint4(csum[c][0], csum[c][1], csum[c][2], csum[c][3]);
                reinterpret_cast<uint4 *>(cs)[1] = make_uint4(csum[c][4], csum[c][5], csum[c][6], csum[c][7]);
                reinterpret_cast<uint4 *>(cm)[0] = make_uint4(cmax[c][0], cmax[c][1], cmax[c][2], cmax[c][3]);
            }
        return;
    }

    // Combine the four waves' column partials (same pixels, consecutive position runs) and write one
    // partial per group.
    __syncthreads();
    if (wave > 0) {
        uint32_t *dst = lds + (size_t)(wave - 1) * CH * 12 * kWave;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
#pragma unroll
            for (int e = 0; e < 8; ++e) dst[(c * 12 + e) * kWave + lane] = csum[c][e];
#pragma unroll
            for (int e = 0; e < 4; ++e) dst[(c * 12 + 8 + e) * kWave + lane] = cmax[c][e];
        }
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
#pragma unroll
            for (int ww = 0; ww < kWavesPerBlock - 1; ++ww) {
                const uint32_t *src = lds + (size_t)ww * CH * 12 * kWave;
#pragma unroll
                for (int e = 0; e < 8; ++e) csum[c][e] += src[(c * 12 + e) * kWave + lane];
#pragma unroll
                for (int e = 0; e < 4; ++e) cmax[c][e] = pk_max_u16(cmax[c][e], src[(c * 12 + 8 + e) * kWave + lane]);
            }
            if (valid[c]) {
                uint32_t *cs = a.colsum + (size_t)group * a.ndet + px[c];
                uint32_t *cm = a.colmax + ((size_t)group * a.ndet + px[c]) / 2;    // the maxima stay packed
                reinterpret_cast<uint4 *>(cs)[0] = make_uint4(csum[c][0], csum[c][1], csum[c][2], csum[c][3]);
                reinterpret_cast<uint4 *>(cs)[1] = make_uint4(csum[c][4], csum[c][5], csum[c][6], csum[c][7]);
                reinterpret_cast<uint4 *>(cm)[0] = make_uint4(cmax[c][0], cmax[c][1], cmax[c][2], cmax[c][3]);
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// FP64 kernel: any pattern length, uint16 or Float32 cube, NW weight rows in Double. One pixel per
// lane per chunk (2- or 4-byte loads, coalesced across the wave).
// ------------------------------------------------------------------------------------------------
template <typename T, int NW>
__global__ __launch_bounds__(kBlock) void fused_f64(F64Launch a) {
    constexpr int CH = kF64Chunks;
    constexpr int ROWS = 16;
    constexpr int LDS_ROWS = kWavesPerBlock * ROWS * kLdsStride;        // doubles
    constexpr int LDS_COLS = (kWavesPerBlock - 1) * CH * 3 * kWave / 2 + 1;  // doubles (sum: 2 words, max: 1)
    __shared__ double lds[LDS_ROWS > LDS_COLS ? LDS_ROWS : LDS_COLS];

    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int slice = blockIdx.x % a.nslices, group = blockIdx.x / a.nslices;
    const int s0 = slice * kF64Slice;
    const T *cube = reinterpret_cast<const T *>(a.cube);

    int px[CH];
    int pxl[CH];
    bool valid[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        px[c] = s0 + c * kWave + lane;
        valid[c] = px[c] < a.ndet;
        pxl[c] = valid[c] ? px[c] : 0;
    }
    double csum[CH];
    float cmax[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) { csum[c] = -0.0; cmax[c] = -__builtin_inff(); }

    RowBatch<double, NW, ROWS> rb{lds + wave * ROWS * kLdsStride};
    double *rowslice = a.rowpart + (size_t)slice * a.npos * NW;

    const int64_t gbeg = (int64_t)group * a.ppg;
    const int64_t gend = min(a.npos, gbeg + a.ppg);
    const int64_t pbeg = min(gend, gbeg + (int64_t)wave * a.ppw);
    const int64_t pend = min(gend, pbeg + a.ppw);

    int b = 0;
    int64_t first = pbeg;
    for (int64_t p = pbeg; p < pend; ++p) {
        const T *row = cube + p * a.ndet;
        // unconditional loads (pixels past the pattern read the row start; see fused_u16_int), and the
        // value of such a pixel is replaced by 0 where it is used, so it adds nothing to the rows; its
        // column sums are never written
        T xv[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) xv[c] = __builtin_nontemporal_load(row + pxl[c]);
        double acc[NW];
#pragma unroll
        for (int v = 0; v < NW; ++v) acc[v] = 0.0;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const float xf = valid[c] ? (float)xv[c] : 0.f;
            const double x = (double)xf;
            csum[c] += x;
            cmax[c] = max_step(cmax[c], xf);
#pragma unroll
            for (int v = 0; v < NW; ++v) acc[v] = fma(x, a.w[(size_t)v * a.ndet + pxl[c]], acc[v]);
        }
        rb.put(b, acc, lane);
        if (++b == rb.B) {
            rb.flush(b, first, rowslice, lane);
            first += b;
            b = 0;
        }
    }
    if (b > 0) rb.flush(b, first, rowslice, lane);

    __syncthreads();
    double *dsum = lds;
    float *dmax = reinterpret_cast<float *>(lds + (kWavesPerBlock - 1) * CH * kWave);
    if (wave > 0) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            dsum[((wave - 1) * CH + c) * kWave + lane] = csum[c];
            dmax[((wave - 1) * CH + c) * kWave + lane] = cmax[c];
        }
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int c = 0; c < CH; ++c) {
#pragma unroll
            for (int ww = 0; ww < kWavesPerBlock - 1; ++ww) {      // waves in position order
                csum[c] += dsum[(ww * CH + c) * kWave + lane];
                cmax[c] = max_step(cmax[c], dmax[(ww * CH + c) * kWave + lane]);
            }
            if (valid[c]) {
                a.colsum[(size_t)group * a.ndet + px[c]] = csum[c];
                a.colmax[(size_t)group * a.ndet + px[c]] = cmax[c];
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Vectorised FP64 kernel: uint16 or Float32 cube whose pattern length is a multiple of 16 bytes, at most
// kVecMasks mask rows, optionally the centre-of-mass rows of unit weight. The same outputs as fused_f64, with:
//  - octets: a block of kVecWaves waves covers kVecWaves consecutive slices of kVecSlice pixels (one per wave)
//    and all its waves walk the same positions of its group, so a compute unit reads whole spans of each
//    pattern in order; each batch of row partials is added across the waves in LDS (one block barrier per
//    batch). With one octet per pattern the block writes the row totals and the centre of mass itself (to
//    mapped host memory, while the pass still streams); with several, one partial per octet goes to rowpart;
//  - 16-byte loads: lane l of chunk c holds pixels s0 + (c*64 + l)*VEC ... + VEC-1 (VEC = 8 uint16 or 4 Float32),
//    CH = 1 or 2 chunks per slice (8 pixels per lane), NB - 1 positions in flight in a register ring;
//  - the mask weights of the lane's pixels in registers as Double, loaded once;
//  - the centre-of-mass rows arithmetic. uint16: exact integer sums per chunk (packed dot products), as in
//    fused_u16_int. Float32: S and sum i*x from the chunk's sum (one detector row per chunk, so i is the chunk's
//    row; i * sum gives the per-pixel products' special values too), sum j*x per pixel as (j0 + e) * x (a pixel's
//    own column: 0 * inf is NaN only where j = 0, as CenterOfMass.swift:58-63 computes it);
//  - column sums and maxima of each wave's pixels in registers (integers for uint16), written once per group.
// TAIL: the pattern length is not a multiple of the octet; chunks past it load the row start and their values are
// replaced by 0 where they are used (their column statistics are never written).
// ------------------------------------------------------------------------------------------------
constexpr int f64_rows(int nw) { return nw < 1 ? 1 : nw <= 6 ? nw : nw <= 8 ? 8 : nw <= 11 ? 11 : 16; }

// An empty asm that reads and writes v: the update that produced v stays where it is written (LLVM would otherwise
// sink the column-statistic updates of every position of an unrolled ring into the loop's last block, keeping all
// of their data live across the row batches).
template <typename V>
__device__ __forceinline__ void keep(V &v) { asm volatile("" : "+v"(v)); }

// Position buffers per wave (NB - 1 positions in flight) and LDS rows per wave of the row batches.
#ifndef CUBO_VEC_NB
#define CUBO_VEC_NB(CH) (8 / (CH) + 1)
#endif
#ifndef CUBO_VEC_ROWS
#define CUBO_VEC_ROWS 32
#endif

template <typename T, int NWM, bool COM, bool TAIL>
__global__ __launch_bounds__(kWave * kVecWaves) void fused_vec_f64(F64Launch a) {
    constexpr bool U16 = sizeof(T) == 2;
    constexpr int VEC = 16 / (int)sizeof(T);
    constexpr int CH = kVecSlice / (kWave * VEC);
    constexpr int NV = f64_rows(NWM + (COM ? 3 : 0));
    constexpr int ROWS = CUBO_VEC_ROWS;
    constexpr int NB = CUBO_VEC_NB(CH);
    __shared__ double lds[kVecWaves * ROWS * kLdsStride];
    __shared__ double part[2 * kVecWaves * ROWS];          // [batch parity][wave][b * NV + v]

    const int lane = threadIdx.x & (kWave - 1), wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int octet = blockIdx.x % a.nslices, group = blockIdx.x / a.nslices;
    const int s0 = (octet * kVecWaves + wave) * kVecSlice;

    int px[CH], pxl[CH];
    bool valid[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        px[c] = s0 + (c * kWave + lane) * VEC;
        valid[c] = !TAIL || px[c] < a.ndet;
        pxl[c] = valid[c] ? px[c] : 0;
    }
    double w[NWM > 0 ? NWM : 1][CH][VEC];
#pragma unroll
    for (int v = 0; v < NWM; ++v)
#pragma unroll
        for (int c = 0; c < CH; ++c)
#pragma unroll
            for (int e = 0; e < VEC; ++e) w[v][c][e] = valid[c] ? a.w[(size_t)v * a.ndet + px[c] + e] : 0.0;
    double icd[CH], jcd[CH];
    uint32_t vmask[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        icd[c] = COM ? (double)(pxl[c] / a.qx) : 0.0;
        jcd[c] = COM ? (double)(pxl[c] % a.qx) : 0.0;
        vmask[c] = valid[c] ? ~0u : 0u;
    }

    // column statistics: uint16 sums (uint32) and packed maxima; Float32 Double sums seeded at -0 and maxima
    uint32_t usum[U16 ? CH : 1][VEC], umax[U16 ? CH : 1][VEC / 2];
    double dsum[U16 ? 1 : CH][VEC];
    float fmaxv[U16 ? 1 : CH][VEC];
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if constexpr (U16) {
                usum[c][e] = 0u;
                if (e < VEC / 2) umax[c][e] = 0u;
            } else {
                dsum[c][e] = -0.0;
                fmaxv[c][e] = -__builtin_inff();
            }
        }

    RowBatch<double, NV, ROWS> rb{lds + wave * ROWS * kLdsStride};
    const int64_t gbeg = (int64_t)group * a.ppg;
    const int64_t pend = min(a.npos, gbeg + a.ppg);

    // The wave totals of a batch of nb positions go to part[parity]; after the barrier the first nb threads add
    // the octet's waves (slice order) for one position each and write its totals, or its octet partials. The
    // parity alternates, so the next batch never overwrites a buffer that is still being read.
    int parity = 0;
    auto combine = [&](int nb, int64_t first_pos) {
        const double t = rb.template total<double>(lane);
        const int r = lane % ROWS, v = r / rb.B, bb = r % rb.B;
        if (lane < ROWS && v < NV && bb < nb) part[(parity * kVecWaves + wave) * ROWS + bb * NV + v] = t;
        lds_barrier();                                   // LDS only: the prefetched positions stay in flight
        if ((int)threadIdx.x < nb) {
            const int pb = threadIdx.x;
            const int64_t pp = first_pos + pb;
            double tot[NV];
#pragma unroll
            for (int vv = 0; vv < NV; ++vv) {
                double sum = 0.0;
#pragma unroll
                for (int ww = 0; ww < kVecWaves; ++ww) sum += part[(parity * kVecWaves + ww) * ROWS + pb * NV + vv];
                tot[vv] = sum;
            }
            if (a.nslices == 1) {
                const RowsOut &o = a.out;
#pragma unroll
                for (int vv = 0; vv < NV; ++vv) o.rows[(int64_t)vv * o.total_pos + o.pos0 + pp] = tot[vv];
                if constexpr (COM) {                                     // the CoM rows follow the masks
                    o.com[o.pos0 + pp] = tot[NWM + 1] / tot[NWM];                 // NaN where S = 0
                    o.com[o.total_pos + o.pos0 + pp] = tot[NWM + 2] / tot[NWM];
                }
            } else {
#pragma unroll
                for (int vv = 0; vv < NV; ++vv) a.rowpart[((int64_t)octet * a.npos + pp) * NV + vv] = tot[vv];
            }
        }
        parity ^= 1;
    };

    auto load = [&](u32x4 (&d)[CH], int64_t p) {
        const T *row = reinterpret_cast<const T *>(a.cube) + p * a.ndet;
#pragma unroll
        for (int c = 0; c < CH; ++c)
            d[c] = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(row + pxl[c]));
    };

    int b = 0;
    int64_t first = gbeg;
    auto process = [&](const u32x4 (&d)[CH]) {
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const uint32_t raw[4] = {d[c].x & vmask[c], d[c].y & vmask[c], d[c].z & vmask[c], d[c].w & vmask[c]};
            double x[VEC];
            if constexpr (U16) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t lo = raw[k] & 0xffffu, hi = raw[k] >> 16;
                    usum[c][2 * k] += lo;
                    usum[c][2 * k + 1] += hi;
                    umax[c][k] = pk_max_u16(umax[c][k], raw[k]);
                    keep(usum[c][2 * k]);
                    keep(usum[c][2 * k + 1]);
                    keep(umax[c][k]);
                    x[2 * k] = (double)lo;
                    x[2 * k + 1] = (double)hi;
                }
                if constexpr (COM) {
                    uint32_t cs = udot2(raw[0], 0x00010001u, 0u);
                    cs = udot2(raw[1], 0x00010001u, cs);
                    cs = udot2(raw[2], 0x00010001u, cs);
                    cs = udot2(raw[3], 0x00010001u, cs);
                    uint32_t jj = udot2(raw[0], 0x00010000u, 0u);
                    jj = udot2(raw[1], 0x00030002u, jj);
                    jj = udot2(raw[2], 0x00050004u, jj);
                    jj = udot2(raw[3], 0x00070006u, jj);
                    const double csd = (double)cs;
                    acc[NWM] += csd;
                    acc[NWM + 1] = fma(icd[c], csd, acc[NWM + 1]);
                    acc[NWM + 2] += fma(jcd[c], csd, (double)jj);     // exact: integers below 2^53
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float xf = __builtin_bit_cast(float, raw[e]);
                    x[e] = (double)xf;
                    dsum[c][e] += x[e];
                    fmaxv[c][e] = max_step(fmaxv[c][e], xf);
                    keep(dsum[c][e]);
                    keep(fmaxv[c][e]);
                }
                if constexpr (COM) {
                    const double cs = (x[0] + x[1]) + (x[2] + x[3]);
                    acc[NWM] += cs;
                    acc[NWM + 1] = fma(icd[c], cs, acc[NWM + 1]);
#pragma unroll
                    for (int e = 0; e < 4; ++e) acc[NWM + 2] = fma(jcd[c] + (double)e, x[e], acc[NWM + 2]);
                }
            }
#pragma unroll
            for (int v = 0; v < NWM; ++v)
#pragma unroll
                for (int e = 0; e < VEC; ++e) acc[v] = fma(x[e], w[v][c][e], acc[v]);
        }
        rb.put(b, acc, lane);
        if (++b == rb.B) {
            combine(b, first);
            first += b;
            b = 0;
        }
    };

    u32x4 buf[NB][CH];
    int64_t p = gbeg;
    if (gbeg < pend) {
        const int64_t last = pend - 1;
#pragma unroll
        for (int j = 0; j < NB - 1; ++j) load(buf[j], min(p + j, last));
        for (; p + NB <= pend; p += NB) {
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                load(buf[(j + NB - 1) % NB], min(p + j + NB - 1, last));
                process(buf[j]);
            }
        }
        const int64_t rest = pend - p;
#pragma unroll
        for (int j = 0; j < NB - 1; ++j)
            if (j < rest) process(buf[j]);
    }
    if (b > 0) combine(b, first);

    // each wave's own pixels: no combining
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        if (!valid[c]) continue;
        double *cs = a.colsum + (size_t)group * a.ndet + px[c];
        float *cm = a.colmax + (size_t)group * a.ndet + px[c];
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            if constexpr (U16) {
                cs[e] = (double)usum[c][e];
                cm[e] = (float)((umax[c][e / 2] >> (16 * (e & 1))) & 0xffffu);
            } else {
                cs[e] = dsum[c][e];
                cm[e] = fmaxv[c][e];
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Second-stage kernels.
// ------------------------------------------------------------------------------------------------
template <typename Acc>
__global__ void finalize_rows(const Acc *rowpart, int nslices, int64_t npos, RowsOut o) {
    finalize_positions<Acc>(rowpart, nslices, npos, o, (int64_t)blockIdx.x * blockDim.x,
                            min(npos, (int64_t)(blockIdx.x + 1) * blockDim.x), threadIdx.x, blockDim.x);
}

// The same with the row count fixed: a position's NV partials of U slices are loaded before any is added, so a
// thread has U * NV loads in flight; the adds keep slice order.
template <typename Acc, int NV>
__global__ __launch_bounds__(256) void finalize_rows_nv(const Acc *rowpart, int nslices, int64_t npos, RowsOut o) {
    constexpr int U = NV <= 4 ? 8 : NV <= 8 ? 4 : 2;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npos) return;
    Acc acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0;
    int sl = 0;
    for (; sl + U <= nslices; sl += U) {
        Acc t[U][NV];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < NV; ++v) t[u][v] = rowpart[((int64_t)(sl + u) * npos + p) * NV + v];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += t[u][v];
    }
    for (; sl < nslices; ++sl)
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += rowpart[((int64_t)sl * npos + p) * NV + v];
    double com3[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double d = (double)acc[v];
        o.rows[(int64_t)v * o.total_pos + o.pos0 + p] = d;
        const int k = v - o.com_base;
        if (o.com_base >= 0 && k >= 0 && k < 3) com3[k] = d;
    }
    if (o.com_base >= 0) {
        o.com[o.pos0 + p] = com3[1] / com3[0];                        // NaN where S = 0 (CenterOfMass.swift:44-45)
        o.com[o.total_pos + o.pos0 + p] = com3[2] / com3[0];
    }
}

// 64 pixels x 16 group phases per block: each thread sums every sixteenth group of one pixel (eight loads in
// flight), LDS adds the sixteen. Integer sums: the order does not change them.
constexpr int kColPhases = 16;
__global__ __launch_bounds__(64 * kColPhases) void accumulate_cols_int(const uint32_t *colsum, const uint32_t *colmax,
                                                                       int ngroups, int ndet, uint64_t *sum,
                                                                       uint32_t *max, bool first, bool finish,
                                                                       double count, double *mean_out,
                                                                       double *max_out) {
    __shared__ uint64_t ps[kColPhases][64];
    __shared__ uint32_t pm[kColPhases][64];
    const int lx = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lx;
    uint64_t s = 0;
    uint32_t m = 0;
    if (i < ndet) {
        // the maxima are packed uint16 pairs: pixel i is the half i % 2 of word i / 2
        const int sh = 16 * (i & 1);
        const size_t half = (size_t)ndet / 2;
        int g = q;
        for (; g + 7 * kColPhases < ngroups; g += 8 * kColPhases) {
            uint32_t vs[8], vm[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                vs[u] = colsum[(size_t)(g + kColPhases * u) * ndet + i];
                vm[u] = (colmax[(size_t)(g + kColPhases * u) * half + i / 2] >> sh) & 0xffffu;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) { s += vs[u]; m = vm[u] > m ? vm[u] : m; }
        }
        for (; g < ngroups; g += kColPhases) {
            s += colsum[(size_t)g * ndet + i];
            const uint32_t v = (colmax[(size_t)g * half + i / 2] >> sh) & 0xffffu;
            m = v > m ? v : m;
        }
    }
    ps[q][lx] = s;
    pm[q][lx] = m;
    __syncthreads();
    if (q == 0 && i < ndet) {
        s = 0;
        m = 0;
        for (int k = 0; k < kColPhases; ++k) {
            s += ps[k][lx];
            m = pm[k][lx] > m ? pm[k][lx] : m;
        }
        if (!first) {
            s += sum[i];
            m = max[i] > m ? max[i] : m;
        }
        if (finish) {
            mean_out[i] = (double)s / count;
            max_out[i] = (double)m;
        } else {
            sum[i] = s;
            max[i] = m;
        }
    }
}

__global__ void accumulate_cols_f64(const double *colsum, const float *colmax, int ngroups, int ndet, double *sum,
                                    float *max, bool first, bool finish, double count, double *mean_out,
                                    double *max_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndet) return;
    // the running sum starts at -0 (an all -0 column keeps its sign) and the maximum at -inf
    double s = first ? -0.0 : sum[i];
    float m = first ? -__builtin_inff() : max[i];
    int g = 0;
    for (; g + 8 <= ngroups; g += 8) {                   // loads batched, adds still in group order
        double vs[8];
        float vm[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            vs[u] = colsum[(size_t)(g + u) * ndet + i];
            vm[u] = colmax[(size_t)(g + u) * ndet + i];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) { s += vs[u]; m = max_step(m, vm[u]); }
    }
    for (; g < ngroups; ++g) {
        s += colsum[(size_t)g * ndet + i];
        m = max_step(m, colmax[(size_t)g * ndet + i]);
    }
    if (finish) {
        mean_out[i] = s / count;
        max_out[i] = (double)m;
    } else {
        sum[i] = s;
        max[i] = m;
    }
}

__global__ void copy_bytes(const u32x4 *src, u32x4 *dst, uint64_t n16, const uint8_t *tail_src, uint8_t *tail_dst,
                           uint32_t tail) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
    if (blockIdx.x == 0 && threadIdx.x < tail) tail_dst[threadIdx.x] = tail_src[threadIdx.x];
}


// ------------------------------------------------------------------------------------------------
// Host launchers.
// ------------------------------------------------------------------------------------------------
int int_slice_pixels(int chunks) { return kWave * chunks * 8; }

#define CUBO_INT_CASES(X) \
    X(4, 0, true) X(4, 0, false) X(4, 1, true) X(4, 2, true) X(4, 4, true) X(2, 8, true) \
    X(4, 1, false) X(4, 2, false) X(4, 4, false) X(2, 8, false) X(1, 16, false) \
    X(2, 0, true) X(2, 1, true) X(2, 2, true) X(2, 4, true) \
    X(2, 1, false) X(2, 2, false) X(2, 4, false) \
    X(1, 0, true) X(1, 1, true) X(1, 2, true) X(1, 4, true) X(1, 8, true) \
    X(1, 1, false) X(1, 2, false) X(1, 4, false) X(1, 8, false)

bool supported_int(int nw, int chunks, bool com_ones) {
#define X(CH, NW, C1) if (chunks == CH && nw == NW && com_ones == C1) return true;
    CUBO_INT_CASES(X)
#undef X
    return false;
}

hipError_t launch_int(const IntLaunch &a, int nw, int chunks, bool com_ones, bool full, hipStream_t s) {
    if (full && a.nslices > kFullSlices) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(full ? a.ngroups : a.nslices * a.ngroups)), block(full ? kWave * a.nslices : kBlock);
#define X(CH, NW, C1) \
    if (chunks == CH && nw == NW && com_ones == C1) { \
        if (full) hipLaunchKernelGGL((fused_u16_int<CH, NW, C1, true>), grid, block, 0, s, a); \
        else hipLaunchKernelGGL((fused_u16_int<CH, NW, C1, false>), grid, block, 0, s, a); \
        return hipGetLastError(); \
    }
    CUBO_INT_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

bool full_without_spills(int nw, int chunks, bool com_ones) {
    hipFuncAttributes at{};
    hipError_t e = hipErrorInvalidValue;
#define X(CH, NW, C1) \
    if (chunks == CH && nw == NW && com_ones == C1) \
        e = hipFuncGetAttributes(&at, reinterpret_cast<const void *>(fused_u16_int<CH, NW, C1, true>));
    CUBO_INT_CASES(X)
#undef X
    return e == hipSuccess && at.localSizeBytes == 0;
}

int blocks_per_cu_int(int nw, int chunks, bool com_ones, bool full, int nslices) {
    int n = 0;
    const int threads = full ? kWave * nslices : kBlock;
#define X(CH, NW, C1) \
    if (chunks == CH && nw == NW && com_ones == C1) { \
        const hipError_t e = full ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_u16_int<CH, NW, C1, true>, \
                                                                                 threads, 0) \
                                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_u16_int<CH, NW, C1, false>, \
                                                                                 threads, 0); \
        if (e != hipSuccess) n = 1; \
    }
    CUBO_INT_CASES(X)
#undef X
    return n;
}

#define CUBO_F64_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(11) X(16)

bool supported_f64(int nw) {
#define X(NW) if (nw == NW) return true;
    CUBO_F64_CASES(X)
#undef X
    return false;
}

// The vectorised kernel applies to a pattern of whole 16-byte pieces, at most kVecMasks masks, and no centre of
// mass or one of unit weight whose detector rows are whole pieces.
bool vec_f64_applies(const F64Launch &a, bool is_u16) {
    const int vec = is_u16 ? 8 : 4;
    return a.vec_ok && a.ndet % vec == 0 && a.nmask >= 0 && a.nmask <= kVecMasks && a.com >= 0 && a.com <= 1 &&
           (!a.com || a.qx % vec == 0) && ((uintptr_t)a.cube & 15) == 0;
}

#define CUBO_VEC_CASES(X) X(0) X(1) X(2) X(3) X(4)

hipError_t launch_f64(const F64Launch &a, bool is_u16, int nw, hipStream_t s) {
    const dim3 grid((unsigned)(a.nslices * a.ngroups)), block(kBlock);
    if (vec_f64_applies(a, is_u16)) {
        if (nw != f64_rows(a.nmask + 3 * a.com)) return hipErrorInvalidValue;
        const bool tail = a.ndet % kVecOctet != 0;
        const dim3 block(kWave * kVecWaves);
#define V(T, M, C, TL) hipLaunchKernelGGL((fused_vec_f64<T, M, C, TL>), grid, block, 0, s, a)
#define X(M) \
        if (a.nmask == M) { \
            if (is_u16) { \
                if (a.com) { if (tail) V(uint16_t, M, true, true); else V(uint16_t, M, true, false); } \
                else { if (tail) V(uint16_t, M, false, true); else V(uint16_t, M, false, false); } \
            } else { \
                if (a.com) { if (tail) V(float, M, true, true); else V(float, M, true, false); } \
                else { if (tail) V(float, M, false, true); else V(float, M, false, false); } \
            } \
            return hipGetLastError(); \
        }
        CUBO_VEC_CASES(X)
#undef X
#undef V
        return hipErrorInvalidValue;
    }
#define X(NW) \
    if (nw == NW) { \
        if (is_u16) hipLaunchKernelGGL((fused_f64<uint16_t, NW>), grid, block, 0, s, a); \
        else hipLaunchKernelGGL((fused_f64<float, NW>), grid, block, 0, s, a); \
        return hipGetLastError(); \
    }
    CUBO_F64_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

int blocks_per_cu_f64(const F64Launch &a, bool is_u16, int nw) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    if (vec_f64_applies(a, is_u16)) {
        // the TAIL instantiations share the register use of the others
#define X(M) \
        if (a.nmask == M) \
            e = is_u16 ? (a.com ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_vec_f64<uint16_t, M, true, false>, kWave * kVecWaves, 0) \
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_vec_f64<uint16_t, M, false, false>, kWave * kVecWaves, 0)) \
                       : (a.com ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_vec_f64<float, M, true, false>, kWave * kVecWaves, 0) \
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_vec_f64<float, M, false, false>, kWave * kVecWaves, 0));
        CUBO_VEC_CASES(X)
#undef X
        return e == hipSuccess && n > 0 ? n : 1;
    }
#define X(NW) \
    if (nw == NW) { \
        e = is_u16 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_f64<uint16_t, NW>, kBlock, 0) \
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fused_f64<float, NW>, kBlock, 0); \
    }
    CUBO_F64_CASES(X)
#undef X
    return e == hipSuccess && n > 0 ? n : 1;
}

int f64_kernel(const F64Launch &a, bool is_u16) { return vec_f64_applies(a, is_u16) ? 2 : 1; }

int f64_slice_pixels(const F64Launch &a, bool is_u16) { return vec_f64_applies(a, is_u16) ? kVecOctet : kF64Slice; }

static unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// Row counts of the integer layouts (padded weight rows, plus three arithmetic CoM rows) and of the FP64 ones.
#define CUBO_FINALIZE_NV(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(11) X(16) X(19)

template <typename Acc>
static hipError_t launch_finalize(const Acc *rowpart, int nslices, int64_t npos, const RowsOut &out, hipStream_t s) {
    const dim3 grid(blocks_for(npos, 256)), block(256);
#define X(NV) \
    if (out.nv == NV) { \
        hipLaunchKernelGGL((finalize_rows_nv<Acc, NV>), grid, block, 0, s, rowpart, nslices, npos, out); \
        return hipGetLastError(); \
    }
    CUBO_FINALIZE_NV(X)
#undef X
    hipLaunchKernelGGL(finalize_rows<Acc>, grid, block, 0, s, rowpart, nslices, npos, out);
    return hipGetLastError();
}

hipError_t launch_finalize_rows_u64(const uint64_t *rowpart, int nslices, int64_t npos, const RowsOut &out,
                                    hipStream_t s) {
    return launch_finalize<uint64_t>(rowpart, nslices, npos, out, s);
}

hipError_t launch_finalize_rows_f64(const double *rowpart, int nslices, int64_t npos, const RowsOut &out,
                                    hipStream_t s) {
    return launch_finalize<double>(rowpart, nslices, npos, out, s);
}

hipError_t launch_accumulate_cols_int(const uint32_t *colsum, const uint32_t *colmax, int ngroups, int ndet,
                                      uint64_t *sum, uint32_t *max, bool first, bool finish, double count,
                                      double *mean_out, double *max_out, hipStream_t s) {
    hipLaunchKernelGGL(accumulate_cols_int, dim3(blocks_for(ndet, 64)), dim3(64 * kColPhases), 0, s, colsum, colmax,
                       ngroups, ndet, sum, max, first, finish, count, mean_out, max_out);
    return hipGetLastError();
}

hipError_t launch_accumulate_cols_f64(const double *colsum, const float *colmax, int ngroups, int ndet,
                                      double *sum, float *max, bool first, bool finish, double count,
                                      double *mean_out, double *max_out, hipStream_t s) {
    hipLaunchKernelGGL(accumulate_cols_f64, dim3(blocks_for(ndet, 256)), dim3(256), 0, s, colsum, colmax, ngroups,
                       ndet, sum, max, first, finish, count, mean_out, max_out);
    return hipGetLastError();
}

hipError_t launch_copy_bytes(const void *src, void *dst, uint64_t n, hipStream_t s) {
    const uint64_t n16 = n / 16;
    const uint32_t tail = (uint32_t)(n % 16);
    hipLaunchKernelGGL(copy_bytes, dim3(4096), dim3(256), 0, s, (const u32x4 *)src, (u32x4 *)dst, n16,
                       (const uint8_t *)src + n16 * 16, (uint8_t *)dst + n16 * 16, tail);
    return hipGetLastError();
}


}  // namespace cubo
