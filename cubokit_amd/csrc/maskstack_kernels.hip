// maskstack_kernels.hip — the mask stack Y = X W as one contraction over the detector, for many masks.
//
// X is the (positions x pixels) cube, W the (pixels x K) stack {1, q_row, q_col, masks...}: K virtual
// images and the centre-of-mass sums in one pass (no reference function does this; the reference calls
// VirtualImaging.image once per mask, Sources/CuboVirtual/CPUReducers.swift:13-39).
//
// Two kernels:
//  - split-FP16 MFMA (uint16 cubes): x = (x & 0xF800) + (x & 0x07FF), both parts exact in FP16; each
//    mask column is scaled by a power of two and split into FP16 hi + lo (22 significant bits). Products
//    are exact in the FP32 MFMA accumulators, which are flushed into FP64 every 256 pixels so FP32
//    rounding cannot build up. The high data part is multiplied only for 64-pixel chunks where the wave
//    holds a count of 2048 or more, the lo weight part only for 32-mask tiles that have one.
//  - plain VALU (uint16 or Float32 cubes): one position per lane, FP32 FMAs over 64-pixel chunks staged
//    in LDS with their weights (read as LDS broadcasts), FP64 across chunks.
#include "maskstack_kernels.hpp"

namespace cubo {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short us2 __attribute__((ext_vector_type(2)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

constexpr int kStackBK = 64;           // pixels per chunk
constexpr int kFlushChunks = 4;        // FP32 -> FP64 every 4 chunks (256 pixels)

__device__ __forceinline__ uint32_t cvt_pair(uint32_t w) {
    // two uint16 below 2^16 whose values are exact in FP16 (callers mask them to <= 2047 or to multiples
    // of 2048) -> two FP16 in one register
    const us2 u = __builtin_bit_cast(us2, w);
    const h2 h = __builtin_convertvector(u, h2);
    return __builtin_bit_cast(uint32_t, h);
}

// One tile: WAVES waves x 32 positions starting at p0. Each lane loads its A operand straight from HBM
// (16 bytes: 8 pixels of one position per k-step), two 64-pixel chunks ahead of the matrix cores; the
// block shares one LDS copy of the weight chunk (double-buffered, one barrier per chunk). Every thread of
// the block calls it; it ends after a barrier, so the block may start another tile at once.
template <int NT>
struct StackLds {
    static constexpr int WCH = NT * 4 * 2 * 64;          // half8 entries of one weight chunk
    half8 w[2][WCH];
};

template <int NT, int WAVES>
__device__ __forceinline__ void stack_tile(const StackLaunch &a, int64_t p0, StackLds<NT> &sh) {
    constexpr int WCH = StackLds<NT>::WCH;
    auto &wlds = sh.w;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t prow = p0 + wave * 32 + (lane & 31);
    const bool live = prow < a.npos;
    const int nchunks = a.nchunks;
    const half8 *wsrc = reinterpret_cast<const half8 *>(a.wb);
    // Within a 64-pixel chunk, k-step s of lane half h holds pixels 32h + 8s + (0..7): each lane reads 64
    // contiguous bytes over the four k-steps, the two halves one full 128-byte line per position. Plain
    // (cached) loads keep the line in L2 between the four k-step loads.
    const uint16_t *xrow = a.cube + (live ? prow : 0) * a.ndet + 32 * (lane >> 5);

    auto load_x = [&](u32x4 (&x)[4], int chunk) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int64_t px = (int64_t)chunk * kStackBK + 32 * (lane >> 5) + 8 * s;
            x[s] = (live && px < a.ndet) ? *reinterpret_cast<const u32x4 *>(xrow + (int64_t)chunk * kStackBK + 8 * s)
                                         : u32x4{0u, 0u, 0u, 0u};
        }
    };
    // The block's weight chunk: each thread carries WPT 16-byte entries from HBM (or L2) to LDS. They are
    // loaded before the data prefetch of the same step, so waiting for them never waits for that prefetch.
    constexpr int WPT = (WCH + 64 * WAVES - 1) / (64 * WAVES);
    half8 wreg[WPT];
    auto fetch_w = [&](int chunk) {
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            const int e = tid + j * 64 * WAVES;
            if (e < WCH) wreg[j] = wsrc[(size_t)chunk * WCH + e];
        }
    };
    auto put_w = [&](int buf) {
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            const int e = tid + j * 64 * WAVES;
            if (e < WCH) wlds[buf][e] = wreg[j];
        }
    };

    float16v acc[NT];
    double accd[NT][16];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
#pragma unroll
        for (int r = 0; r < 16; ++r) { acc[n][r] = 0.f; accd[n][r] = 0.0; }
    }

    auto compute = [&](const u32x4 (&x)[4], int chunk) {
        uint32_t any = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s) any |= (x[s].x | x[s].y | x[s].z | x[s].w) & 0xf800f800u;
        const bool use_hi = __builtin_amdgcn_ballot_w64(any != 0) != 0;   // wave-uniform
        const half8 *wl = wlds[chunk & 1];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t w[4] = {x[s].x, x[s].y, x[s].z, x[s].w};
            u32x4 lo, hi;
            lo.x = cvt_pair(w[0] & 0x07ff07ffu); lo.y = cvt_pair(w[1] & 0x07ff07ffu);
            lo.z = cvt_pair(w[2] & 0x07ff07ffu); lo.w = cvt_pair(w[3] & 0x07ff07ffu);
            const half8 alo = __builtin_bit_cast(half8, lo);
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                if (n * 32 >= a.kn) continue;                  // a tile with no mask of this stack
                const half8 bh = wl[((n * 4 + s) * 2 + 0) * 64 + lane];
                acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bh, acc[n], 0, 0, 0);
                if (a.tile_lo[n]) {
                    const half8 bl = wl[((n * 4 + s) * 2 + 1) * 64 + lane];
                    acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bl, acc[n], 0, 0, 0);
                }
            }
            if (use_hi) {
                hi.x = cvt_pair(w[0] & 0xf800f800u); hi.y = cvt_pair(w[1] & 0xf800f800u);
                hi.z = cvt_pair(w[2] & 0xf800f800u); hi.w = cvt_pair(w[3] & 0xf800f800u);
                const half8 ahi = __builtin_bit_cast(half8, hi);
#pragma unroll
                for (int n = 0; n < NT; ++n) {
                    if (n * 32 >= a.kn) continue;
                    const half8 bh = wl[((n * 4 + s) * 2 + 0) * 64 + lane];
                    acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bh, acc[n], 0, 0, 0);
                    if (a.tile_lo[n]) {
                        const half8 bl = wl[((n * 4 + s) * 2 + 1) * 64 + lane];
                        acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bl, acc[n], 0, 0, 0);
                    }
                }
            }
        }
        if ((chunk + 1) % kFlushChunks == 0 || chunk + 1 == nchunks) {
#pragma unroll
            for (int n = 0; n < NT; ++n) {
#pragma unroll
                for (int r = 0; r < 16; ++r) { accd[n][r] += (double)acc[n][r]; acc[n][r] = 0.f; }
            }
        }
    };

    // Chunk c's weights are in wlds[c & 1] before the barrier that precedes compute(c); each step fetches
    // the next chunk's weights, then the data two chunks ahead, multiplies, and stores the weights into the
    // other buffer (read last by compute(c - 1), which every wave finished before the previous barrier).
    u32x4 X0[4], X1[4], X2[4];
    fetch_w(0);
    load_x(X0, 0);
    if (nchunks > 1) load_x(X1, 1);
    put_w(0);
    __syncthreads();
    auto step = [&](const u32x4 (&xc)[4], u32x4 (&xn)[4], int c) {
        if (c + 1 < nchunks) fetch_w(c + 1);
        if (c + 2 < nchunks) load_x(xn, c + 2);
        compute(xc, c);
        if (c + 1 < nchunks) put_w((c + 1) & 1);
        __syncthreads();
    };
    int c = 0;
    for (; c + 3 <= nchunks; c += 3) {
        step(X0, X2, c);
        step(X1, X0, c + 1);
        step(X2, X1, c + 2);
    }
    if (c < nchunks) step(X0, X2, c);
    if (c + 1 < nchunks) step(X1, X0, c + 1);

    // 32x32 accumulator layout: register r of lane l holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
    const int64_t pbase = p0 + wave * 32;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int col = n * 32 + (lane & 31);
        if (col >= a.kn) continue;
        const double inv = a.inv_scale[col];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t p = pbase + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (p < a.npos) a.y[p * a.k + a.k0 + col] = accd[n][r] * inv;
        }
    }
}

template <int NT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void mask_stack_mfma(StackLaunch a) {
    __shared__ StackLds<NT> sh;
    stack_tile<NT, WAVES>(a, (int64_t)blockIdx.x * 32 * WAVES, sh);
}

// Persistent batch: every block takes tiles from one work counter until all cubes are done. A tile is
// (cube, 32 * WAVES positions); first[c] is cube c's first tile number.
template <int NT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void batch_stack_mfma(const StackLaunch *cubes, const int64_t *first, int ncubes,
                                                              int64_t total, unsigned long long *counter) {
    __shared__ StackLds<NT> sh;
    __shared__ int64_t item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = (int64_t)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t item = item_s;
        __syncthreads();                                   // every thread has read item_s
        if (item >= total) break;                          // block-uniform: every block leaves
        int lo = 0, hi = ncubes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (first[mid] <= item) lo = mid; else hi = mid - 1;
        }
        stack_tile<NT, WAVES>(cubes[lo], (item - first[lo]) * 32 * WAVES, sh);
    }
}

__global__ void reset_counter(unsigned long long *c) { *c = 0ull; }

// ------------------------------------------------------------------------------------------------
// Plain VALU path: a lane per position.
// ------------------------------------------------------------------------------------------------
constexpr int kPlainBM = 256;
constexpr int kPlainBK = 64;                 // one 128-byte line of a uint16 row per chunk

template <typename T>
struct PlainRow { static constexpr int value = kPlainBK + (sizeof(T) == 2 ? 2 : 1); };   // pad: lanes hit distinct banks

template <typename T, int K>
__global__ __launch_bounds__(256) void mask_stack_plain(PlainStackLaunch a) {
    constexpr int ROW = PlainRow<T>::value;
    constexpr int PIECES = kPlainBK * sizeof(T) / 16;          // 16-byte pieces per position per chunk
    constexpr int XPT = kPlainBM * PIECES / 256;                // per thread
    constexpr int ELEMS = 16 / sizeof(T);
    constexpr int WPT = (kPlainBK * K / 4 + 255) / 256;         // float4 weights per thread
    __shared__ T tile[kPlainBM * ROW];                          // the chunk in the cube's element type
    __shared__ float4 wt[kPlainBK * K / 4];                     // its weights, read as broadcasts
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * kPlainBM;
    const T *cube = reinterpret_cast<const T *>(a.cube);
    const float4 *w4src = reinterpret_cast<const float4 *>(a.w);
    const bool vec = a.ndet % ELEMS == 0;                       // 16-byte loads stay aligned and in the row

    u32x4 xr[XPT];
    float4 wr[WPT];
    auto fetch = [&](int64_t c0) {
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            const int e = tid + 256 * j;
            const int64_t px = c0 + (e * 4) / K;
            if (e < kPlainBK * K / 4) wr[j] = px < a.ndet ? w4src[c0 * (K / 4) + e] : float4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int j = 0; j < XPT; ++j) {
            const int q = tid + 256 * j, r = q / PIECES, pc = q % PIECES;
            const int64_t p = p0 + r, px = c0 + pc * ELEMS;
            u32x4 v{0u, 0u, 0u, 0u};
            if (p < a.npos && px < a.ndet) {
                if (vec) {
                    v = *reinterpret_cast<const u32x4 *>(cube + p * a.ndet + px);
                } else {
                    T t[ELEMS];
#pragma unroll
                    for (int e = 0; e < ELEMS; ++e) t[e] = px + e < a.ndet ? cube[p * a.ndet + px + e] : (T)0;
                    v = __builtin_bit_cast(u32x4, t);
                }
            }
            xr[j] = v;
        }
    };
    auto put = [&]() {
#pragma unroll
        for (int j = 0; j < WPT; ++j) {
            const int e = tid + 256 * j;
            if (e < kPlainBK * K / 4) wt[e] = wr[j];
        }
#pragma unroll
        for (int j = 0; j < XPT; ++j) {
            const int q = tid + 256 * j, r = q / PIECES, pc = q % PIECES;
            const T *t = reinterpret_cast<const T *>(&xr[j]);
#pragma unroll
            for (int e = 0; e < ELEMS; ++e) tile[r * ROW + pc * ELEMS + e] = t[e];
        }
    };

    float acc[K];
    double accd[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { acc[k] = 0.f; accd[k] = 0.0; }
    const int nchunks = (int)((a.ndet + kPlainBK - 1) / kPlainBK);
    fetch(0);
    for (int ch = 0; ch < nchunks; ++ch) {
        put();                                                   // chunk ch: registers -> LDS
        __syncthreads();
        if (ch + 1 < nchunks) fetch((int64_t)(ch + 1) * kPlainBK);   // chunk ch + 1 in flight meanwhile
#pragma unroll 2
        for (int c = 0; c < kPlainBK; ++c) {
            const float x = (float)tile[tid * ROW + c];
            const float4 *w4 = &wt[c * K / 4];
#pragma unroll
            for (int k = 0; k < K / 4; ++k) {
                const float4 w = w4[k];
                acc[4 * k] = __builtin_fmaf(x, w.x, acc[4 * k]);
                acc[4 * k + 1] = __builtin_fmaf(x, w.y, acc[4 * k + 1]);
                acc[4 * k + 2] = __builtin_fmaf(x, w.z, acc[4 * k + 2]);
                acc[4 * k + 3] = __builtin_fmaf(x, w.w, acc[4 * k + 3]);
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) { accd[k] += (double)acc[k]; acc[k] = 0.f; }
        __syncthreads();                                         // every lane is done with the chunk in LDS
    }
    const int64_t p = p0 + tid;
    if (p < a.npos) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k < a.kn) a.y[p * a.k + a.k0 + k] = accd[k];
    }
}

hipError_t launch_mask_stack_mfma(const StackLaunch &a, int nt, hipStream_t s) {
    constexpr int W = 8;
    const dim3 grid((unsigned)((a.npos + 32 * W - 1) / (32 * W))), block(64 * W);
    switch (nt) {
        case 1: hipLaunchKernelGGL((mask_stack_mfma<1, W>), grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL((mask_stack_mfma<2, W>), grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int batch_blocks_per_cu(int waves) {
    int n = 0;
    hipError_t e = waves == 8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_stack_mfma<2, 8>, 512, 0)
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_stack_mfma<2, 4>, 256, 0);
    return e == hipSuccess && n > 0 ? n : 1;
}

hipError_t launch_batch_stack(const StackLaunch *cubes, const int64_t *first, int ncubes, int64_t total,
                              unsigned long long *counter, int waves, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, counter);
    if (waves == 8)
        hipLaunchKernelGGL((batch_stack_mfma<2, 8>), dim3(blocks), dim3(512), 0, s, cubes, first, ncubes, total, counter);
    else if (waves == 4)
        hipLaunchKernelGGL((batch_stack_mfma<2, 4>), dim3(blocks), dim3(256), 0, s, cubes, first, ncubes, total, counter);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_mask_stack_plain(const PlainStackLaunch &a, bool is_u16, int kpad, hipStream_t s) {
    const dim3 grid((unsigned)((a.npos + kPlainBM - 1) / kPlainBM)), block(256);
#define CUBO_PLAIN(K) \
    if (kpad == K) { \
        if (is_u16) hipLaunchKernelGGL((mask_stack_plain<uint16_t, K>), grid, block, 0, s, a); \
        else hipLaunchKernelGGL((mask_stack_plain<float, K>), grid, block, 0, s, a); \
        return hipGetLastError(); \
    }
    CUBO_PLAIN(8) CUBO_PLAIN(16) CUBO_PLAIN(32)
#undef CUBO_PLAIN
    return hipErrorInvalidValue;
}

}  // namespace cubo
