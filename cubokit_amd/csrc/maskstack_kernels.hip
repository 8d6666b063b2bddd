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
#include "device_sync.hpp"
#include "maskstack_kernels.hpp"

#include <type_traits>

namespace cubo {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
// Global-memory views of pointers that were read from memory (the batch descriptors): without the
// address space the compiler emits flat loads, whose completion it can only track together with LDS
// traffic, and every wait then drains both.
#define CUBO_GLOBAL __attribute__((address_space(1)))
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

template <int NT, int WAVES, bool DEEP>
__device__ __forceinline__ void stack_tile(const StackLaunch &a, int64_t p0, StackLds<NT> &sh) {
    constexpr int WCH = StackLds<NT>::WCH;
    auto &wlds = sh.w;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t prow = p0 + wave * 32 + (lane & 31);
    const bool live = prow < a.npos;
    const int nchunks = a.nchunks;
    const CUBO_GLOBAL half8 *wsrc = (const CUBO_GLOBAL half8 *)a.wb;
    // Within a 64-pixel chunk, k-step s of lane half h holds pixels 32h + 8s + (0..7): each lane reads 64
    // contiguous bytes over the four k-steps, the two halves one full 128-byte line per position. Plain
    // (cached) loads keep the line in L2 between the four k-step loads.
    // Every load is unconditional (loads under a lane mask make the compiler wait for all loads in flight,
    // which drains the prefetch): a position past the cube reads the first row and a pixel past the pattern
    // reads its row's start. Neither needs masking: the weights of pixels past the pattern are zero and
    // the split parts of any uint16 are finite in FP16, so they add exact zeros, and the rows of positions
    // past the cube are never written.
    const CUBO_GLOBAL uint16_t *xrow = (const CUBO_GLOBAL uint16_t *)a.cube + (live ? prow : 0) * a.ndet;

    auto load_x = [&](u32x4 (&x)[4], int chunk) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int64_t px = (int64_t)chunk * kStackBK + 32 * (lane >> 5) + 8 * s;
            x[s] = *(const CUBO_GLOBAL u32x4 *)(xrow + (px < a.ndet ? px : 0));
        }
    };
    // The block's weight chunk: each thread carries WPT 16-byte entries from HBM (or L2) to LDS. They are
    // loaded before the data prefetch of the same step, so waiting for them never waits for that prefetch.
    // DEEP keeps two weight chunks and three data chunks ahead in registers (for blocks with the registers
    // to spare): a step then waits on loads issued two steps earlier.
    constexpr int WPT = (WCH + 64 * WAVES - 1) / (64 * WAVES);
    constexpr int NW = DEEP ? 2 : 1;       // weight chunks in registers
    constexpr int NX = DEEP ? 4 : 3;       // data chunks in registers (the current one and NX - 1 ahead)
    half8 wreg[NW][WPT];
    static_assert(WCH % (64 * WAVES) == 0, "every thread carries the same number of weight entries");
    auto fetch_w = [&](int set, int chunk) {
#pragma unroll
        for (int j = 0; j < WPT; ++j) wreg[set][j] = wsrc[(size_t)chunk * WCH + tid + j * 64 * WAVES];
    };
    auto put_w = [&](int set, int buf) {
#pragma unroll
        for (int j = 0; j < WPT; ++j) wlds[buf][tid + j * 64 * WAVES] = wreg[set][j];
    };

    float16v acc[NT];
    double accd[NT][16];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
#pragma unroll
        for (int r = 0; r < 16; ++r) { acc[n][r] = 0.f; accd[n][r] = 0.0; }
    }

    // Both tiles' weight fragments of a k-step are read from LDS before its MFMAs, so the reads overlap;
    // a tile with no mask of this stack multiplies zero weights (exact zeros), and the lo parts are
    // multiplied when any tile has one (any_lo is uniform over the launch).
    const bool any_lo = (a.tile_lo[0] | (NT > 1 ? a.tile_lo[1] : 0)) != 0;
    auto compute = [&](const u32x4 (&x)[4], int chunk) {
        uint32_t any = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s) any |= (x[s].x | x[s].y | x[s].z | x[s].w) & 0xf800f800u;
        const bool use_hi = __builtin_amdgcn_ballot_w64(any != 0) != 0;   // wave-uniform
        const half8 *wl = wlds[chunk & 1];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            half8 bh[NT], bl[NT];
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                bh[n] = wl[((n * 4 + s) * 2 + 0) * 64 + lane];
                bl[n] = wl[((n * 4 + s) * 2 + 1) * 64 + lane];
            }
            const uint32_t w[4] = {x[s].x, x[s].y, x[s].z, x[s].w};
            u32x4 lo, hi;
            lo.x = cvt_pair(w[0] & 0x07ff07ffu); lo.y = cvt_pair(w[1] & 0x07ff07ffu);
            lo.z = cvt_pair(w[2] & 0x07ff07ffu); lo.w = cvt_pair(w[3] & 0x07ff07ffu);
            const half8 alo = __builtin_bit_cast(half8, lo);
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bh[n], acc[n], 0, 0, 0);
            if (any_lo) {
#pragma unroll
                for (int n = 0; n < NT; ++n) acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bl[n], acc[n], 0, 0, 0);
            }
            if (use_hi) {
                hi.x = cvt_pair(w[0] & 0xf800f800u); hi.y = cvt_pair(w[1] & 0xf800f800u);
                hi.z = cvt_pair(w[2] & 0xf800f800u); hi.w = cvt_pair(w[3] & 0xf800f800u);
                const half8 ahi = __builtin_bit_cast(half8, hi);
#pragma unroll
                for (int n = 0; n < NT; ++n) acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bh[n], acc[n], 0, 0, 0);
                if (any_lo) {
#pragma unroll
                    for (int n = 0; n < NT; ++n)
                        acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bl[n], acc[n], 0, 0, 0);
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

    // Chunk c's weights are in wlds[c & 1] before the barrier that precedes compute(c); step c fetches the
    // weights NW chunks ahead into the register set chunk c's weights left, then the data NX - 1 chunks
    // ahead, multiplies, and stores chunk c + 1's weights into the other LDS buffer (read last by
    // compute(c - 1), which every wave finished before the previous barrier).
    // The prologue issues the loads of the steps before chunk 0 in the order the steps issue them (step v:
    // the weights of chunk v + NW, then the data of chunk v + NX - 1), so every wait in the loop counts the
    // same loads on its first pass as on later ones.
    u32x4 X[NX][4];
#pragma unroll
    for (int v = -(NX - 1); v < 0; ++v) {
        if (v + NW >= 0) fetch_w(v + NW, min(v + NW, nchunks - 1));
        __builtin_amdgcn_sched_barrier(0);
        load_x(X[v + NX - 1], min(v + NX - 1, nchunks - 1));
    }
    put_w(0, 0);
    lds_barrier();
    // The loads are unconditional (past the last chunk they repeat it, and the extra weight copies land in
    // registers or in the buffer no wave reads any more), and each weight fetch is issued before the data
    // loads of its step: the wait for the weights before put_w then leaves the data prefetch in flight.
    // Step P of an unrolled group handles a chunk c with c = P modulo NW and NX, so the register sets are
    // chosen at compile time.
    auto step = [&](auto phase, int c) {
        constexpr int P = decltype(phase)::value;
        fetch_w(P % NW, min(c + NW, nchunks - 1));
        __builtin_amdgcn_sched_barrier(0);
        load_x(X[(P + NX - 1) % NX], min(c + NX - 1, nchunks - 1));
        compute(X[P % NX], c);
        put_w((P + 1) % NW, (c + 1) & 1);
        lds_barrier();
    };
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    using P2 = std::integral_constant<int, 2>;
    using P3 = std::integral_constant<int, 3>;
    static_assert(NX % NW == 0 && (NX == 3 || NX == 4), "one unrolled group of NX steps");
    int c = 0;
    for (; c + NX <= nchunks; c += NX) {
        step(P0{}, c);
        step(P1{}, c + 1);
        step(P2{}, c + 2);
        if constexpr (NX == 4) step(P3{}, c + 3);
    }
    if (c < nchunks) step(P0{}, c);
    if (c + 1 < nchunks) step(P1{}, c + 1);
    if constexpr (NX == 4) if (c + 2 < nchunks) step(P2{}, c + 2);

    // 32x32 accumulator layout: register r of lane l holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
    const int64_t pbase = p0 + wave * 32;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int col = n * 32 + (lane & 31);
        if (col >= a.kn) continue;
        const double inv = ((const CUBO_GLOBAL double *)a.inv_scale)[col];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t p = pbase + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (p < a.npos) ((CUBO_GLOBAL double *)a.y)[p * a.k + a.k0 + col] = accd[n][r] * inv;
        }
    }
}

template <int NT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void mask_stack_mfma(StackLaunch a) {
    __shared__ StackLds<NT> sh;
    stack_tile<NT, WAVES, WAVES == 8>(a, (int64_t)blockIdx.x * 32 * WAVES, sh);
}

// Persistent batch: every block takes items from one work counter until all are done. Item i is tile
// items[i].tile (32 * WAVES positions) of cube items[i].cube; the host chooses the order.
template <int NT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void batch_stack_mfma(const StackLaunch *__restrict__ cubes,
                                                              const BatchItem *__restrict__ items,
                                                              int64_t total, unsigned long long *counter) {
    __shared__ StackLds<NT> sh;
    __shared__ int item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = (int)atomicAdd(counter, 1ull);     // total < 2^31 (checked on the host)
        __syncthreads();
        // uniform from here: the item, the cube it falls in and that cube's descriptor live in scalar registers
        const int item = __builtin_amdgcn_readfirstlane(item_s);
        __syncthreads();                                   // every thread has read item_s
        if (item >= total) break;                          // block-uniform: every block leaves
        const BatchItem it = items[item];
        const StackLaunch a = cubes[it.cube];
        const int64_t p0 = (int64_t)it.tile * 32 * WAVES;
        // a cube of at most 32 masks has its weights in the one-tile layout: half the matrix-core work
        if (NT == 2 && a.kn <= 32)
            stack_tile<1, WAVES, WAVES == 8>(a, p0, *reinterpret_cast<StackLds<1> *>(&sh));
        else
            stack_tile<NT, WAVES, WAVES == 8>(a, p0, sh);
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
        lds_barrier();
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
        lds_barrier();                                           // every lane is done with the chunk in LDS
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

hipError_t launch_batch_stack(const StackLaunch *cubes, const BatchItem *items, int64_t total,
                              unsigned long long *counter, int waves, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, counter);
    if (waves == 8)
        hipLaunchKernelGGL((batch_stack_mfma<2, 8>), dim3(blocks), dim3(512), 0, s, cubes, items, total, counter);
    else if (waves == 4)
        hipLaunchKernelGGL((batch_stack_mfma<2, 4>), dim3(blocks), dim3(256), 0, s, cubes, items, total, counter);
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
