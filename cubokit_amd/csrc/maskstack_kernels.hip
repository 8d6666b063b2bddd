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
//    rounding cannot build up. The high data part is multiplied only for 64-pixel chunks whose block
//    holds a count of 2048 or more.
//  - plain VALU (uint16 or Float32 cubes): one position per lane, FP32 FMAs over 32-pixel chunks
//    staged in LDS, weights as wave-uniform scalar loads, FP64 across 32-pixel chunks.
#include "maskstack_kernels.hpp"

namespace cubo {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short us2 __attribute__((ext_vector_type(2)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

constexpr int kStackBM = 128;          // positions per block (32 per wave)
constexpr int kStackBK = 64;           // pixels per chunk
constexpr int kFlushChunks = 4;        // FP32 -> FP64 every 4 chunks (256 pixels)
constexpr int kRowHalves = kStackBK + 8;   // LDS row stride in halves (16-byte pad)

__device__ __forceinline__ uint32_t cvt_pair(uint32_t w) {
    // two uint16 below 2^16 whose values are exact in FP16 (callers mask them to <= 2047 or to multiples
    // of 2048) -> two FP16 in one register
    const us2 u = __builtin_bit_cast(us2, w);
    const h2 h = __builtin_convertvector(u, h2);
    return __builtin_bit_cast(uint32_t, h);
}

template <int NT>
__global__ __launch_bounds__(256) void mask_stack_mfma(StackLaunch a) {
    __shared__ _Float16 lo_t[kStackBM * kRowHalves];
    __shared__ _Float16 hi_t[kStackBM * kRowHalves];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t p0 = (int64_t)blockIdx.x * kStackBM;
    const int nchunks = a.nchunks;

    // Loader mapping: thread t covers positions t/8 + 32j (j = 0..3), 16 bytes at pixel (t % 8) * 8.
    const int lrow = tid >> 3, lcol = (tid & 7) * 8;
    auto load_x = [&](u32x4 (&x)[4], int chunk) {
        const int px = chunk * kStackBK + lcol;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t p = p0 + lrow + 32 * j;
            x[j] = (p < a.npos && px < a.ndet)
                       ? __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(a.cube + p * a.ndet + px))
                       : u32x4{0u, 0u, 0u, 0u};
        }
    };

    float16v acc[NT];
    double accd[NT][16];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
#pragma unroll
        for (int r = 0; r < 16; ++r) { acc[n][r] = 0.f; accd[n][r] = 0.0; }
    }

    u32x4 x[4];
    load_x(x, 0);
    for (int chunk = 0; chunk < nchunks; ++chunk) {
        // 1. convert the chunk held in registers and stage it in LDS
        uint32_t any_hi = 0;
        __syncthreads();                               // previous chunk's LDS reads are done
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t w[4] = {x[j].x, x[j].y, x[j].z, x[j].w};
            uint32_t lo[4], hi[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                lo[k] = cvt_pair(w[k] & 0x07ff07ffu);
                const uint32_t h = w[k] & 0xf800f800u;
                any_hi |= h;
                hi[k] = cvt_pair(h);
            }
            const int row = lrow + 32 * j;
            *reinterpret_cast<u32x4 *>(&lo_t[row * kRowHalves + lcol]) = u32x4{lo[0], lo[1], lo[2], lo[3]};
            *reinterpret_cast<u32x4 *>(&hi_t[row * kRowHalves + lcol]) = u32x4{hi[0], hi[1], hi[2], hi[3]};
        }
        const bool use_hi = __syncthreads_or(any_hi != 0);
        // 2. prefetch the next chunk while the matrix cores work on this one
        if (chunk + 1 < nchunks) load_x(x, chunk + 1);
        // 3. MFMA over the chunk: 4 k-steps of 16 pixels
        const half8 *wb = reinterpret_cast<const half8 *>(a.wb) + (size_t)chunk * NT * 4 * 2 * 64;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int arow = wave * 32 + (lane & 31);
            const int acol = s * 16 + (lane >> 5) * 8;
            const half8 alo = *reinterpret_cast<const half8 *>(&lo_t[arow * kRowHalves + acol]);
            const half8 ahi = *reinterpret_cast<const half8 *>(&hi_t[arow * kRowHalves + acol]);
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const half8 bh = wb[((n * 4 + s) * 2 + 0) * 64 + lane];
                const half8 bl = wb[((n * 4 + s) * 2 + 1) * 64 + lane];
                acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bh, acc[n], 0, 0, 0);
                acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bl, acc[n], 0, 0, 0);
                if (use_hi) {
                    acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bh, acc[n], 0, 0, 0);
                    acc[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bl, acc[n], 0, 0, 0);
                }
            }
        }
        // 4. flush FP32 into FP64
        if ((chunk + 1) % kFlushChunks == 0 || chunk + 1 == nchunks) {
#pragma unroll
            for (int n = 0; n < NT; ++n) {
#pragma unroll
                for (int r = 0; r < 16; ++r) { accd[n][r] += (double)acc[n][r]; acc[n][r] = 0.f; }
            }
        }
    }
    // 32x32 accumulator layout: register r of lane l holds row 8*(r/4) + 4*(l/32) + r%4, column l%32.
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int col = n * 32 + (lane & 31);
        if (col >= a.k) continue;
        const double inv = a.inv_scale[col];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = wave * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int64_t p = p0 + row;
            if (p < a.npos) a.y[p * a.k + col] = accd[n][r] * inv;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Plain VALU path: a lane per position.
// ------------------------------------------------------------------------------------------------
constexpr int kPlainBM = 256;
constexpr int kPlainBK = 32;
constexpr int kPlainRow = kPlainBK + 1;     // floats per LDS row (odd word stride: lanes hit distinct banks)

template <typename T, int K>
__global__ __launch_bounds__(256) void mask_stack_plain(PlainStackLaunch a) {
    __shared__ float tile[kPlainBM * kPlainRow];
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * kPlainBM;
    const T *cube = reinterpret_cast<const T *>(a.cube);
    float acc[K];
    double accd[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { acc[k] = 0.f; accd[k] = 0.0; }
    for (int64_t c0 = 0; c0 < a.ndet; c0 += kPlainBK) {
        __syncthreads();
        // stage 256 positions x 32 pixels: each pass of the block covers 8 positions x 32 pixels
        for (int e = tid; e < kPlainBM * kPlainBK; e += 256) {
            const int r = e / kPlainBK, c = e % kPlainBK;
            const int64_t p = p0 + r, px = c0 + c;
            tile[r * kPlainRow + c] = (p < a.npos && px < a.ndet) ? (float)cube[p * a.ndet + px] : 0.f;
        }
        __syncthreads();
        const int n = (int)min((int64_t)kPlainBK, a.ndet - c0);
        for (int c = 0; c < n; ++c) {
            const float x = tile[tid * kPlainRow + c];
            const float *w = a.w + (c0 + c) * K;          // wave-uniform: scalar loads
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] = __builtin_fmaf(x, w[k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) { accd[k] += (double)acc[k]; acc[k] = 0.f; }
    }
    const int64_t p = p0 + tid;
    if (p < a.npos) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k < a.kn) a.y[p * a.k + a.k0 + k] = accd[k];
    }
}

hipError_t launch_mask_stack_mfma(const StackLaunch &a, int nt, hipStream_t s) {
    const dim3 grid((unsigned)((a.npos + kStackBM - 1) / kStackBM)), block(256);
    switch (nt) {
        case 1: hipLaunchKernelGGL(mask_stack_mfma<1>, grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL(mask_stack_mfma<2>, grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL(mask_stack_mfma<4>, grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
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
    CUBO_PLAIN(8) CUBO_PLAIN(16) CUBO_PLAIN(32) CUBO_PLAIN(64)
#undef CUBO_PLAIN
    return hipErrorInvalidValue;
}

}  // namespace cubo
