// reduce_kernels.hpp — launch interface of the fused streaming reduction kernels (reduce_kernels.hip).
//
// One pass over a resident cube yields, per scan position, the dot products of the pattern with a stack
// of weight rows (virtual-detector masks and the centre-of-mass weights w, w*i, w*j), and, per detector
// pixel, the partial sums and maxima over the positions of each position group. Two kernel families:
//   - exact integer (uint16 cube, non-negative integral weights): per-lane uint32 partials from packed
//     16-bit dot products, uint64 totals, so every output is the exact integer;
//   - FP64 (uint16 or Float32 cube, any finite weights): per-lane double accumulation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubo {

// Lanes per wave and waves per block of the streaming kernels.
constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
// Most slices per pattern for the integer kernel's FULL layout (one wave per slice in one block).
constexpr int kFullSlices = 8;

// Row totals of a launch: rows[v][pos0 + p] for v < nv and, when com_base >= 0, the centre of mass
// com[0][pos0 + p] = rows[com_base + 1] / rows[com_base] and com[1][...] = rows[com_base + 2] / rows[com_base]
// (correctly rounded Double divisions; NaN where the intensity is 0). The pointers may be device
// pointers of mapped host memory, so the totals reach the host while the kernel still streams.
struct RowsOut {
    double *rows;
    double *com;
    int64_t pos0, total_pos;
    int32_t nv;
    int32_t com_base;
};

struct IntLaunch {
    const uint16_t *cube;   // first position of this launch
    int64_t npos;           // positions in this launch
    int32_t ndet;           // pixels per pattern (multiple of 8)
    int32_t qx;             // detector columns (multiple of 8 for the arithmetic CoM rows)
    int32_t nslices;        // pixel slices per pattern
    int32_t ngroups;        // position groups
    int32_t ppg;            // positions per group
    int32_t ppw;            // positions per wave
    const uint32_t *wpk;    // [nw][ndet/2] weights as packed uint16 pairs
    uint64_t *rowpart;      // [nslices][npos][nv]
    uint32_t *colsum;       // [ngroups][ndet]
    uint32_t *colmax;       // [ngroups][ndet / 2]: packed uint16 pairs, the lower pixel in the low half
    RowsOut out;            // row totals (written by the kernel itself in the FULL layout)
};

struct F64Launch {
    const void *cube;       // first position of this launch (uint16_t or float)
    int64_t npos;
    int32_t ndet;
    int32_t nslices;
    int32_t ngroups;
    int32_t ppg;
    int32_t ppw;
    const double *w;        // [nw][ndet] weight rows in Double
    double *rowpart;        // [nslices][npos][nw]
    double *colsum;         // [ngroups][ndet]
    float *colmax;          // [ngroups][ndet]
    int32_t nmask;          // mask rows (the first rows of w)
    int32_t com;            // 1: the three rows after the masks are the centre of mass of unit weight
    int32_t qx;             // detector columns
    int32_t vec_ok;         // 0: always the one-pixel-per-lane kernel (a test knob)
    RowsOut out;            // the vectorised kernel with one octet (nslices = 1): row totals written by the kernel
};

// Pixels per lane of the integer kernel (chunks of 8) and of the FP64 kernel (one pixel per chunk).
int int_slice_pixels(int chunks);
constexpr int kF64Chunks = 16;
constexpr int kF64Slice = kWave * kF64Chunks;

// Returns hipSuccess or the launch error. nw / com_ones select the instantiation; the caller checks
// that (nw, chunks) is one of the compiled combinations with supported_int().
bool supported_int(int nw, int chunks, bool com_ones);
// full: one block per position group with one wave per slice (nslices <= kFullSlices), row totals written
// by the kernel; otherwise blocks of 4 waves per (slice, group) and launch_finalize_rows_u64 afterwards.
hipError_t launch_int(const IntLaunch &a, int nw, int chunks, bool com_ones, bool full, hipStream_t s);
// The FULL instantiation keeps everything in registers (its 512-thread blocks cap a lane at 256 VGPRs).
bool full_without_spills(int nw, int chunks, bool com_ones);
// Resident blocks per compute unit of an instantiation (0 if it is not compiled).
int blocks_per_cu_int(int nw, int chunks, bool com_ones, bool full, int nslices);
// Most masks of the vectorised FP64 kernel (weights in registers); more masks, or weighted centre-of-mass rows, run on
// the one-pixel-per-lane kernel.
constexpr int kVecMasks = 4;
constexpr int kVecSlice = kWave * 8;      // a wave's slice: 8 pixels per lane
constexpr int kVecWaves = 8;              // waves per block: one octet of slices
constexpr int kVecOctet = kVecWaves * kVecSlice;
// Pixels per slice of the kernel a launch would run (for the vectorised kernel, per octet: nslices counts octets
// and rowpart holds one partial per octet).
int f64_slice_pixels(const F64Launch &a, bool is_u16);
// Resident blocks per compute unit of the kernel a launch would run (its cube pointer, ndet, nmask, com and qx set).
int blocks_per_cu_f64(const F64Launch &a, bool is_u16, int nw);
bool supported_f64(int nw);
// 1: fused_f64 (one pixel per lane per chunk), 2: fused_vec_f64 (16-byte loads, weights in registers).
int f64_kernel(const F64Launch &a, bool is_u16);
hipError_t launch_f64(const F64Launch &a, bool is_u16, int nw, hipStream_t s);

// Row totals over slices, one thread per position, into `out`.
hipError_t launch_finalize_rows_u64(const uint64_t *rowpart, int nslices, int64_t npos, const RowsOut &out,
                                    hipStream_t s);
hipError_t launch_finalize_rows_f64(const double *rowpart, int nslices, int64_t npos, const RowsOut &out,
                                    hipStream_t s);
// Column partials of one launch, in group order, into the running sums and maxima. first: the sums start
// from this launch (nothing is read from sum/max); finish: the launch is the last one, so mean = sum / count
// and max go to mean_out / max_out (possibly mapped host memory) instead of sum / max.
hipError_t launch_accumulate_cols_int(const uint32_t *colsum, const uint32_t *colmax, int ngroups, int ndet,
                                      uint64_t *sum, uint32_t *max, bool first, bool finish, double count,
                                      double *mean_out, double *max_out, hipStream_t s);
hipError_t launch_accumulate_cols_f64(const double *colsum, const float *colmax, int ngroups, int ndet,
                                      double *sum, float *max, bool first, bool finish, double count,
                                      double *mean_out, double *max_out, hipStream_t s);
// dst = src for n bytes (16-byte aligned buffers): 16-byte loads and stores, grid-stride.
hipError_t launch_copy_bytes(const void *src, void *dst, uint64_t n, hipStream_t s);

}  // namespace cubo
