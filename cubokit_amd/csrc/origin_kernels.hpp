// origin_kernels.hpp — launch interface of the per-position origin kernels (origin_kernels.hip).
//
// Two estimators, each over a tile of scan positions of the resident cube:
//   - py4DSTEM's get_origin (Sources/CuboVirtual/OriginMeasurement.swift:52-95): scipy's gaussian_filter
//     (sigma = r, truncate 4, mode nearest, axis 0 then axis 1) in Double, the first argmax (the first NaN if any),
//     then the centre of mass of the pattern within r * rscale of it;
//   - the iterated-window estimator (IteratedWindowOrigin.swift:136-180, Metal cubo_measure_origin): Float32, the
//     brightest bin x bin block, then up to `passes` centres of mass of a window re-centred on its estimate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubo {

struct Py4Launch {
    const void *cube;           // the tile's first pattern
    bool u16;
    int64_t npos;               // positions in the tile
    int32_t qy, qx;
    const double *weights;      // [2R + 1] normalised Gaussian weights
    int32_t R;
    double limit;               // r * rscale: the centre-of-mass disk (strict <)
    const uint8_t *take;        // [npos] or NULL: positions to measure (NaN elsewhere)
    int32_t fast_center;        // the filtered argmax itself
    double *scratch;            // [gridDim.x][qy * qx]: the axis-0 pass of each block
    double *out_row, *out_col;  // [npos]
    double *filtered;           // optional: the whole filtered pattern of the tile's first position, or NULL
    // optional: measure only positions list[0 .. *nlist) (the matrix-core kernel's undecided patterns)
    const int32_t *list;
    const int32_t *nlist;
    // the matrix-core kernel (py4_origin_mfma): the filter matrix split into FP16 hi/lo, [4][n][n] (G hi, G lo, and
    // G with permuted columns hi, lo), the scale of its second-pass input, the relative margin of the candidates,
    // and where it lists the patterns it leaves to the exact kernel
    const _Float16 *gmat;
    float pass2_scale;
    float candidate_margin;
    float slack;                // the absolute margin of the candidates (second-pass inputs below FP16's normals)
    int32_t *undecided;
    int32_t *n_undecided;
};

struct IWLaunch {
    const void *cube;
    bool u16;
    int64_t npos;
    int32_t qy, qx;
    float radius, radius_scale, margin;
    int32_t passes;
    int32_t binned;
    double *out_row, *out_col;
};

// Blocks of the get_origin kernel (each loops over positions with its own scratch pattern).
int py4_blocks(int cus, int64_t npos);
hipError_t launch_py4_origin(const Py4Launch &a, int blocks, hipStream_t s);
// The matrix-core get_origin (py4_origin_mfma): uint16 patterns of kPyMfmaSide x kPyMfmaSide pixels, no filtered
// pattern asked for. It measures every taken position but those it lists in undecided (n_undecided, cleared by the
// caller), which launch_py4_origin with list = undecided then measures exactly.
constexpr int kPyMfmaSide = 128;
bool py4_mfma_supported(const Py4Launch &a);
hipError_t launch_py4_origin_mfma(const Py4Launch &a, int blocks, hipStream_t s);
hipError_t launch_iterated_window(const IWLaunch &a, hipStream_t s);

}  // namespace cubo
