// maskstack_kernels.hpp — launch interface of the mask-stack contraction (maskstack_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubo {

struct StackLaunch {
    const uint16_t *cube;     // first position
    int64_t npos;
    int64_t ndet;             // multiple of 8
    int32_t nchunks;          // ceil(ndet / 64)
    int32_t k;                // masks in the whole stack (the row stride of y)
    int32_t k0, kn;           // this pass: masks [k0, k0 + kn), kn <= 32 * ntiles
    const void *wb;           // FP16 hi/lo weights in MFMA B-operand order [chunk][ntile][kstep][hi,lo][lane][8]
    const double *inv_scale;  // [kn] 1 / (power-of-two column scale) of this pass's masks
    int32_t tile_lo[2];       // per 32-mask tile: 1 if any weight needs the FP16 lo part
    double *y;                // [npos][k]
};

struct PlainStackLaunch {
    const void *cube;         // uint16_t or float, first position
    int64_t npos;
    int64_t ndet;
    int32_t k;                // masks in the whole stack (the row stride of y)
    int32_t k0, kn;           // this pass: masks [k0, k0 + kn)
    const float *w;           // [ndet rounded up to 64][kpad] FP32 weights of this pass, pixel-major
    double *y;                // [npos][k]
};

hipError_t launch_mask_stack_mfma(const StackLaunch &a, int ntiles, hipStream_t s);
// One work item of a persistent batch: tile `tile` (32 * waves positions) of cube `cube`.
struct BatchItem {
    int32_t cube, tile;
};
// Persistent batch of uint16 cubes (at most 64 masks each): resets the work counter with a kernel, then
// launches `blocks` blocks of `waves` waves that take items[0, total) in order until none is left.
int batch_blocks_per_cu(int waves);
hipError_t launch_batch_stack(const StackLaunch *cubes, const BatchItem *items, int64_t total,
                              unsigned long long *counter, int waves, int blocks, hipStream_t s);
hipError_t launch_mask_stack_plain(const PlainStackLaunch &a, bool is_u16, int kpad, hipStream_t s);

}  // namespace cubo
