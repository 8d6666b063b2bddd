// shift_kernels.hpp — launch interface of the shifted mask stack and its adjoint (shift_kernels.hip).
//
// Per position p the masks are moved by the bilinear shift of MaskShift.bilinear
// (Sources/CuboVirtual/ShiftedImaging.swift:43-73): with d = origin_p - mean(origin), integer parts
// (ir, ic) = floor(d) wrapped into the detector and fractions (wr, wc), the shifted mask is the blend of the
// four periodic rolls of the mask by the corner shifts s0 = (ir, ic), s1 = (ir + 1, ic), s2 = (ir, ic + 1),
// s3 = (ir + 1, ic + 1) with weights (1-wr)(1-wc), wr(1-wc), (1-wr)wc, wr wc. So
//     Y[p, k] = sum_c w_c(p) C_c[p, k],   C_c[p, k] = sum_u x_p[u] m_k[(u - s_c) mod (qy, qx)],
// and every C_c is an ordinary mask-stack contraction with a rolled mask. Positions are sorted by (ir, ic)
// into tiles of one class, so a tile multiplies the exact uint16 data by four rolled copies of the split-FP16
// weights, on the matrix cores.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubo {

// A tile: up to 32 * WAVES positions of one shift class, perm[start .. start + count), over the chunks c0 .. c1 - 1.
// A whole tile (c0 = 0, c1 = nchunks, part = -1) writes its outputs; a split tile (one chunk range of a tile split
// across several blocks) writes partial outputs at part (256 positions each), added by launch_shift_split_sum.
struct ShiftTile {
    int32_t start, count;
    int32_t slot[4];          // rolled weight tables of the four corners
    int32_t c0, c1, part;
};

struct ShiftLaunch {
    const uint16_t *cube;
    int64_t npos, ndet;       // ndet: multiple of 8
    int32_t nchunks;          // ceil(ndet / 64)
    int32_t k;                // masks (at most 32)
    int32_t any_lo;           // some weight needs its FP16 lo part
    int32_t row_chunks;       // qx / 64 when qx is a multiple of 64 (a roll by one row = row_chunks chunks), else 0
    const void *wroll;        // [slot][chunk][kstep][hi, lo][lane][8] FP16, the split weights rolled by the slot's shift
    int64_t slot_stride;      // 16-byte entries per slot
    const double *inv_scale;  // [k] 1 / (power-of-two column scale)
    const ShiftTile *tiles;
    const int32_t *perm;      // positions sorted by shift class
    const double *wr, *wc;    // [npos] bilinear fractions
    double *y;                // forward: [npos][k]
    const double *dy;         // backward: [npos][k] dL/dY
    double *gr, *gc;          // backward: [npos] dL/d(shift row), dL/d(shift column), before the mean removal
    double *part;             // split tiles' partial outputs: [part][256][k] (Y), [part][256][2] (origin adjoint) or
                              // [part][256][3][k] (Y and the Jacobian rows)
    double *jac;              // kShiftYJac: [2][npos][k] dY/dwr and dY/dwc
};

// What one pass of the shifted stack computes.
enum ShiftPass : int {
    kShiftY = 0,              // Y
    kShiftOrigin = 1,         // gr, gc from dL/dY (the origin adjoint)
    kShiftYJac = 2,           // Y, and the Jacobian rows dY/dwr, dY/dwc kept for an origin adjoint without a pass
};

// Tiles of 32 * kShiftWaves positions.
constexpr int kShiftWaves = 8;
// rounded: MaskShift.rounded, one roll per class (tiles use slot[0] only); kShiftY only.
hipError_t launch_shift_stack(const ShiftLaunch &a, int ntiles, int pass, bool rounded, hipStream_t s);
// The split tiles' partial outputs added in split order into the pass's outputs: ngroups groups of nsplit consecutive
// split tiles from first_split (each group one tile's position range, its parts consecutive).
hipError_t launch_shift_split_sum(const ShiftLaunch &a, int first_split, int ngroups, int nsplit, int pass,
                                  hipStream_t s);
// The origin adjoint from kept Jacobian rows jac [2][npos][k]: gr[p] = sum_k dy[p][k] jac[0][p][k], gc likewise.
hipError_t launch_shift_origin_from_jac(const double *jac, const double *dy, int64_t npos, int k, double *gr, double *gc,
                                        hipStream_t s);

// Any cube (uint16 or Float32, any detector shape): each position's masks moved by the bilinear or the rounded
// shift and dotted with the pattern in Double, as MaskShift's CPU loop does (ShiftedImaging.swift:55-71): the blend
// a*w00 + b*w10 + c*w01 + d*w11 of the four rolls, Double(x) * m, TwoSum per lane. A wave per position, the masks
// one after another (the pattern is re-read from the cache). The forward pass only.
struct ShiftValu {
    const void *cube;
    bool u16;
    int64_t npos;
    int32_t qy, qx, k;
    const float *masks;       // [k][qy * qx]
    const int32_t *shift;     // [npos][2] (ir, ic) in [0, qy) x [0, qx)
    const double *wr, *wc;    // [npos] fractions (0 for the rounded shift)
    double *y;                // [npos][k]
};
hipError_t launch_shift_valu(const ShiftValu &a, bool rounded, hipStream_t s);

// out[slot] = base rolled by shift[slot] = (tr, tc): the entry of pixel u holds the base weight of pixel
// ((ui - tr) mod qy, (uj - tc) mod qx). base is split_weights' one-tile layout.
hipError_t launch_roll_weights(const void *base, void *out, const int32_t *shifts, int nslots, int nchunks, int64_t ndet,
                               int qy, int qx, hipStream_t s);

// dL/dm[k][v] = sum_p dY[p, k] xs_p[v], xs_p[v] = sum_c w_c(p) x_p[(v + s_c) mod (qy, qx)] (the pattern moved back
// by the shift, bilinearly): one thread per pixel, position groups in parallel, FP32 within 16 positions, FP64
// across them; partial[g][k][v] per group g.
struct ShiftGradW {
    const uint16_t *cube;
    int64_t npos;
    int32_t qy, qx;
    int32_t kp;               // masks padded to the instantiation (4, 8, 16, 24 or 32)
    const int32_t *shift;     // [npos][2] (ir, ic)
    const float *wts;         // [npos][4] corner weights
    const float *dy;          // [npos][kp] dL/dY in FP32
    double *partial;          // [ngroups][kp][ndet]
    int32_t ngroups;
    int64_t ppg;              // positions per group
};
hipError_t launch_shift_grad_w(const ShiftGradW &a, hipStream_t s);

// The same adjoint on the matrix cores, for detectors whose rows hold a multiple of 32 pixels, at most 256: block
// (band, g) owns a band of output rows (shift_grad_rows_band) and position group g, one wave per output row and
// 32 output pixels. dL/dm[k][v] =
// sum_p sum_c (dL/dY[p, k] w_c(p)) x_p[v + s_c(p)] over the four corners c of the bilinear shift: for each corner
// the A operand is the pattern value at that neighbour of pixel v (32 pixels x 16 positions, uint16 split exactly
// into FP16 lo and hi parts) and the B operand dL/dY times the corner weight (FP16 hi + lo, columns scaled by
// powers of two). Per batch of 16 positions the band + 1 input rows they read (from (a0 + ir) mod qy on) and the
// batch's B operands are staged in LDS once for the band (double-buffered); two MFMAs per corner and k-step (four
// when the data have high parts), flushed to FP64 every 32 positions.
struct ShiftGradRows {
    const uint16_t *cube;
    int64_t npos;
    int32_t qy, qx;
    const int32_t *shift;     // [npos][2] (ir, ic)
    const void *dyw;          // [ceil(npos / 16)][corner][hi, lo][lane][8] FP16, zero past the positions and masks
    const double *dy_inv;     // [32] 1 / column scale (0 past the masks)
    double *partial;          // [ngroups][32][ndet]
    int32_t ngroups;
    int64_t ppg;              // positions per group, a multiple of 16
};
bool shift_grad_rows_supported(int qx);
// Output rows per block of shift_grad_rows for a row width (the grid is ceil(qy / band) x ngroups).
int shift_grad_rows_band(int qx);
// The B operands of shift_grad_rows on the device: column kk of dy [npos][k] times the position's corner weight
// ((1-wr)(1-wc), wr(1-wc), (1-wr)wc, wr wc for the corners (0,0), (1,0), (0,1), (1,1)) times scale[kk] (a power of
// two), rounded to Float then split into FP16 hi and lo, at [k-step][corner][hi, lo][lane][8] (lane l: mask l % 32,
// positions 16 s + 8 (l / 32) + e), zero past the positions and the masks.
// order (optional): the position of each of the nslots slots, -1 for a padding slot (zero B operands); without it
// slot i is position i and nslots = npos.
hipError_t launch_split_dy(const double *dy, const double *wr, const double *wc, int64_t npos, int k,
                           const double *scale, const int32_t *order, int64_t nslots, void *dyw, hipStream_t s);
hipError_t launch_shift_grad_rows(const ShiftGradRows &a, hipStream_t s);

// The same adjoint for rows of 64, 128, 192 or 256 pixels, with the positions in batches of 16 of one column shift ic
// (the host's order, padded per shift): block (band, g) owns shift_grad_t_band(qx) output rows and the batches of
// group g, each wave two 32-pixel tiles of one row. Per batch the band + 1 pattern rows each slot reads are staged
// in LDS as FP16 lo (and, for batches with counts of 1024 or more, hi) parts, transposed to [row][pixel][slot], so
// each corner's A operand (8 positions of one pixel, the column moved by ic) is one aligned 16-byte read and the B
// operands are split_dy's, in batch order.
struct ShiftGradT {
    const uint16_t *cube;
    int32_t qy, qx;
    int64_t nbatch;           // batches of 16 slots
    const int32_t *slots;     // [nbatch][16][2] (position x qy < 2^32, ir); a padding slot repeats a position of its batch
    const int32_t *bic;       // [nbatch] the batch's column shift
    const void *dyw;          // [nbatch][corner][hi, lo][lane][8] FP16 (split_dy with the batch order)
    const double *dy_inv;     // [32] 1 / column scale (0 past the masks)
    double *partial;          // [ngroups][32][qy * qx]
    int32_t ngroups;
    int64_t bpg;              // batches per group
};
// Output rows per block (0: not supported); the kernel needs qy >= that many rows.
int shift_grad_t_band(int qx);
// Blocks of shift_grad_t that run at once on a compute unit.
int shift_grad_t_blocks_per_cu(int qx);
bool shift_grad_t_supported(int qy, int qx);
hipError_t launch_shift_grad_t(const ShiftGradT &a, hipStream_t s);
// out[i] = sum_g partial[g][i] in group order, i < n.
hipError_t launch_sum_groups(const double *partial, int ngroups, int64_t n, double *out, hipStream_t s);

}  // namespace cubo
