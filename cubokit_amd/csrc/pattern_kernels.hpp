// pattern_kernels.hpp — launch interface of the selected-area pattern kernels (pattern_kernels.hip).
//
// Per detector pixel, the mean or the maximum over a list of scan positions of the pattern, each optionally
// multiplied by a per-position weight and rolled periodically by a per-position integer shift
// (SelectedDiffraction.swift:110-179 and 225-271, ShiftedImaging.swift:125-167). One pass over the selected
// patterns of the resident cube: lanes own 16-byte pixel vectors and walk the positions, so every pattern is
// read once, in full cache lines, and no pattern outside the selection is read.
//   - integer: uint16 cube, no weights: uint32 per-lane sums and packed 16-bit maxima, exact;
//   - FP64: any other case: each value (times its weight) in Double, compensated sums (TwoSum) per lane and
//     across lanes, and the maximum in Double with the reference's NaN and tie rules.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cubo {

enum { kPatternMean = 0, kPatternMax = 1 };

struct PatternLaunch {
    const void *cube;         // the cube's first element
    int64_t ndet;
    int32_t qy, qx;           // detector extents (the roll)
    int64_t first;            // the launch's first position (no list) or first list entry
    int64_t count;            // positions or list entries in this launch
    const int32_t *list;      // the selected positions in scan order, or NULL for positions first .. first + count
    const double *w;          // [npos] weights by position, or NULL
    const uint32_t *shift;    // [npos] (row << 16 | column) roll of each pattern, both in [0, extent), or NULL
    int32_t ngroups;
    int64_t ppg;              // entries per group
    void *sum;                // [ngroups][ndet] uint32 (integer) or double2 (sum, correction)
    void *max;                // [ngroups][ndet] uint32 (integer) or double
};

struct PatternKind {
    bool u16;                 // cube element type (else float)
    bool integer;             // the exact integer family (uint16, no weights)
    int stat;                 // kPatternMean or kPatternMax
    int vec;                  // pixels per lane: 16 bytes of elements, or 1
    bool shift;
    bool first_wins;          // of equal maxima the first is kept (np.max), else the later (Swift.max)
};

// Pixels per lane for a cube: 16 bytes of elements when every pattern starts 16-byte aligned (and, with a
// roll, the detector rows hold whole vectors), else 1.
int pattern_vec(bool u16, int64_t ndet, int qx, bool shift);
// Pixels covered by one block (4 waves over the same pixels, each a quarter of the group's entries).
int pattern_block_pixels(int vec);
hipError_t launch_pattern(const PatternLaunch &a, const PatternKind &k, hipStream_t s);

// Running totals over launches, one thread per pixel, groups combined in order. first: start from the seeds
// (0, or +0 and -inf); finish: write value / count (mean) or the maximum to out (NaN canonical) instead of the
// running state. integer: run_sum uint64, run_max uint32; FP64: run_sum double2, run_max double.
hipError_t launch_pattern_accumulate(const PatternKind &k, const void *part_sum, const void *part_max, int ngroups,
                                     int64_t ndet, void *run_sum, void *run_max, bool first, bool finish, double count,
                                     double *out, hipStream_t s);

// PatternShift.fourier (SelectedDiffraction.swift:273-285, ShiftedImaging.swift:145-153): each pattern moved by
// Re(IFFT2(FFT2(p) exp(-2 pi i (sc fc + sr fr)))), (sr, sc) the map's mean minus the origin, fr, fc NumPy's fftfreq.
// A batch of entries is expanded to complex, transformed by the batched FP64 FFT (dpc_kernels.hpp), multiplied by
// the phase ramp, transformed back and reduced, in entry order, into one group partial of the pattern kernels'
// FP64 layout (sum and correction, maximum), which launch_pattern_accumulate folds into the running totals.
struct FourierBatch {
    const void *cube;         // position p's pattern at cube + p * ndet * element size
    bool u16;
    int64_t ndet;
    int32_t qy, qx;
    int64_t first;            // the batch's first entry
    int32_t batch;
    const int32_t *list;      // entry -> position, or NULL (the entry is the position)
    const double *w;          // [npos] weights, or NULL
    const double *sr, *sc;    // [npos] the shift in pixels
    const double *fr, *fc;    // fftfreq(qy), fftfreq(qx)
    double2 *data;            // [batch][ndet]
    double2 *part_sum;        // [ndet]
    double *part_max;         // [ndet]
    int stat;
    bool first_wins;
};
hipError_t launch_fourier_expand(const FourierBatch &b, hipStream_t s);

// The same shift and reduction in one kernel for power-of-two detectors of at most 16384 pixels (rows >= 16,
// 16 <= columns <= 256): block g takes the entries first + g ppg .. first + (g + 1) ppg - 1 in order, one pattern
// at a time, with the pattern's half spectrum (rows x (columns / 2 + 1), FP64 complex) in LDS. Rows are
// transformed two at a time as one complex FFT, the columns in place, the phase ramp made Hermitian (so the real
// part of the inverse is the inverse of a real pattern) and the inverse taken the same way back; each lane adds
// the shifted values of its own pixels (TwoSum, or the maximum) in registers and the block writes one group
// partial of the layout launch_pattern_accumulate reads.
struct FourierFused {
    const void *cube;         // position p's pattern at cube + p * ndet * element size
    bool u16;
    int32_t qy, qx;
    int64_t ndet;
    int64_t first, count;     // entries first .. first + count - 1
    const int32_t *list;      // entry -> position, or NULL
    const double *w;          // [npos] weights, or NULL
    const double *sr, *sc;    // [npos] the shift in pixels
    const double2 *tw;        // [max(qy, qx)] exp(-2 pi i m / max(qy, qx))
    int32_t ngroups;
    int64_t ppg;              // entries per group
    double2 *part_sum;        // [ngroups][ndet]
    double *part_max;         // [ngroups][ndet]
    int stat;
    bool first_wins;
};
bool fourier_fused_supported(int qy, int qx);
hipError_t launch_fourier_fused(const FourierFused &a, hipStream_t s);
hipError_t launch_fourier_phase(const FourierBatch &b, hipStream_t s);
hipError_t launch_fourier_reduce(const FourierBatch &b, hipStream_t s);

}  // namespace cubo
