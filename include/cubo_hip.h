/*
 * cubo_hip.h — C ABI of the MI355X (gfx950) backend of CuboKit's first-processing hot path.
 *
 * This is the boundary a third `ComputePath.hip` binds, beside `.cpu` and `.metal`
 * (Sources/CuboVirtual/Results.swift:13-19). Plain C: opaque handles, pointers and sizes, int status
 * codes (0 = success, negative = failure, the shape of the package's existing C targets,
 * Sources/CStrainLAPACK/LeastSquares.c:4-14), a thread-local message for the last failure, and
 * caller-owned host output buffers. No callbacks cross the ABI: cancellation and progress are the
 * caller's loop over cubo_hip_reduce_rows calls, one per scan tile.
 *
 * Threading: every cube handle owns one HIP stream; calls on different handles may run from
 * different threads; calls on one handle must not overlap.
 */
#ifndef CUBO_HIP_H
#define CUBO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The Swift side maps them to CuboVirtualError.noHIPDevice / .hip(String) and
 * ResidentCubeError.unavailable (Sources/CuboVirtual/Results.swift:65-75,
 * Sources/CuboVirtual/ResidentCube.swift:14-21). A refusal is never a fallback. */
enum {
    CUBO_HIP_OK = 0,
    CUBO_HIP_ERR_NO_DEVICE = -1,
    CUBO_HIP_ERR_OUT_OF_MEMORY = -2,
    CUBO_HIP_ERR_EXCEEDS_CAP = -3,
    CUBO_HIP_ERR_INVALID_ARGUMENT = -4,
    CUBO_HIP_ERR_HIP = -5,
    CUBO_HIP_ERR_UNSUPPORTED = -6,
    CUBO_HIP_ERR_INCOMPLETE = -7,
    CUBO_HIP_ERR_EXCEEDS_BUFFER_LIMIT = -8,
    CUBO_HIP_ERR_FILE = -9,    /* a CuboError of the file readers (Sources/CuboCore/CuboError.swift); the message
                                  starts with the case: cannotOpen, notNPY, fortranOrderNotSupported,
                                  bigEndianNotSupported, unsupportedElementType, wrongRank, fileTooShort,
                                  misalignedOffset, invalidShape */
    CUBO_HIP_ERR_CANCELLED = -10  /* the caller's cancel flag was set: Swift's CancellationError
                                     (Sources/CuboPhase/ParallelScores.swift:20-38); nothing was written */
};

/* Element types held on the device: the native detector counts (uint16) or Float32, the type every
 * other element type becomes on read (Sources/CuboCore/MappedCube.swift:56-83). */
enum { CUBO_HIP_U16 = 1, CUBO_HIP_F32 = 2 };

typedef struct cubo_hip_cube cubo_hip_cube;

/* ---- devices (replaces MetalContext.shared, Sources/CuboVirtual/MetalReducers.swift:21-58) ---- */
typedef struct {
    uint64_t hbm_bytes;        /* total device memory */
    uint64_t free_bytes;       /* free device memory now (hipMemGetInfo) */
    uint64_t max_alloc_bytes;  /* largest single allocation the library will attempt */
    int32_t compute_units;
    int32_t wavefront_size;
    char name[64];
    char arch[32];             /* e.g. "gfx950" */
} cubo_hip_device_info;

int cubo_hip_device_count(int *count);
int cubo_hip_get_device_info(int device, cubo_hip_device_info *info);

/* Message of the last failure on the calling thread (always NUL-terminated). */
int cubo_hip_last_error(char *buffer, size_t size);

/* ---- resident cube (replaces ResidentCube, Sources/CuboVirtual/ResidentCube.swift:31-128) ---- */

/* Reasons for cubo_hip_availability, in the reference's order of checks
 * (ResidentCube.swift:93-103): the memory cap first, then the device, then the largest buffer. */
enum {
    CUBO_HIP_AVAILABLE = 0,
    CUBO_HIP_UNAVAILABLE_NO_DEVICE = 1,
    CUBO_HIP_UNAVAILABLE_EXCEEDS_MEMORY_CAP = 2,
    CUBO_HIP_UNAVAILABLE_EXCEEDS_BUFFER_LIMIT = 3
};

/* Fraction of the device's HBM one resident cube may take (ResidentCube.memoryCapFraction is 1/2 of
 * host memory on Apple's unified memory; an MI355X cube lives in its own 288 GB of HBM). */
#define CUBO_HIP_MEMORY_CAP_FRACTION 0.9

/* availability(of:physicalMemory:) (ResidentCube.swift:93-103). hbm_bytes_override = 0 uses the
 * device's HBM; a non-zero value injects a memory size (the reference's injected physicalMemory,
 * Tests/CuboVirtualTests/ResidentCubeAvailabilityTests.swift:21-37). Writes the reason, the cube's
 * bytes and the cap. */
int cubo_hip_availability(int device, const int64_t shape[4], int dtype, uint64_t hbm_bytes_override,
                          int *reason, uint64_t *bytes, uint64_t *cap);

/* init(loading:) (ResidentCube.swift:43-58): allocates one device buffer for the cube in its native
 * dtype. shape = [scanRows, scanColumns, detectorRows, detectorColumns] (CubeShape.swift:11-58). */
int cubo_hip_cube_create(int device, const int64_t shape[4], int dtype, cubo_hip_cube **out);
/* Copies scan rows [row0, row0 + nrows) from host memory in the cube's dtype (the tile-by-tile load,
 * ResidentCube.swift:49-55). */
int cubo_hip_cube_upload_rows(cubo_hip_cube *cube, int64_t row0, int64_t nrows, const void *host);
/* readScanTile (ResidentCube.swift:113-117): copies rows out to host memory. */
int cubo_hip_cube_read_rows(cubo_hip_cube *cube, int64_t row0, int64_t nrows, void *host);
/* readPattern(row:column:) (ResidentCube.swift:111-116) for n scan positions (row * scanColumns + column) at once:
 * n patterns in the cube's dtype, one after another, into host memory. Positions outside the scan are refused
 * before any copy. */
int cubo_hip_cube_read_patterns(cubo_hip_cube *cube, const int64_t *indices, int64_t n, void *host);
int cubo_hip_cube_destroy(cubo_hip_cube *cube);
/* init(loading:) from another resident cube: a new cube on the same device with the same shape, dtype and
 * contents, copied on the device by one streaming kernel (every byte read once and written once). */
int cubo_hip_cube_clone(const cubo_hip_cube *source, cubo_hip_cube **out);
/* byteCount (ResidentCube.swift:109). */
int cubo_hip_cube_byte_count(const cubo_hip_cube *cube, uint64_t *bytes);

/* ---- ingestion: .npy and raw files, and host tiles, into HBM (NPYFile.header,
 * Sources/CuboCore/NPYFile.swift:28-130; MappedCube, Sources/CuboCore/MappedCube.swift:98-164; the staging case of
 * TileFeed, Sources/CuboVirtual/MetalReducers.swift:494-573) ---- */

/* Stored element types, in ElementType's order (MappedCube.swift:15-17). uint16 is kept as uint16 on the device;
 * every other type becomes Float32 there, converted on the device as ElementType.convert does (float64 and the
 * integers round to nearest even, MappedCube.swift:56-83). */
enum {
    CUBO_HIP_ELEMENT_F32 = 0, CUBO_HIP_ELEMENT_F64 = 1, CUBO_HIP_ELEMENT_U8 = 2, CUBO_HIP_ELEMENT_I8 = 3,
    CUBO_HIP_ELEMENT_U16 = 4, CUBO_HIP_ELEMENT_I16 = 5, CUBO_HIP_ELEMENT_U32 = 6, CUBO_HIP_ELEMENT_I32 = 7,
    CUBO_HIP_ELEMENT_U64 = 8, CUBO_HIP_ELEMENT_I64 = 9
};
typedef struct {
    int32_t major_version, minor_version;
    int32_t element_type;      /* CUBO_HIP_ELEMENT_* */
    int32_t rank;              /* at most 32 */
    int64_t shape[32];
    int64_t data_offset;       /* byte offset of the first element */
} cubo_hip_npy_header;
/* NPYFile.header(of:) (NPYFile.swift:28-71): versions 1.0, 2.0 and 3.0, the reference's dict grammar and refusals.
 * Host only: no device is needed. */
int cubo_hip_npy_read_header(const char *path, cubo_hip_npy_header *header);
/* MappedCube(url:dataOffset:elementType:shape:) + ResidentCube(loading:): the file region
 * [data_offset, data_offset + cube bytes) into a new resident cube. Worker threads read the next chunk into one
 * pinned buffer while the previous chunk's copy to HBM (and conversion) runs from the other. The file checks
 * (misalignedOffset, cannotOpen, invalidShape, fileTooShort) come before any device call. */
int cubo_hip_cube_load_file(int device, const char *path, int64_t data_offset, int32_t element_type,
                            const int64_t shape[4], cubo_hip_cube **out);
/* MappedCube(npy:) + ResidentCube(loading:): the header's rank must be 4 (wrongRank otherwise). */
int cubo_hip_cube_load_npy(int device, const char *path, cubo_hip_cube **out);
/* A cube whose scan rows are fed tile by tile from host memory instead of being held whole (the reference's
 * staging buffer for sources that are not resident): a two-slot device ring of tile_rows scan rows each. Its
 * tiles are reduced by cubo_hip_reduce_rows and cubo_hip_pattern_rows like a resident cube's; calls that read the
 * whole cube refuse it. */
int cubo_hip_cube_create_streamed(int device, const int64_t shape[4], int dtype, int64_t tile_rows,
                                  cubo_hip_cube **out);
/* Copies scan rows [row0, row0 + nrows) of `host` (stored as element_type, e.g. a mapped file) to the device,
 * asynchronously: through two pinned buffers and a copy stream, converted on the device where the type is not the
 * cube's. On a resident cube they land in place; on a streamed cube in the next ring slot (nrows <= tile_rows).
 * Later launches on the cube wait for the copy; the call returns once `host` has been read, so the next tile's
 * host read overlaps this tile's copy and reduction. */
int cubo_hip_cube_feed_rows(cubo_hip_cube *cube, int64_t row0, int64_t nrows, const void *host, int32_t element_type);

/* ---- the fused reduction: VirtualImaging.image, VirtualDiffraction.mean/max/meanAndMax and
 * CenterOfMass.measure in one pass over the cube (VirtualImaging.swift:20-66,
 * CenterOfMass.swift:27-47). ---- */

typedef struct {
    const float *masks;        /* n_masks x detectorCount detector masks (DetectorMask.make output), or NULL */
    int32_t n_masks;
    int32_t want_mean_max;     /* column statistics: mean and max patterns */
    int32_t want_com;          /* centre of mass: S, Mr, Mc per position */
    const float *com_weights;  /* detectorCount CoM weights, or NULL for all ones (CenterOfMass.swift:31-33) */
} cubo_hip_reduce_request;

typedef struct {
    double *images;            /* n_masks x scanCount, row-major per mask, or NULL */
    double *mean;              /* detectorCount, or NULL */
    double *max;               /* detectorCount, or NULL */
    double *com_intensity;     /* scanCount: S = sum w P */
    double *com_moment_row;    /* scanCount: sum (w i) P */
    double *com_moment_column; /* scanCount: sum (w j) P */
    double *com_row;           /* scanCount: Mr / S (NaN where S = 0) */
    double *com_column;        /* scanCount: Mc / S */
} cubo_hip_reduce_result;

typedef struct cubo_hip_reduction cubo_hip_reduction;

/* A reduction is begun, fed scan-row tiles in scan order (the caller asks isCancelled before each
 * and reports progress after each, Sources/CuboVirtual/Interruption.swift:15-25), and ended, which
 * waits for the device and writes the caller's buffers. Every row must be fed exactly once before
 * end; cancel frees a reduction without results. */
int cubo_hip_reduce_begin(cubo_hip_cube *cube, const cubo_hip_reduce_request *request, cubo_hip_reduction **out);
int cubo_hip_reduce_rows(cubo_hip_reduction *r, int64_t row0, int64_t nrows);
int cubo_hip_reduce_end(cubo_hip_reduction *r, cubo_hip_reduce_result *result);
int cubo_hip_reduce_cancel(cubo_hip_reduction *r);
/* Reuse without reallocation, for callers that run the same request on the same cube repeatedly:
 * results waits and writes the caller's buffers but keeps the reduction; reset starts a new pass
 * (every row again); release frees it. reduce_end = results + release. */
int cubo_hip_reduce_results(cubo_hip_reduction *r, cubo_hip_reduce_result *result);
/* The same results without a copy: pointers into the reduction's pinned host memory, valid until the next
 * reset, rows or release call on it (NULL for what was not requested). */
typedef struct {
    const double *images, *mean, *max, *com_intensity, *com_moment_row, *com_moment_column, *com_row, *com_column;
} cubo_hip_result_view;
int cubo_hip_reduce_view(cubo_hip_reduction *r, cubo_hip_result_view *view);
int cubo_hip_reduce_reset(cubo_hip_reduction *r);
int cubo_hip_reduce_release(cubo_hip_reduction *r);
/* One call: begin, all rows as one tile, end. */
int cubo_hip_reduce(cubo_hip_cube *cube, const cubo_hip_reduce_request *request, cubo_hip_reduce_result *result);

/* cubo_hip_reduce without device allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same kernels and results as cubo_hip_reduce on a resident cube, all rows as one tile, with every scratch
 * buffer (weight rows, partials, results before they reach `result`) in a static device array of 64 MiB per
 * device, held from the library's first such call on a device until the process ends. A request that needs more
 * is refused with CUBO_HIP_ERR_EXCEEDS_BUFFER_LIMIT before any device work (the message gives the size it needs);
 * streamed cubes are refused (CUBO_HIP_ERR_UNSUPPORTED). Calls on one device share that array: they run one at a time (a lock held until
 * the results are in `result`), so calls that overlap on one device wait for each other; the array keeps the last
 * call's contents, and every call overwrites what it reads. Synchronous: returns once `result` is written. */
int cubo_hip_reduce_noalloc(cubo_hip_cube *cube, const cubo_hip_reduce_request *request,
                            cubo_hip_reduce_result *result);

/* Which kernel family a request runs on: 1 = exact integer (uint16 cube, non-negative integral
 * weights), 2 = FP64 accumulation. Lets tests assert the path they exercise. */
int cubo_hip_reduce_path(cubo_hip_cube *cube, const cubo_hip_reduce_request *request, int *path);

/* Device time of the fused streaming kernel over the last `cubo_hip_reduce_rows` calls since the
 * previous query, from HIP events on the cube's stream: total milliseconds and launch count. */
int cubo_hip_cube_kernel_time(cubo_hip_cube *cube, double *total_ms, int64_t *launches);

/* ---- selected-area and centre-shifted patterns: VirtualDiffraction.pattern(of:statistic:scan:shift:) for the
 * mean and the max over a .mask or .weights selection or every position, each pattern optionally moved by
 * PatternShift.rounded (Sources/CuboVirtual/SelectedDiffraction.swift:46-179, 225-271; the Metal path
 * MetalReducers.weightedPattern, MetalReducers.swift:575-617), and VirtualDiffraction.meanAndMax(of:shift:)
 * with .rounded or .fourier, one statistic per pass (ShiftedImaging.swift:125-167). The median is not on the
 * device path. ---- */
enum { CUBO_HIP_PATTERN_MEAN = 0, CUBO_HIP_PATTERN_MAX = 1 };
enum {
    CUBO_HIP_TIES_FIRST = 0,   /* of equal maxima the first in scan order is kept (np.max; pattern(of:...),
                                  SelectedDiffraction.swift:161, 260) */
    CUBO_HIP_TIES_LATER = 1    /* the later is kept (Swift.max; meanAndMax(of:shift:), ShiftedImaging.swift:160) */
};
typedef struct {
    int32_t statistic;         /* CUBO_HIP_PATTERN_MEAN or CUBO_HIP_PATTERN_MAX */
    const uint8_t *take;       /* scanCount flags, non-zero = selected (ScanSelection.mask), or NULL */
    const double *weights;     /* scanCount finite weights (ScanSelection.weights), or NULL; not with take */
    const double *origin_row;  /* PatternShift.rounded: the origin map, scanCount each, or both NULL: each */
    const double *origin_col;  /*   pattern is rolled by round_half_even(mean - origin) before it is combined */
    int32_t max_ties;          /* CUBO_HIP_TIES_FIRST or CUBO_HIP_TIES_LATER */
    int32_t shift_kind;        /* CUBO_HIP_SHIFT_ROUNDED (PatternShift.rounded) or CUBO_HIP_SHIFT_FOURIER
                                  (PatternShift.fourier: Re(IFFT2(FFT2(p) exp(-2 pi i (sc fc + sr fr)))) with the
                                  unrounded shift, FP64 on the device, detectors up to 2048 pixels a side) */
} cubo_hip_pattern_request;
enum { CUBO_HIP_SHIFT_ROUNDED = 0, CUBO_HIP_SHIFT_FOURIER = 1 };
/* Mean: sum of value (times weight) / divisor, the divisor being the selected count, the Neumaier sum of the
 * weights in scan order, or scanCount. Max: of value (times weight), seeded at -infinity, NaN winning.
 * Refusals (CUBO_HIP_ERR_INVALID_ARGUMENT), as SelectedDiffraction.swift:60-102 and ShiftedImaging.swift:99-106:
 * non-finite weights, an empty mask, weights summing to zero for the mean, a non-finite origin map.
 * A session takes scan-row tiles in order like the fused reduction and writes detectorCount Doubles. */
typedef struct cubo_hip_pattern_session cubo_hip_pattern_session;
int cubo_hip_pattern_begin(cubo_hip_cube *cube, const cubo_hip_pattern_request *request, cubo_hip_pattern_session **out);
int cubo_hip_pattern_rows(cubo_hip_pattern_session *pattern, int64_t row0, int64_t nrows);
/* results waits and copies (the session is kept), reset starts a new pass, release frees, end = results + release,
 * cancel frees without results. */
int cubo_hip_pattern_results(cubo_hip_pattern_session *pattern, double *out);
int cubo_hip_pattern_reset(cubo_hip_pattern_session *pattern);
int cubo_hip_pattern_release(cubo_hip_pattern_session *pattern);
int cubo_hip_pattern_end(cubo_hip_pattern_session *pattern, double *out);
int cubo_hip_pattern_cancel(cubo_hip_pattern_session *pattern);
/* One call: begin, all rows as one tile, end. */
int cubo_hip_pattern(cubo_hip_cube *cube, const cubo_hip_pattern_request *request, double *out);
/* 1 = exact integer kernel (uint16 cube, no weights), 2 = FP64 (compensated sums). */
int cubo_hip_pattern_path(cubo_hip_cube *cube, const cubo_hip_pattern_request *request, int *path);

/* ---- origins: py4DSTEM's get_origin (OriginMeasurement.measure, Sources/CuboVirtual/OriginMeasurement.swift:52-95)
 * and the iterated-window estimator (OriginMeasurement.measure(of:radius:radiusScale:estimator:...),
 * IteratedWindowOrigin.swift:136-241; the Metal kernel cubo_measure_origin, MetalReducers.swift:179-266), per scan
 * position on the device, and ProbeSize.estimate (ProbeSize.swift:25-58) on the host ---- */
typedef struct {
    double radius, center_row, center_column, threshold;
} cubo_hip_probe_estimate;
/* py4DSTEM's get_probe_size step for step in Double (NumPy's linspace, gradient, median and pairwise sums): the
 * radius of the bright disk of a pattern (normally the maximum pattern). Host only. "probeNotFound" refusals. */
int cubo_hip_probe_size(const double *pattern, int32_t rows, int32_t cols, double lower_threshold,
                        double upper_threshold, int32_t count, cubo_hip_probe_estimate *out);
enum { CUBO_HIP_ORIGIN_PY4DSTEM = 0, CUBO_HIP_ORIGIN_ITERATED_WINDOW = 1 };
typedef struct {
    int32_t estimator;         /* CUBO_HIP_ORIGIN_* */
    double radius;             /* the probe radius in pixels (sigma of the Gaussian, or the window's radius) */
    double radius_scale;       /* 1.2 */
    const uint8_t *scan_mask;  /* py4DSTEM: scanCount flags, origins only where set (NaN elsewhere), or NULL */
    int32_t fast_center;       /* py4DSTEM: the filtered argmax itself, no centre of mass */
    int32_t passes;            /* iterated window: 0 ... 1000 (4) */
    double minimum_margin;     /* iterated window: >= 0 (1.5) */
    int32_t binned_coarse;     /* iterated window: the brightest round(radius)-sized block (1) or pixel (0) */
} cubo_hip_origin_request;
/* The origins of scan rows [row0, row0 + nrows) into out_row / out_col (scanCount Doubles each; only the tile's
 * entries are written), so a caller walks the tiles with cancellation and progress between them. */
int cubo_hip_origins(cubo_hip_cube *cube, const cubo_hip_origin_request *request, int64_t row0, int64_t nrows,
                     double *out_row, double *out_col);
/* GaussianFilter.apply (OriginMeasurement.swift:127-135) of one pattern: scipy's gaussian_filter with mode nearest,
 * in Double on the device (the filter the get_origin kernel runs). */
int cubo_hip_gaussian_filter_pattern(cubo_hip_cube *cube, int64_t row, int64_t column, double sigma, double *out);

/* ---- the mask stack: Y = X W for K weight rows in one pass, X the (scanCount x detectorCount) cube.
 * {1, q_row, q_col, masks...} gives K virtual images and the centre-of-mass sums together. No reference
 * counterpart (the reference runs VirtualImaging.image once per mask, CPUReducers.swift:13-39). ---- */
enum {
    CUBO_HIP_STACK_AUTO = 0,        /* the matrix-core method where it applies, else plain */
    CUBO_HIP_STACK_PLAIN = 1,       /* VALU: FP32 FMAs, FP64 across 32-pixel chunks; any cube */
    CUBO_HIP_STACK_MFMA_SPLIT = 2   /* uint16 cube, split-FP16 operands on MFMA, FP32 accumulators flushed to
                                       FP64 every 256 pixels; at most 256 masks (passes of 64) */
};
typedef struct cubo_hip_stack_plan cubo_hip_stack_plan;
/* masks: k x detectorCount FP32 weights, row-major. */
int cubo_hip_stack_plan_create(cubo_hip_cube *cube, const float *masks, int32_t k, int32_t method,
                               cubo_hip_stack_plan **out);
int cubo_hip_stack_plan_method(const cubo_hip_stack_plan *plan, int32_t *method);
/* y: scanCount x k, row-major, caller-owned host memory. */
int cubo_hip_stack_plan_run(cubo_hip_stack_plan *plan, double *y);
int cubo_hip_stack_plan_destroy(cubo_hip_stack_plan *plan);
int cubo_hip_mask_stack(cubo_hip_cube *cube, const float *masks, int32_t k, int32_t method, double *y);

/* ---- batches: mask stacks over many resident uint16 cubes of uneven scan and detector size in ONE
 * persistent launch, replayed from a hipGraph. No reference counterpart: the reference reduces one cube
 * per call (CuboVirtual/VirtualImaging.swift:13-30); a session that screens many acquisitions calls it
 * in a loop. Each item: a uint16 cube whose pattern length is a multiple of 8, 1..64 masks (FP32,
 * k x detectorCount row-major), results as in cubo_hip_mask_stack (split-FP16 matrix-core method). All
 * cubes on one device; they must stay alive and unchanged in shape while the batch exists. ---- */
typedef struct cubo_hip_batch cubo_hip_batch;
int cubo_hip_batch_create(int32_t n, cubo_hip_cube *const *cubes, const float *const *masks, const int32_t *k,
                          cubo_hip_batch **out);
/* Times every launch shape (waves per block 8 or 4, resident blocks per compute unit, items in the given
 * order (0), largest detector first (1) or interleaved so that writing results to host memory overlaps
 * reading the cubes (2)) on the device, keeps the fastest for this batch and for every later batch of the
 * same shapes in this process, and rebuilds the graph. Any output pointer may be NULL. */
int cubo_hip_batch_tune(cubo_hip_batch *batch, int32_t *waves, int32_t *blocks_per_cu, int32_t *order);
/* One graph replay: counter reset, the persistent launch, copies of every result. y[i]: scanCount_i x k_i
 * Doubles, caller-owned host memory (NULL entries or y == NULL skip the copy out). */
int cubo_hip_batch_run(cubo_hip_batch *batch, double *const *y);
/* Cube i's results of the last run in place, scanCount_i x k_i Doubles in the batch's pinned host buffer (no
 * copy): valid until the next cubo_hip_batch_run or cubo_hip_batch_destroy. An error before the first run. */
int cubo_hip_batch_view(const cubo_hip_batch *batch, int32_t i, const double **y);
/* Device time of the graph replays since the previous query: total milliseconds and replay count. */
int cubo_hip_batch_time(cubo_hip_batch *batch, double *total_ms, int64_t *runs);
int cubo_hip_batch_destroy(cubo_hip_batch *batch);

/* ---- the shifted mask stack and its adjoint (differentiation). Forward: per position p the k masks are
 * moved by MaskShift.bilinear (Sources/CuboVirtual/ShiftedImaging.swift:43-73, 109-122) with the shift
 * origin_p - mean(origin) (row and column), and Y[p][j] = sum_u x_p[u] shifted_mask_j,p[u]. The reference
 * builds the shifted mask per position on the CPU (ShiftedImaging.swift:64-73) and has no adjoint; the
 * centre of mass relative to the origin is the stack {w, w*q_row, w*q_col} shifted the same way.
 * Backward (the vector-Jacobian product of Y, re-reading the cube): given dL/dY, dmasks = dL/dmasks
 * (k x detectorCount, Double) and dL/dorigin (row and column, scanCount each; inside a bilinear cell, through
 * the mean). Any output pointer of backward may be NULL (origin rows: both or neither). uint16 cubes whose
 * pattern length is a multiple of 8, 1..32 masks (FP32, k x detectorCount), finite origins.
 * With cubo_hip_shift_keep_jacobian(plan, 1), the bilinear forward also keeps, on the device, the Jacobian
 * rows dY/d(shift row) and dY/d(shift column) (2 x scanCount x k Doubles). A backward on the same origin map, with the cube unchanged since that forward, then takes dL/dorigin
 * from them (a dot product per position), and only dL/dmasks re-reads the cube. Otherwise the origin adjoint
 * is a pass over the cube of its own. ---- */
typedef struct cubo_hip_shift_plan cubo_hip_shift_plan;
int cubo_hip_shift_plan_create(cubo_hip_cube *cube, const float *masks, int32_t k, cubo_hip_shift_plan **out);
/* y: scanCount x k Double, row-major. */
int cubo_hip_shift_forward(cubo_hip_shift_plan *plan, const double *origin_row, const double *origin_col, double *y);
/* MaskShift.rounded (ShiftedImaging.swift:66): each position's masks rolled by wrap(round_half_even(origin - mean)),
 * no blend; one contraction per shift class on the matrix cores. Exact for uint16 cubes and integral masks. */
int cubo_hip_shift_forward_rounded(cubo_hip_shift_plan *plan, const double *origin_row, const double *origin_col,
                                   double *y);
int cubo_hip_shift_backward(cubo_hip_shift_plan *plan, const double *origin_row, const double *origin_col,
                            const double *dy, double *dmasks, double *d_origin_row, double *d_origin_col);
/* Device milliseconds since the previous query: [forward kernel, origin-adjoint kernel, mask-adjoint kernels],
 * and the number of backward calls. */
int cubo_hip_shift_kernel_time(cubo_hip_shift_plan *plan, double *ms3, int64_t *calls);
/* keep = 1: the bilinear forward keeps the Jacobian rows for the next backward (that pass costs a little more than
 * a plain forward: two more sums per output, and the rows written); 0, the default: it does not, and their memory
 * is freed. No counterpart in the reference, which has no adjoint. */
int cubo_hip_shift_keep_jacobian(cubo_hip_shift_plan *plan, int32_t keep);
/* The number of backward calls since the previous query whose dL/dorigin came from the Jacobian rows kept by the
 * forward (no pass over the cube). No counterpart in the reference, which has no adjoint; for tests and callers
 * that account for the passes. */
int cubo_hip_shift_jacobian_uses(cubo_hip_shift_plan *plan, int64_t *uses);
int cubo_hip_shift_plan_destroy(cubo_hip_shift_plan *plan);

/* ---- the centre-of-mass field's reference and rotation (CuboPhase), on the device in Double ----
 * Both calls are thread-safe: they share one stream and one set of device buffers per device (kept between
 * calls), so concurrent calls on one device run one after another. */

/* CenterOfMassReference's cases (Sources/CuboPhase/CenterOfMassReference.swift:13-22). */
enum {
    CUBO_HIP_REFERENCE_NONE = 0,      /* .none: zero */
    CUBO_HIP_REFERENCE_PLANE = 1,     /* .planeFit: the least-squares plane over the finite positions */
    CUBO_HIP_REFERENCE_CONSTANT = 2,  /* .constant: the NumPy mean of each component */
    CUBO_HIP_REFERENCE_GIVEN = 3      /* .given(field): given_row / given_col, e.g. origins measured by
                                         cubo_hip_origins (py4DSTEM's force_com_shifts) */
};
/* CenterOfMassReference.normalize (CenterOfMassReference.swift:27-84): fx/fy get the reference (the plane is refused
 * when degenerate: det <= tol * max(suu svv, 1)), nx/ny = (measured - reference) with NaN -> 0 and +-inf -> +-max
 * finite, times sampling. All arrays rows x cols; given_row/given_col are read only for CUBO_HIP_REFERENCE_GIVEN (the
 * caller checks that the given field has the measured field's shape, :46-48). tol must be finite and >= 0 (:30-32). */
int cubo_hip_phase_normalize(int device, const double *mx, const double *my, int32_t rows, int32_t cols,
                             int32_t reference, double sampling, double tol, const double *given_row,
                             const double *given_col, double *fx, double *fy, double *nx, double *ny);
/* RotationSolver.solve (Sources/CuboPhase/RotationSolver.swift:48-150): the criterion for the field rotated by an
 * angle (deg * (pi / 180); components swapped first when transposed) is the mean over the interior of
 * |d_row y' - d_col x'|, or of the divergence |d_row x' + d_col y'| when maximize. The three branches of :66-90:
 *  - force_degrees NULL, force_transpose -1: both curves over angles_deg (nangles entries each); the first best angle
 *    wins and the plain orientation wins ties;
 *  - force_degrees NULL, force_transpose 0 or 1: only that orientation's curve over angles_deg (RotationSolver.curve,
 *    :36-44); the other curve is not evaluated;
 *  - *force_degrees, force_transpose -1: both orientations scored at that angle only (one entry each);
 *  - *force_degrees, force_transpose 0 or 1: nothing scored.
 * n_curve / n_curve_t receive the number of entries written to curve / curve_t (0 for a curve not evaluated; it may
 * then be NULL). angles_deg is not read when an angle is forced. cx/cy = the field rotated by the chosen angle and
 * orientation (RotationSolution.corrected). Refusals (CUBO_HIP_ERR_INVALID_ARGUMENT): a scan under 3 x 3, a
 * non-finite field, no angles, a non-finite angle (invalidAngles). cancel: NULL or a flag another thread may set,
 * read before any work and once more after the scores (the reference asks isCancelled at the start and at the end,
 * ParallelScores.swift:20-38): CUBO_HIP_ERR_CANCELLED, no output written. */
int cubo_hip_phase_rotation(int device, const double *nx, const double *ny, int32_t rows, int32_t cols,
                            const double *angles_deg, int32_t nangles, int32_t maximize, const double *force_degrees,
                            int32_t force_transpose, const volatile int32_t *cancel, double *degrees,
                            int32_t *transposed, int32_t *n_curve, int32_t *n_curve_t, double *curve, double *curve_t,
                            double *cx, double *cy);

/* ---- iterative DPC (DPC.reconstruct, Sources/CuboPhase/DPC.swift:26-102) on the device, in Double, and
 * its adjoint. A plan fixes the scan (rows x cols), the padding (the grid is round(rows * padding) x
 * round(cols * padding), half-even, at most 2^28 points; sides above 2048 use four-step transforms for powers of two
 * and Bluestein's chirp convolution otherwise) and the sampling. Fields fx, fy are
 * rows x cols; phases on the padded grid are padded_rows x padded_cols, row-major. ---- */
typedef struct cubo_hip_dpc cubo_hip_dpc;
int cubo_hip_dpc_create(int device, int32_t rows, int32_t cols, double padding, double sampling_row,
                        double sampling_col, cubo_hip_dpc **out);
int cubo_hip_dpc_padded_shape(const cubo_hip_dpc *dpc, int32_t *rows, int32_t *cols);
/* DPC.reconstruct: from phase 0, at most max_iter iterations; an iteration whose error grew halves the step
 * instead when backtrack is set (DPC.swift:83-90); stops when the step falls below stop. phase: rows x cols;
 * errors: the accepted iterations' errors; accepted: 1/0 per iteration (max_iter entries each). */
int cubo_hip_dpc_reconstruct(cubo_hip_dpc *dpc, const double *fx, const double *fy, int32_t max_iter, double step,
                             double stop, int32_t backtrack, double *phase, double *errors, int32_t *accepted,
                             int32_t *n_iter, double *final_step);
/* One accepted iteration from phi (padded grid): phi_out = phi + step * update, and the error of phi. */
int cubo_hip_dpc_step(cubo_hip_dpc *dpc, const double *phi, const double *fx, const double *fy, double step,
                      double *phi_out, double *error);
/* Its vector-Jacobian product: from dL/dphi_out and dL/derror, dL/dphi, dL/dfx, dL/dfy. No reference
 * counterpart (the reference has no adjoint). */
int cubo_hip_dpc_step_backward(cubo_hip_dpc *dpc, const double *phi, const double *fx, const double *fy, double step,
                               const double *g_phi_out, double g_error, double *g_phi, double *g_fx, double *g_fy);
/* nsteps accepted iterations with a fixed step from phase 0 (the phase after them, cropped to the scan, and
 * their errors: phase and errors may be NULL), then the adjoint of the whole chain: given dL/dphase
 * (rows x cols) and dL/derrors (nsteps, or NULL for none), dL/dfx and dL/dfy. The iterates stay on the device. */
int cubo_hip_dpc_chain_backward(cubo_hip_dpc *dpc, const double *fx, const double *fy, int32_t nsteps, double step,
                                const double *g_phase, const double *g_errors, double *phase, double *errors,
                                double *g_fx, double *g_fy);
/* Device milliseconds of reconstruct / step / chain calls since the previous query, and their count. */
int cubo_hip_dpc_time(cubo_hip_dpc *dpc, double *ms, int64_t *calls);
int cubo_hip_dpc_destroy(cubo_hip_dpc *dpc);

/* ---- synthetic cubes for benchmarks and tests (no reference counterpart: the reference's bench
 * cubes are generated in Python, tools/benchmarks/bench.py:17-27). A disk probe with a sigmoid edge,
 * Gaussian rings, a per-position deflection and Poisson counts, generated on the device from a
 * counter-based hash of (seed, element index). ---- */
typedef struct {
    uint64_t seed;
    double center_row, center_column;  /* undeflected probe centre (detector pixels) */
    double radius;                     /* disk radius */
    double edge;                       /* sigmoid edge width (0: hard edge) */
    double counts;                     /* mean counts per pixel inside the disk */
    double background;                 /* mean counts per pixel everywhere */
    int32_t n_rings;                   /* up to 4 Gaussian rings */
    double ring_q[4], ring_amplitude[4], ring_sigma[4];
    int32_t n_waves;                   /* up to 8 deflection waves: d = a sin(2 pi (fr x + fc y) + phase) */
    double wave_row_amp[8], wave_col_amp[8], wave_fr[8], wave_fc[8], wave_phase[8];
    int32_t poisson;                   /* 1: Poisson counts, 0: the expected value */
} cubo_hip_synth_params;

int cubo_hip_cube_fill_synthetic(cubo_hip_cube *cube, const cubo_hip_synth_params *params);

#ifdef __cplusplus
}
#endif
#endif
