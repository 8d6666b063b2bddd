// cube.hpp — the resident cube handle shared by the host translation units of libcubo_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

#include "timing.hpp"

namespace cubo {
struct Feed;
}

struct cubo_hip_cube {
    int device;
    int64_t shape[4];       // [scanRows, scanColumns, detectorRows, detectorColumns]
    int dtype;              // CUBO_HIP_U16 or CUBO_HIP_F32
    int64_t npos, ndet;     // scan positions, detector pixels
    uint64_t bytes;
    void *data;             // one HBM buffer, storage order
    hipStream_t stream;     // every launch on this cube goes to this stream
    cubo::LaunchTimer timer;   // device time of the fused-kernel launches, until queried
    // Counts the writes to the resident buffer (uploads, feeds, synthesis): results kept from a pass over the cube
    // (the shifted stack's Jacobian rows) hold only while it is unchanged.
    uint64_t generation = 0;
    // A streamed cube holds no whole-cube buffer: `data` is NULL and the scan-row tiles fed by
    // cubo_hip_cube_feed_rows live in the feed's two-slot device ring. Every cube that has been fed owns a feed
    // (pinned double buffers, a copy stream).
    bool streamed = false;
    cubo::Feed *feed = nullptr;
};

#include "errors.hpp"

namespace cubo {

inline int element_size(int dtype) { return dtype == CUBO_HIP_U16 ? 2 : dtype == CUBO_HIP_F32 ? 4 : 0; }

inline int rows_range(const cubo_hip_cube *c, int64_t row0, int64_t nrows) {
    if (row0 < 0 || nrows < 0 || row0 + nrows > c->shape[0])
        return fail(CUBO_HIP_ERR_INVALID_ARGUMENT, "rows [%lld, %lld) outside the scan's %lld rows", (long long)row0,
                    (long long)(row0 + nrows), (long long)c->shape[0]);
    return CUBO_HIP_OK;
}

// The device address of scan rows [row0, row0 + nrows): in the resident buffer, or in the ring slot a streamed
// cube was fed them into (an error when they are not staged). Defined in cubo_io.hip.
int rows_data(const cubo_hip_cube *c, int64_t row0, int64_t nrows, const void **out);
// Refuses a streamed cube for calls that read the whole cube.
inline int require_resident(const cubo_hip_cube *c, const char *what) {
    if (c->streamed)
        return fail(CUBO_HIP_ERR_UNSUPPORTED, "%s needs a resident cube, not a streamed one", what);
    return CUBO_HIP_OK;
}
// Releases a cube's feed (waits for its copies). Defined in cubo_io.hip.
void free_feed(cubo_hip_cube *c);

inline int compute_units(int dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

}  // namespace cubo
