// cube.hpp — the resident cube handle shared by the host translation units of libcubo_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

struct cubo_hip_cube {
    int device;
    int64_t shape[4];       // [scanRows, scanColumns, detectorRows, detectorColumns]
    int dtype;              // CUBO_HIP_U16 or CUBO_HIP_F32
    int64_t npos, ndet;     // scan positions, detector pixels
    uint64_t bytes;
    void *data;             // one HBM buffer, storage order
    hipStream_t stream;     // every launch on this cube goes to this stream
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timed;   // fused-kernel launches not yet queried
};
