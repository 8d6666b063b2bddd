// device_sync.hpp — block barrier for LDS hand-offs.
#pragma once

#include <hip/hip_runtime.h>

namespace cubo {

// A block barrier that orders LDS accesses only. __syncthreads() also orders global memory, so the
// compiler waits for every global load in flight before it (s_waitcnt vmcnt(0)), which drains a
// register prefetch at each barrier. Here only LDS operations complete first (lgkmcnt(0)); global loads
// stay in flight and are waited for where their registers are used.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace cubo
