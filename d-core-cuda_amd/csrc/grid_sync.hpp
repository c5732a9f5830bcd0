// Bounded grid-wide barrier for cooperatively launched kernels.
//
// A sense-free arrival counter plus a generation word, with agent-scope
// release/acquire fences so that every global write made before the
// barrier is visible to every workgroup after it (across XCDs).  Waiting is
// bounded by wall-clock time: a workgroup that waits longer than
// `timeout_ticks` raises the abort word, and every workgroup that sees the
// abort word leaves its barrier with `false`, so no wavefront can wait
// forever even if the grid were not co-resident.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dcx {

enum : uint32_t {
    kSyncArrive = 0,   // arrivals at the current barrier
    kSyncGen = 1,      // barrier generation
    kSyncAbort = 2,    // nonzero: leave every barrier with false
    kSyncError = 3,    // error code reported by the kernel (0 = none)
    kSyncWords = 4,
    kSyncResetWords = 3  // zeroed before each launch; the error word persists
};

// Error codes written to word kSyncError.
enum : uint32_t { kErrNone = 0, kErrTimeout = 1, kErrInconsistent = 2 };

struct GridSync {
    uint32_t* w;             // kSyncWords words, zeroed before the launch
    uint32_t nblocks;        // gridDim.x
    uint64_t timeout_ticks;  // wall_clock64 ticks one barrier may take
};

__device__ inline void sync_fail(const GridSync& g, uint32_t code) {
    atomicMax(g.w + kSyncError, code);
    __hip_atomic_store(g.w + kSyncAbort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline bool sync_aborted(const GridSync& g) {
    return __hip_atomic_load(g.w + kSyncAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// Every thread of every workgroup must call this.  Returns false when the
// launch is being aborted.
__device__ inline bool grid_barrier(const GridSync& g) {
    __shared__ uint32_t s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen =
            __hip_atomic_load(g.w + kSyncGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(g.w + kSyncArrive, 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_AGENT);
        if (arrived + 1 == g.nblocks) {
            __hip_atomic_store(g.w + kSyncArrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(g.w + kSyncGen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(g.w + kSyncGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                   gen) {
                if (sync_aborted(g)) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > g.timeout_ticks) {
                    sync_fail(g, kErrTimeout);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (sync_aborted(g)) ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

}  // namespace dcx
