// Bounded barrier for a team of workgroups (up to the whole grid).
//
// Flag barrier: at its e-th barrier every workgroup stores e into its own
// flag word (no two workgroups write the same word, so arrivals do not
// serialise on one atomic counter); the threads of rank 0 poll all flags in
// parallel, and once every flag has reached e rank 0 stores e into the
// release word that the other workgroups poll.  Agent-scope release and
// acquire fences make every global write made before the barrier visible
// to every workgroup after it (across XCDs).  Waiting is bounded by wall
// clock: a workgroup that waits longer than `timeout_ticks` raises the
// abort word, and every workgroup that sees the abort word leaves its
// barrier with `false`, so no wavefront can wait forever even if the team
// were not co-resident.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dcx {

enum : uint32_t {
    kSyncRelease = 0,  // epoch released by rank 0
    kSyncAbort = 1,    // nonzero: leave every barrier with false
    kSyncError = 2,    // error code reported by the kernel (0 = none)
    kSyncFlags = 3,    // first per-rank arrival flag
    kSyncResetWords = 2  // words 0-1 zeroed before each launch (flags too)
};

// Words needed for a team of n workgroups.
__host__ __device__ inline uint32_t sync_words(uint32_t n) { return kSyncFlags + n; }

// Error codes written to word kSyncError.
enum : uint32_t { kErrNone = 0, kErrTimeout = 1, kErrInconsistent = 2 };

struct GridSync {
    uint32_t* w;             // sync_words(nblocks) words; release word and flags zero at start
    uint32_t nblocks;        // workgroups in the team
    uint32_t rank;           // this workgroup's rank in the team
    uint64_t timeout_ticks;  // wall_clock64 ticks one barrier may take
    uint32_t epoch;          // barriers passed so far (same in every member)
};

__device__ inline void sync_fail(const GridSync& g, uint32_t code) {
    atomicMax(g.w + kSyncError, code);
    __hip_atomic_store(g.w + kSyncAbort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline bool sync_aborted(const GridSync& g) {
    return __hip_atomic_load(g.w + kSyncAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// Every thread of every workgroup of the team must call this.  Returns
// false when the launch is being aborted.
__device__ inline bool grid_barrier(GridSync& g) {
    __shared__ uint32_t s_ok;
    const uint32_t e = ++g.epoch;
    if (threadIdx.x == 0) s_ok = 1;
    // every wavefront publishes its own stores before the flag goes up
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(g.w + kSyncFlags + g.rank, e, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (g.rank == 0) {
        // every thread of rank 0 watches a strided subset of the flags
        const uint64_t t0 = wall_clock64();
        bool ok = true;
        for (uint32_t i = threadIdx.x; i < g.nblocks && ok; i += blockDim.x) {
            while (__hip_atomic_load(g.w + kSyncFlags + i, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) < e) {
                if (sync_aborted(g)) {
                    ok = false;
                    break;
                }
                if (wall_clock64() - t0 > g.timeout_ticks) {
                    sync_fail(g, kErrTimeout);
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (!ok) s_ok = 0;
        __syncthreads();
        if (threadIdx.x == 0 && s_ok) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(g.w + kSyncRelease, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (threadIdx.x == 0) {
        const uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(g.w + kSyncRelease, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
               e) {
            if (sync_aborted(g)) {
                s_ok = 0;
                break;
            }
            if (wall_clock64() - t0 > g.timeout_ticks) {
                sync_fail(g, kErrTimeout);
                s_ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (sync_aborted(g)) s_ok = 0;
    }
    __syncthreads();
    return s_ok != 0;
}

}  // namespace dcx
