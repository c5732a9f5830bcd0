// Bounded barrier for a team of workgroups (up to the whole grid).
//
// Two-level arrival: each workgroup increments the counter of its group of
// kSyncFan workgroups; the last arrival of a group increments the top
// counter; the last top arrival resets the counters and bumps the
// generation word that every waiting workgroup polls.  Agent-scope
// release/acquire fences make every global write made before the barrier
// visible to every workgroup after it (across XCDs).  Waiting is bounded by
// wall-clock time: a workgroup that waits longer than `timeout_ticks`
// raises the abort word, and every workgroup that sees the abort word
// leaves its barrier with `false`, so no wavefront can wait forever even if
// the team were not co-resident.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dcx {

enum : uint32_t {
    kSyncArrive = 0,   // arrivals (top level) at the current barrier
    kSyncGen = 1,      // barrier generation
    kSyncAbort = 2,    // nonzero: leave every barrier with false
    kSyncError = 3,    // error code reported by the kernel (0 = none)
    kSyncGroups = 4,   // first group counter
    kSyncFan = 16,     // workgroups per group counter
    kSyncResetWords = 3  // zeroed before each launch; the error word persists
};

// Words needed for a team of n workgroups.
__host__ __device__ inline uint32_t sync_words(uint32_t n) {
    return kSyncGroups + (n + kSyncFan - 1) / kSyncFan;
}

// Error codes written to word kSyncError.
enum : uint32_t { kErrNone = 0, kErrTimeout = 1, kErrInconsistent = 2 };

struct GridSync {
    uint32_t* w;             // sync_words(nblocks) words, zeroed before first use
    uint32_t nblocks;        // workgroups in the team
    uint32_t rank;           // this workgroup's rank in the team
    uint64_t timeout_ticks;  // wall_clock64 ticks one barrier may take
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
__device__ inline bool grid_barrier(const GridSync& g) {
    __shared__ uint32_t s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen =
            __hip_atomic_load(g.w + kSyncGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t grp = g.rank / kSyncFan;
        const uint32_t ngroups = (g.nblocks + kSyncFan - 1) / kSyncFan;
        const uint32_t in_grp = min((uint32_t)kSyncFan, g.nblocks - grp * kSyncFan);
        bool last = false;
        const uint32_t a = __hip_atomic_fetch_add(g.w + kSyncGroups + grp, 1u, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        if (a + 1 == in_grp) {
            __hip_atomic_store(g.w + kSyncGroups + grp, 0u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t t = __hip_atomic_fetch_add(g.w + kSyncArrive, 1u, __ATOMIC_ACQ_REL,
                                                      __HIP_MEMORY_SCOPE_AGENT);
            last = t + 1 == ngroups;
        }
        if (last) {
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
