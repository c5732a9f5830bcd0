// Bounded barrier for a team of workgroups (up to the whole grid).
//
// Flag barrier: at its e-th barrier every workgroup stores e into its own
// flag word (no two workgroups write the same word, so arrivals do not
// serialise on one atomic counter); the threads of rank 0 poll all flags in
// parallel, and once every flag has reached e rank 0 stores e into the
// release word that the other workgroups poll.  Every wavefront waits
// until its stores are performed, then thread 0 issues the workgroup's one
// agent-scope release fence (an L2 write-back on CDNA), so every global
// write made before the team barrier is visible to every workgroup after
// it (across XCDs); thread 0's agent-scope acquire fence followed by the
// workgroup barrier covers the way out.  The fences dominate the cost, so
// there is one pair per workgroup and barrier, and teams are kept as small
// as their work allows.  Waiting is bounded by
// wall clock: a workgroup that waits longer than `timeout_ticks` raises the
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
    kSyncDone = 2,     // passes a team has finished (team_resync)
    kSyncEpoch = 3,    // rank 0's barrier epoch at the end of its last pass
    kSyncError = 4,    // error code reported by the kernel (0 = none)
    kSyncFlags = 5,    // first per-rank arrival flag
    kSyncResetWords = 4  // words 0-3 zeroed before each launch (flags too)
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

// Cross-workgroup words (barrier flags, phase counters) are read and
// written only with read-modify-write atomics, resolved at one point for
// the whole device.  Each one is issued WITH return and its old value
// consumed before the caller goes on: a no-return atomic counts as done
// when the L2 has accepted it, and on gfx950 one was observed to take
// effect after the next barrier (a counter reset landing on top of the
// next phase's increments).  Consuming the returned value (here: a store to
// an LDS sink, which must wait for it) means the operation has been
// performed before anything that follows in program order.
__device__ __forceinline__ void consume(uint32_t x) {
    __shared__ uint32_t s_sink;
    *(volatile uint32_t*)&s_sink = x;
}
__device__ __forceinline__ uint32_t coherent_read(uint32_t* p) { return atomicAdd(p, 0u); }
__device__ __forceinline__ void coherent_write(uint32_t* p, uint32_t v) { consume(atomicExch(p, v)); }

__device__ inline void sync_fail(const GridSync& g, uint32_t code) {
    consume(atomicMax(g.w + kSyncError, code));
    coherent_write(g.w + kSyncAbort, 1u);
}

__device__ inline bool sync_aborted(const GridSync& g) { return coherent_read(g.w + kSyncAbort) != 0; }

// Every thread of every workgroup of the team must call this.  Returns
// false when the launch is being aborted.  fenced = false skips the L2
// write-back and invalidation: only for callers whose data crossing the
// barrier is written and read with agent-coherent accesses (atomics,
// agent-scope atomic loads and stores), which the stores' completion
// (waited for below) already makes visible.
__device__ inline bool grid_barrier(GridSync& g, bool fenced = true) {
    __shared__ uint32_t s_ok;
    const uint32_t e = ++g.epoch;
    if (threadIdx.x == 0) s_ok = 1;
    // Every wavefront waits for its own stores to be performed (they are in
    // the L2 once counted), then thread 0 alone writes the L2 back: one
    // write-back per workgroup and barrier.  (A write-back per wavefront
    // costs the same for each and, with hundreds of workgroups syncing,
    // kept every XCD's L2 busy flushing.)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (fenced) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        coherent_write(g.w + kSyncFlags + g.rank, e);
    }
    if (g.rank == 0) {
        // every thread of rank 0 watches a strided subset of the flags
        const uint64_t t0 = wall_clock64();
        bool ok = true;
        for (uint32_t i = threadIdx.x; i < g.nblocks && ok; i += blockDim.x) {
            for (uint32_t spin = 1; coherent_read(g.w + kSyncFlags + i) < e; ++spin) {
                // the abort word and the clock are checked every few polls:
                // each check is another round trip on the waiting path
                if ((spin & 15) == 0) {
                    if (sync_aborted(g)) {
                        ok = false;
                        break;
                    }
                    if (wall_clock64() - t0 > g.timeout_ticks) {
                        sync_fail(g, kErrTimeout);
                        ok = false;
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (!ok) s_ok = 0;
        __syncthreads();
        // rank 0 released its own writes before raising its flag; the
        // release word carries no data of its own
        if (threadIdx.x == 0 && s_ok) coherent_write(g.w + kSyncRelease, e);
    } else if (threadIdx.x == 0) {
        const uint64_t t0 = wall_clock64();
        for (uint32_t spin = 1; coherent_read(g.w + kSyncRelease) < e; ++spin) {
            if ((spin & 15) == 0) {
                if (sync_aborted(g)) {
                    s_ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > g.timeout_ticks) {
                    sync_fail(g, kErrTimeout);
                    s_ok = 0;
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (threadIdx.x == 0 && fenced) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __builtin_amdgcn_s_waitcnt(0);  // invalidation done before anyone reads
    }
    // A member that failed raised the abort word before arriving here; the
    // others learn it at their next wait (every wait polls the abort word),
    // so the barrier itself adds no round trip to check it.
    __syncthreads();
    return s_ok != 0;
}

// Joins a team again after a pass from which members may have left early
// (a pass narrows its team as its alive list shrinks, and the narrowed
// team's barriers advance only the epochs of its members).  Rank 0 calls
// it once its own narrowed team has finished the pass: it publishes its
// epoch and the pass count `seq`; every other member waits for `seq` and
// adopts that epoch, so the next barrier has the same number in every
// member.  Every thread of every member must call it.
__device__ inline bool team_resync(GridSync& g, uint32_t seq) {
    __shared__ uint32_t s_ep, s_ok2;
    if (threadIdx.x == 0) {
        s_ok2 = 1;
        s_ep = g.epoch;
        if (g.rank == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __builtin_amdgcn_s_waitcnt(0);
            coherent_write(g.w + kSyncEpoch, g.epoch);
            coherent_write(g.w + kSyncDone, seq);
        } else {
            const uint64_t t0 = wall_clock64();
            while (coherent_read(g.w + kSyncDone) < seq) {
                if (sync_aborted(g)) {
                    s_ok2 = 0;
                    break;
                }
                if (wall_clock64() - t0 > g.timeout_ticks) {
                    sync_fail(g, kErrTimeout);
                    s_ok2 = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __builtin_amdgcn_s_waitcnt(0);
            s_ep = coherent_read(g.w + kSyncEpoch);
        }
    }
    __syncthreads();
    g.epoch = s_ep;
    const bool ok = s_ok2 != 0;
    __syncthreads();  // s_ep, s_ok2 are read before anyone reuses them
    return ok;
}

}  // namespace dcx
