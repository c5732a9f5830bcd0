// Peel kernels for D-core decomposition on CDNA4 (gfx950, wave64).
//
// One pass computes, for a fixed k, every vertex's level l[v] =
// max{l : v in D_{k,l}} (0 outside the (k,0)-core), which is what the
// reference's scan/process pair computes level by level
// (main.cu:27-141, driven by KListInplace main.cu:171-195).  The same
// pass with the two adjacency directions swapped and k = 0 gives the
// in-core numbers kmaxes (dcore main.cu:218-220).
//
// Work organisation (not the reference's):
//   * l-rows run on the core-ordered graph (core_graph.hip): members of the
//     (k,0)-core are ranks [0, V_k) and the member neighbours of a member
//     are a prefix of its sorted list, found once per pass by binary
//     search, so a pass reads exactly the core's edges and never tests
//     membership;
//   * a removed vertex gets its two live-degree counters poisoned (set far
//     below zero), so later decrements can never hit a threshold and the
//     per-edge work is one adjacency read plus one atomic -- no alive-flag
//     load;
//   * each level seeds from a compacted alive list instead of a scan over
//     all V, and the next level number jumps straight to the smallest live
//     key degree (levels that would remove nothing are skipped);
//   * frontier vertices are expanded edge-parallel: a tile of up to BLK
//     vertices is prefix-summed over its list lengths in LDS and every
//     thread takes one edge (load-balanced search), so hubs do not stall a
//     wavefront;
//   * frontier and alive-list appends are compacted with a 64-lane ballot
//     and popcount, one atomic per wavefront;
//   * phase counters are triple-buffered so each phase needs one barrier;
//   * a pass runs on a team of workgroups: one workgroup (workgroup
//     barriers, counters in LDS), or several with the bounded team barrier
//     of grid_sync.hpp; a team shrinks to its first workgroup once the alive
//     list is short, so the long tail of small levels costs no team
//     barriers.  Many passes share one persistent launch (k_rows_teams).
#include <hip/hip_runtime.h>

#include <climits>

#include "chain_device.hpp"
#include "dcx_internal.hpp"
#include "grid_sync.hpp"

namespace dcx {
namespace {

constexpr int W = 64;
constexpr int kBlk = 512;
constexpr int kUnroll = 4;
constexpr int kSeedUnroll = 4;
// core-layout ranks whose decrements a workgroup combines in LDS (per side);
// in the original-id layout, slots of the hashed table (keys, then counts)
constexpr uint32_t kCombRanks = 4096;
constexpr uint32_t kCombLog2 = 12;
constexpr uint32_t kCombEmpty = 0xffffffffu;
// edges in flight per thread in the LDS-resident pass (its atomics are
// cheap, so the list reads' latency is what remains to hide)
constexpr int kLdsUnroll = 8;
// packed bound of an LDS-resident member meaning "none below 2^16 - 1"
constexpr uint32_t kLdsUbNone = 0xffffu;
constexpr int32_t kPoison = INT_MIN / 2;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (W - 1)); }

// Loads and stores of pass state that other workgroups of a team read or
// wrote in earlier phases.  C (a multi-workgroup team): agent-coherent
// (atomic relaxed at agent scope), so the phase barriers need no L2
// write-back or invalidation; otherwise plain (one workgroup shares its
// L1).
template <bool C, typename T>
__device__ __forceinline__ T ld(const T* p) {
    if constexpr (C)
        return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *p;
}
template <bool C, typename T>
__device__ __forceinline__ void st(T* p, T v) {
    if constexpr (C)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

// Append `val` to arr[] for every lane with `pred`, one atomic per wavefront.
// Every lane of the wavefront must reach this call.
template <bool C = false>
__device__ __forceinline__ void wave_append(bool pred, uint32_t val, uint32_t* arr,
                                            uint32_t* ctr) {
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    const int leader = __ffsll(m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    base = __shfl(base, leader, W);
    if (pred) st<C>(&arr[base + __popcll(m & ((1ull << lane_id()) - 1ull))], val);
}

// wave_append that also adds the appended vertices' list lengths `len` to
// the 64-bit total *etot (one more atomic per wavefront that appends).
template <bool C = false>
__device__ __forceinline__ void wave_append_len(bool pred, uint32_t val, uint32_t len,
                                                uint32_t* arr, uint32_t* ctr,
                                                unsigned long long* etot) {
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    const int leader = __ffsll(m) - 1;
    uint32_t l = pred ? len : 0u;
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) l += (uint32_t)__shfl_xor(l, d, W);
    uint32_t base = 0;
    if (lane_id() == leader) {
        base = atomicAdd(ctr, (uint32_t)__popcll(m));
        consume((uint32_t)atomicAdd(etot, (unsigned long long)l));
    }
    base = __shfl(base, leader, W);
    if (pred) st<C>(&arr[base + __popcll(m & ((1ull << lane_id()) - 1ull))], val);
}

// Wavefront sum of x into the 64-bit total *dst (performed before returning).
__device__ __forceinline__ void wave_sum_to(uint32_t x, unsigned long long* dst) {
    unsigned long long y = x;
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) y += __shfl_xor(y, d, W);
    if (lane_id() == 0 && y) consume((uint32_t)atomicAdd(dst, y));
}

// Workgroup-aggregated appends of up to U values per thread to two lists
// (pa -> arra, pb -> arrb): one global atomic per list per call instead of
// one per wavefront.  Every thread of the workgroup must call it (it
// contains barriers); scr: LDS, 2 * (BLK / W) + 2 words.
template <int BLK, bool C, int U>
__device__ __forceinline__ void block_append2(const bool (&pa)[U], uint32_t* arra, uint32_t* ctra,
                                              const bool (&pb)[U], uint32_t* arrb, uint32_t* ctrb,
                                              const uint32_t (&val)[U], uint32_t* scr) {
    constexpr int NW = BLK / W;
    const int w = (int)(threadIdx.x / W), ln = lane_id();
    const unsigned long long lt = (1ull << ln) - 1ull;
    unsigned long long ma[U], mb[U];
    uint32_t ca = 0, cb = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        ma[u] = __ballot(pa[u]);
        mb[u] = __ballot(pb[u]);
        ca += (uint32_t)__popcll(ma[u]);
        cb += (uint32_t)__popcll(mb[u]);
    }
    if (ln == 0) {
        scr[w] = ca;
        scr[NW + w] = cb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ta = 0, tb = 0;
        for (int i = 0; i < NW; ++i) {
            const uint32_t x = scr[i], y = scr[NW + i];
            scr[i] = ta;
            scr[NW + i] = tb;
            ta += x;
            tb += y;
        }
        scr[2 * NW] = ta ? atomicAdd(ctra, ta) : 0u;
        scr[2 * NW + 1] = tb ? atomicAdd(ctrb, tb) : 0u;
    }
    __syncthreads();
    uint32_t oa = scr[2 * NW] + scr[w], ob = scr[2 * NW + 1] + scr[NW + w];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (pa[u]) st<C>(&arra[oa + __popcll(ma[u] & lt)], val[u]);
        oa += (uint32_t)__popcll(ma[u]);
        if (pb[u]) st<C>(&arrb[ob + __popcll(mb[u] & lt)], val[u]);
        ob += (uint32_t)__popcll(mb[u]);
    }
    __syncthreads();  // scr is reused by the next call
}

// wave_append into arr[0, cap): appends beyond cap are counted but dropped.
template <bool C>
__device__ __forceinline__ void wave_append_capped(bool pred, uint32_t val, uint32_t* arr,
                                                   uint32_t* ctr, uint32_t cap) {
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    const int leader = __ffsll(m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    base = __shfl(base, leader, W);
    const uint32_t idx = base + __popcll(m & ((1ull << lane_id()) - 1ull));
    if (pred && idx < cap) st<C>(&arr[idx], val);
}

// Frontier vertices whose lists hold at least kLongList edges are kept in a
// list of their own (fl, fl_cap entries; slot word 8 counts them), so that
// every member of a team can take a slice of each such list instead of one
// member walking it alone.  The rest, and long ones beyond fl_cap, go to G.
// Both count their list edges into *etot (the workgroup's LDS total, added
// to slot words 4..5 once per phase).
constexpr uint32_t kLongList = 4096;
// Appends go per workgroup (one atomic per list and workgroup) instead of
// per wavefront when kAggMembers or more workgroups append to the same
// counters: from kAggEdges frontier edges on in a push, always in a seed.
constexpr uint64_t kAggEdges = 1u << 16;
constexpr uint32_t kAggMembers = 64;
template <bool C>
__device__ __forceinline__ void append_frontier(bool pred, uint32_t v, uint32_t len, uint32_t* G,
                                                uint32_t* fl, uint32_t fl_cap, uint32_t* slot,
                                                unsigned long long* etot) {
    const bool lng = pred && len >= kLongList;
    bool placed = false;
    const unsigned long long m = __ballot(lng);
    if (m) {
        const int leader = __ffsll(m) - 1;
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(&slot[8], (uint32_t)__popcll(m));
        base = __shfl(base, leader, W);
        if (lng) {
            const uint32_t idx = base + __popcll(m & ((1ull << lane_id()) - 1ull));
            if (idx < fl_cap) {
                st<C>(&fl[idx], v);
                placed = true;
            }
        }
        wave_sum_to(placed ? len : 0u, etot);
    }
    wave_append_len<C>(pred && !placed, v, len, G, &slot[0], etot);
}

// Wavefront minimum, then one atomicMin per wavefront (performed before
// returning: its result is consumed, see coherent_write).
__device__ __forceinline__ void wave_min_to(uint32_t x, uint32_t* dst) {
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) x = min(x, (uint32_t)__shfl_xor(x, d, W));
    if (lane_id() == 0 && x != UINT_MAX) consume(atomicMin(dst, x));
}

__device__ __forceinline__ void wave_max_to(int32_t x, int32_t* dst) {
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) x = max(x, __shfl_xor(x, d, W));
    if (lane_id() == 0) consume((uint32_t)atomicMax(dst, x));
}

template <int BLK>
struct TileLds {
    uint32_t koff[BLK];
    uint32_t coff[BLK];
    uint32_t klen[BLK];
    uint32_t scan[BLK];
    uint32_t wsum[BLK / W];
    uint32_t total;
    uint32_t nitems;
};

// Block-wide exclusive scan; writes the total to *total.  Contains
// barriers: every thread of the block must call it.
template <int BLK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* wsum,
                                                    uint32_t* total) {
    constexpr int NW = BLK / W;
    const int ln = lane_id(), w = (int)(threadIdx.x / W);
    uint32_t inc = x;
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        uint32_t y = __shfl_up(inc, d, W);
        if (ln >= d) inc += y;
    }
    if (ln == W - 1) wsum[w] = inc;
    __syncthreads();
    if (w == 0) {
        uint32_t t = ln < NW ? wsum[ln] : 0u;
#pragma unroll
        for (int d = 1; d < NW; d <<= 1) {
            uint32_t y = __shfl_up(t, d, W);
            if (ln >= d) t += y;
        }
        if (ln < NW) wsum[ln] = t;
        if (ln == NW - 1) *total = t;
    }
    __syncthreads();
    return (w ? wsum[w - 1] : 0u) + inc - x;
}

// Item owning edge e: the largest i < nitems with scan[i] <= e.
template <int BLK>
__device__ __forceinline__ uint32_t tile_find(const TileLds<BLK>& T, uint32_t e) {
    uint32_t lo = 0, hi = T.nitems - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (T.scan[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// First index of a sorted list whose entry is >= bound.
__device__ __forceinline__ uint32_t lower_bound(const uint32_t* a, uint32_t n, uint32_t bound) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < bound)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Phase counters: three slots (triple-buffered by phase), each in its own
// 128-byte line, touched only with RMW atomics (coherent_read/_write of
// grid_sync.hpp).  Words: [0] frontier appends, [1] alive-list appends,
// [2] smallest key degree seen above the level, [3] the leader's phase,
// [4..5] list edges of the new frontier, [6..7] list edges of the alive
// vertices (first level only), [8] frontier vertices with long lists,
// [9] alive vertices with long lists (first level only).
// Called by threads 0..9 of one wavefront: each resets one word, so the
// resets are one round trip instead of a chain.
__device__ __forceinline__ void slot_reset(uint32_t* s, uint32_t lane) {
    if (lane < 10 && lane != 3) coherent_write(s + lane, lane == 2 ? (uint32_t)UINT_MAX : 0u);
}

__device__ __forceinline__ unsigned long long* slot_u64(uint32_t* s, int w) {
    return reinterpret_cast<unsigned long long*>(s + w);
}

struct SlotView {
    uint32_t nF, nA, mn, nL, nLA;
    uint64_t fedges, aedges;
};

// One coherent read of a slot per workgroup, shared through LDS; every
// thread must call it (contains workgroup barriers).
__device__ __forceinline__ SlotView slot_read(uint32_t* s, uint32_t* lds) {
    // nine lanes read one word each (one round trip)
    if (threadIdx.x < 9) {
        const uint32_t w = threadIdx.x < 3 ? threadIdx.x : threadIdx.x + 1;
        lds[threadIdx.x] = coherent_read(s + w);
    }
    __syncthreads();
    SlotView v;
    v.nF = lds[0];
    v.nA = lds[1];
    v.mn = lds[2];
    v.fedges = (uint64_t)lds[3] | ((uint64_t)lds[4] << 32);
    v.aedges = (uint64_t)lds[5] | ((uint64_t)lds[6] << 32);
    v.nL = lds[7];
    v.nLA = lds[8];
    __syncthreads();
    return v;
}

__device__ __forceinline__ void store_agent(int32_t* p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Marks x removed in phase `ph`: both counters poisoned, removal phase
// recorded for the pull phases' frontier test.
__device__ __forceinline__ void poison(const PassScratch& S, uint32_t x, uint32_t ph) {
    store_agent(S.kd + x, kPoison);
    store_agent(S.cd + x, kPoison);
    __hip_atomic_store(S.st + x, ph, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lengths of v's K and C lists inside the pass (member prefixes in the core
// layout, whole lists otherwise).
__device__ __forceinline__ void list_lens(const PassSpec& P, const PassScratch& S, uint32_t v,
                                          uint32_t& kl, uint32_t& cl) {
    kl = P.core ? S.rk[v] : (uint32_t)P.K.deg[v];
    cl = P.core ? S.rc[v] : (uint32_t)P.C.deg[v];
}

// One full peel pass.  MULTI: a team of workgroups (gs != nullptr; `slots`
// in global memory); otherwise the calling workgroup alone (`slots` may be
// in LDS).  A team narrows as its alive list shrinks: at the start of a
// level it keeps ceil(alive / shrink_at) members (once that is at most
// half of them, or one), so its barriers stay as cheap as the remaining
// work allows; the others return early and the caller joins them again
// (team_resync) before the scratch is reused.  Logic errors
// (counters that do not add up) are reported into *err (single workgroup)
// or the team's error word instead of looping.
// Decrement combiners of a push phase (LDS, per workgroup): none; by rank
// (the first kCombRanks core ranks, core layout only); hashed (any ids).
enum : int { kCombNone = 0, kCombByRank = 1, kCombHashed = 2 };

template <int BLK, bool MULTI, int COMB>
__device__ void run_pass(const PassSpec& P, const PassScratch& S, uint32_t* slots,
                         TileLds<BLK>& T, int32_t* lmax_out, GridSync* gs, uint32_t* err,
                         uint32_t shrink_at, uint32_t* comb) {
    const uint32_t tid = threadIdx.x;
    uint32_t tw = MULTI ? gs->rank : 0u;
    uint32_t tn = MULTI ? gs->nblocks : 1u;
    bool solo = !MULTI || tn == 1;
    // pass state (declared first: the lambdas below read it)
    uint32_t removed = 0, s = 0, levels = 0, nA = 0, mnall_dbg = 0;
    int32_t L = 0;
    // Phase barriers are unfenced: everything the team hands over between
    // phases goes through agent-coherent accesses (ld/st<MULTI>, atomics).
    // The set-up barrier and the last one are fenced (plain stores of the
    // set-up and of the output rows).
    auto barrier = [&](bool fenced = false) -> bool {
        if (MULTI && !solo) {
            if (tw == 0 && tid == 0) coherent_write(slots + 3, s);
            return grid_barrier(*gs, fenced);
        }
        __syncthreads();
        return true;
    };
    bool failed = false;
    auto fail = [&](uint32_t code) {
        failed = true;
        if (tid == 0 && P.fail_rec && atomicCAS(P.fail_rec, 0u, 1u) == 0u) {
            P.fail_rec[1] = tw;
            P.fail_rec[2] = tn;
            P.fail_rec[3] = s;
            P.fail_rec[4] = levels;
            P.fail_rec[5] = removed;
            P.fail_rec[6] = P.nmembers;
            P.fail_rec[7] = nA;
            P.fail_rec[8] = mnall_dbg;
            P.fail_rec[9] = (uint32_t)L;
            P.fail_rec[10] = (uint32_t)P.k;
            P.fail_rec[11] = MULTI ? gs->rank : 0u;
            P.fail_rec[12] = MULTI ? gs->epoch : 0u;
            P.fail_rec[13] = coherent_read(slots + 3);  // leader's last phase
            P.fail_rec[14] = MULTI ? coherent_read(gs->w + kSyncRelease) : 0u;
            P.fail_rec[15] = MULTI ? gs->nblocks : 0u;
        }
        if (tid == 0) {
            if (MULTI)
                sync_fail(*gs, code);
            else
                consume(atomicMax(err, code));
        }
    };
    int32_t* const lval = P.core ? S.lv : P.lval;
    const uint64_t t_start = wall_clock64();
    uint64_t t_prev = t_start;
    // phase trace, written by the team leader after each barrier
    auto trace = [&](uint32_t phase, bool seed, bool pull, int32_t lvl, uint32_t nfr,
                     uint32_t nal) {
        if (!P.trace || tw != 0 || tid != 0) return;
        const uint64_t now = wall_clock64();
        if (phase < P.trace_cap) {
            uint32_t* e = P.trace + kTraceWords * (1 + (size_t)phase);
            e[kTraceLevel] = ((uint32_t)lvl & kTraceLevelMask) | (seed ? kTraceSeedBit : 0u) |
                             (pull ? kTracePullBit : 0u);
            e[kTraceFrontier] = nfr;
            e[kTraceAlive] = nal;
            e[kTraceTicks] = (uint32_t)min<uint64_t>(now - t_prev, 0xffffffffull);
        }
        P.trace[0] = min(phase + 1, P.trace_cap);
        t_prev = now;
    };

    __shared__ uint32_t s_rd[12];
    __shared__ uint32_t s_part[2];
    __shared__ uint32_t s_scr[2 * (BLK / W) + 2];
    // list-edge totals of the phase, summed in LDS and added to the slot
    // once per workgroup (per-wavefront global atomics on two words
    // serialised big phases)
    __shared__ unsigned long long s_etot, s_atot;
    if (tid == 0) s_etot = s_atot = 0;
    auto flush_totals = [&](uint32_t* slot) {
        __syncthreads();
        if (tid == 0) {
            if (s_etot) consume((uint32_t)atomicAdd(slot_u64(slot, 4), s_etot));
            if (s_atot) consume((uint32_t)atomicAdd(slot_u64(slot, 6), s_atot));
            s_etot = s_atot = 0;
        }
    };
    if (tw == 0) {
        slot_reset(slots, tid);
        slot_reset(slots + kSlotStride, tid);
        slot_reset(slots + 2 * kSlotStride, tid);
    }

    // ---- set-up: live degrees (member prefixes in the core layout)
    const uint32_t nm = P.nmembers;
    for (uint32_t i = tw * BLK + tid; i < nm; i += tn * BLK) {
        if (P.core) {
            const uint32_t rk = lower_bound(P.K.nbr + P.K.off[i], (uint32_t)P.K.deg[i], nm);
            const uint32_t rc = lower_bound(P.C.nbr + P.C.off[i], (uint32_t)P.C.deg[i], nm);
            S.rk[i] = rk;
            S.rc[i] = rc;
            S.kd[i] = (int32_t)rk;
            S.cd[i] = (int32_t)rc;
        } else {
            S.kd[i] = P.K.deg[i];
            S.cd[i] = P.C.deg[i];
        }
        lval[i] = -1;
        S.st[i] = ~0u;
    }
    // long alive list (first quarter) and its partial counts (next two)
    const uint32_t la_cap = nm / 4;
    uint32_t* const LA = S.lp;
    uint32_t* const PK = S.lp + la_cap;
    uint32_t* const PC = S.lp + 2 * la_cap;
    for (uint32_t i = tw * BLK + tid; i < 2 * la_cap; i += tn * BLK) PK[i] = 0;
    if (!barrier(true)) return;

    const uint32_t* A = nullptr;  // first level: the members in id order
    uint32_t *B = S.a0, *Bn = S.a1, *F = S.f0, *G = S.f1;
    // long-list frontier vertices: FL (current) and GL (next), nL of them
    const uint32_t fl_cap = nm / 2;
    uint32_t *FL = S.fl, *GL = S.fl + fl_cap, nL = 0;
    nA = nm;
    const int32_t k = P.k;
    // Without membership restriction the constraint can fail before any
    // removal (KList semantics on the whole graph, main.cu:144-169).
    const bool check_c = !P.core && k > 0;
    // List edges of the alive vertices and of the current frontier: every
    // member of the team derives the same values from the slots, so all of
    // them choose the same kind of cascade phase.
    uint64_t pa = 0, fe = 0;
    uint32_t nLA = 0;  // long alive list length (0 when it overflowed)
    const int32_t* const ub = P.ub;
    while (removed < nm) {
        // Each removal needs at most two level iterations (a level number
        // taken from a stale minimum seeds nothing, the next one seeds the
        // true minimum), so more than 2*nm+2 means the counters are wrong.
        if (++levels > 2 * nm + 2) {
            fail(kErrInconsistent);
            break;
        }
        if (MULTI && !solo && shrink_at > 0) {
            // members worth keeping: one per shrink_at alive vertices or per
            // 16 * shrink_at edges of their lists, whichever asks for more
            // (a dense core has few vertices but long lists)
            const uint64_t by_v = ((uint64_t)nA + shrink_at - 1) / shrink_at;
            const uint64_t by_e = (pa + 16ull * shrink_at - 1) / (16ull * shrink_at);
            const uint32_t want = (uint32_t)max<uint64_t>(1, min<uint64_t>(tn, max(by_v, by_e)));
            if (want == 1 || 2 * want <= tn) {
                // The narrowed team resets phase slots without the members
                // that leave; they must all have finished reading the last
                // slot first, so the whole team meets once more here.
                if (!barrier()) return;
                if (tw >= want) return;
                tn = want;
                gs->nblocks = want;
                solo = want == 1;
            }
        }
        // ---- seed level L from the alive list; survivors go to B
        uint32_t* cur = slots + kSlotStride * (s % 3);
        if (tw == 0) slot_reset(slots + kSlotStride * ((s + 1) % 3), tid);
        uint32_t mn = UINT_MAX;
        const bool first = A == nullptr;
        const bool sagg = tn >= kAggMembers;  // the same in every member
        // kSeedUnroll independent list entries per thread per step
        for (uint32_t base = tw * BLK * kSeedUnroll; base < nA; base += tn * BLK * kSeedUnroll) {
            uint32_t v[kSeedUnroll];
            bool has[kSeedUnroll];
#pragma unroll
            for (int u = 0; u < kSeedUnroll; ++u) {
                const uint32_t i = base + u * BLK + tid;
                has[u] = i < nA;
                v[u] = has[u] ? (A ? ld<MULTI>(&A[i]) : i) : 0u;
            }
            int32_t d[kSeedUnroll];
            bool alive[kSeedUnroll];
#pragma unroll
            for (int u = 0; u < kSeedUnroll; ++u) {
                alive[u] = has[u] && ld<MULTI>(&lval[v[u]]) == -1;
                d[u] = alive[u] ? ld<MULTI>(&S.kd[v[u]]) : 0;
                // the bound acts like a key degree that never drops
                if (ub && alive[u]) d[u] = min(d[u], ub[v[u]]);
            }
            bool to_f[kSeedUnroll], to_b[kSeedUnroll];
#pragma unroll
            for (int u = 0; u < kSeedUnroll; ++u) {
                const bool seed =
                    alive[u] && (d[u] <= L || (check_c && ld<MULTI>(&S.cd[v[u]]) < k));
                const bool keep = alive[u] && !seed;
                uint32_t len = 0;
                if (seed || (first && keep)) {
                    uint32_t kl, cl;
                    list_lens(P, S, v[u], kl, cl);
                    len = kl + cl;
                }
                if (seed) {
                    st<MULTI>(&lval[v[u]], L);
                    poison(S, v[u], s);
                }
                if (keep) mn = min(mn, (uint32_t)d[u]);
                // kept vertices with long lists: their pulls are sliced
                if (first) wave_append_capped<MULTI>(keep && len >= kLongList, v[u], LA, &cur[9], la_cap);
                // seeds with long lists go to their own list, the rest and
                // the survivors are appended per workgroup below
                append_frontier<MULTI>(seed && (!sagg || len >= kLongList), v[u], len, F, FL,
                                       fl_cap, cur, &s_etot);
                to_f[u] = sagg && seed && len < kLongList;
                to_b[u] = sagg && keep;
                if (!sagg)
                    wave_append<MULTI>(keep, v[u], B, &cur[1]);
                else
                    wave_sum_to(to_f[u] ? len : 0u, &s_etot);
                if (first) wave_sum_to(keep ? len : 0u, &s_atot);
            }
            if (sagg)
                block_append2<BLK, MULTI, kSeedUnroll>(to_f, F, &cur[0], to_b, B, &cur[1], v,
                                                       s_scr);
        }
        wave_min_to(mn, &cur[2]);
        flush_totals(cur);
        if (!barrier()) return;
        const SlotView sv = slot_read(cur, s_rd);
        uint32_t nF = sv.nF;
        if (first) nLA = sv.nLA <= la_cap ? sv.nLA : 0u;  // overflow: no slicing
        nL = min(sv.nL, fl_cap);
        nA = sv.nA;
        uint32_t mnall = sv.mn;
        fe = sv.fedges;
        pa = first ? sv.aedges : pa - fe;

        trace(s, true, false, L, nF + nL, nA);
        A = B;
        {
            uint32_t* t = B;
            B = Bn;
            Bn = t;
        }
        ++s;
        removed += nF + nL;

        // ---- cascade inside level L until no vertex crosses a threshold
        // (none when the seeds took every alive vertex: their decrements
        // would only reach removed ones -- at the last level of a dense
        // core that is millions of list entries read for nothing)
        if (nA == 0) nF = nL = 0;
        while (nF + nL > 0) {
            uint32_t* nxt = slots + kSlotStride * (s % 3);
            if (tw == 0) slot_reset(slots + kSlotStride * ((s + 1) % 3), tid);
            uint32_t mn2 = UINT_MAX;
            // the same in every member
            const bool agg = fe >= kAggEdges && tn >= kAggMembers;
            // kUnroll edges held by this thread: x[u] loses a neighbour on
            // the constraint side (ks[u]) or the key side; ok[u] = a real edge
            auto visit = [&](uint32_t(&x)[kUnroll], bool(&ks)[kUnroll], bool(&ok)[kUnroll]) {
                bool claim[kUnroll];
                // Half of all edge visits reach a neighbour that
                // left earlier (every edge is visited from both
                // ends); a plain load of its level skips those
                // atomics.  A stale -1 only costs the atomic, which
                // then finds a poisoned counter.
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) ok[u] = ok[u] && lval[x[u]] == -1;
                if (COMB == kCombByRank) {
                    // the first kCombRanks ranks are the densest
                    // core's vertices, which most frontier edges hit:
                    // their decrements are summed in LDS and applied
                    // once per workgroup and phase (flush below)
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u)
                        if (ok[u] && x[u] < kCombRanks) {
                            atomicAdd(&comb[(ks[u] ? kCombRanks : 0u) + x[u]], 1u);
                            ok[u] = false;
                        }
                } else if (COMB == kCombHashed) {
                    // any layout: a hashed table of
                    // (vertex, side) keys, first come first served; a hub
                    // hit by thousands of frontier edges takes a slot early
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u)
                        if (ok[u]) {
                            const uint32_t key = (x[u] << 1) | (ks[u] ? 1u : 0u);
                            const uint32_t h = (key * 2654435761u) >> (32 - kCombLog2);
                            uint32_t cur = comb[h];
                            if (cur == kCombEmpty) {
                                cur = atomicCAS(&comb[h], kCombEmpty, key);
                                if (cur == kCombEmpty) cur = key;
                            }
                            if (cur == key) {
                                atomicAdd(&comb[kCombRanks + h], 1u);
                                ok[u] = false;
                            }
                        }
                }
                int32_t old[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    // K-side edge: x loses a C-neighbour (constraint
                    // degree); C-side edge: x loses a K-neighbour (key)
                    old[u] = ok[u] ? atomicSub(ks[u] ? &S.cd[x[u]] : &S.kd[x[u]], 1) : 0;
                }
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    claim[u] = false;
                    if (ok[u]) {
                        if (ks[u]) {
                            claim[u] = k > 0 && old[u] == k;
                        } else if (old[u] == L + 1) {
                            claim[u] = true;
                        } else if (old[u] - 1 > L) {
                            mn2 = min(mn2, (uint32_t)(old[u] - 1));
                        }
                    }
                    if (claim[u]) {
                        claim[u] = atomicCAS(&lval[x[u]], -1, L) == -1;
                        if (claim[u]) poison(S, x[u], s);
                    }
                }
                bool to_g[kUnroll], none[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    uint32_t len = 0;
                    if (claim[u]) {
                        uint32_t a, b;
                        list_lens(P, S, x[u], a, b);
                        len = a + b;
                    }
                    // a big frontier claims many vertices per step: those
                    // are appended per workgroup; a small one per wavefront
                    // (no extra barriers on the latency-bound path)
                    append_frontier<MULTI>(claim[u] && (!agg || len >= kLongList), x[u], len, G,
                                           GL, fl_cap, nxt, &s_etot);
                    to_g[u] = agg && claim[u] && len < kLongList;
                    none[u] = false;
                    if (agg) wave_sum_to(to_g[u] ? len : 0u, &s_etot);
                }
                if (agg)
                    block_append2<BLK, MULTI, kUnroll>(to_g, G, &nxt[0], none, G, &nxt[0], x,
                                                       s_scr);
            };
            // a pull reads every alive vertex's lists and every entry of
            // the level's alive list (most of which may have left by now)
            const bool pull = P.pull_ratio != 0 && pa + nA <= (uint64_t)P.pull_ratio * fe;
            if (pull) {
                // Pull: every alive vertex of the level's list counts its
                // neighbours removed in the previous phase and updates its
                // own counters -- no atomics on shared counters, which is
                // what a dense frontier (many edges into few vertices)
                // needs.  The counts equal the decrements a push would
                // apply, so the claims and the minimum are the same.  One
                // wavefront per alive vertex: coalesced list reads, the
                // count is a popcount of ballots.
                const uint32_t sprev = s - 1;
                const uint32_t nw = BLK / W;
                const uint32_t lane = (uint32_t)lane_id();
                for (uint32_t i = tw * nw + tid / W; i < nA; i += tn * nw) {
                    const uint32_t x = ld<MULTI>(&A[i]);
                    if (ld<MULTI>(&lval[x]) != -1) continue;  // removed earlier in this level
                    uint32_t kl, cl;
                    list_lens(P, S, x, kl, cl);
                    if (nLA && kl + cl >= kLongList) continue;  // sliced below
                    const uint32_t* kn = P.K.nbr + P.K.off[x];
                    const uint32_t* cn = P.C.nbr + P.C.off[x];
                    uint32_t ck = 0, cc = 0;  // the same in every lane
                    for (uint32_t jb = 0; jb < kl + cl; jb += kUnroll * W) {
                        bool hit[kUnroll], isk[kUnroll];
#pragma unroll
                        for (int u = 0; u < kUnroll; ++u) {
                            const uint32_t e = jb + u * W + lane;
                            isk[u] = e < kl;
                            const bool ok = e < kl + cl;
                            const uint32_t y = !ok ? 0u : isk[u] ? kn[e] : cn[e - kl];
                            hit[u] = ok && ld<MULTI>(&S.st[y]) == sprev;
                        }
#pragma unroll
                        for (int u = 0; u < kUnroll; ++u) {
                            ck += (uint32_t)__popcll(__ballot(hit[u] && isk[u]));
                            cc += (uint32_t)__popcll(__ballot(hit[u] && !isk[u]));
                        }
                    }
                    bool claim = false;
                    if ((ck | cc) && lane == 0) {
                        // x's own counters: nothing else writes them in a
                        // pull phase
                        const int32_t nk = ld<MULTI>(&S.kd[x]) - (int32_t)ck;
                        const int32_t nc = ld<MULTI>(&S.cd[x]) - (int32_t)cc;
                        claim = nk <= L || (k > 0 && nc < k);
                        if (claim) {
                            st<MULTI>(&lval[x], L);
                            poison(S, x, s);
                        } else {
                            if (ck) {
                                store_agent(S.kd + x, nk);
                                mn2 = min(mn2, (uint32_t)nk);
                            }
                            if (cc) store_agent(S.cd + x, nc);
                        }
                    }
                    append_frontier<MULTI>(claim, x, kl + cl, G, GL, fl_cap, nxt, &s_etot);
                }
                if (nLA) {
                    // long alive lists: every member counts the hits in its
                    // slice of each and adds them to the list's partial
                    // counts; after a team barrier one wavefront per list
                    // applies them
                    for (uint32_t li = 0; li < nLA; ++li) {
                        const uint32_t x = ld<MULTI>(&LA[li]);
                        if (ld<MULTI>(&lval[x]) != -1) continue;  // the same for the workgroup
                        uint32_t kl, cl;
                        list_lens(P, S, x, kl, cl);
                        const uint32_t tot = kl + cl;
                        const uint32_t lo = (uint32_t)((uint64_t)tot * tw / tn);
                        const uint32_t hi = (uint32_t)((uint64_t)tot * (tw + 1) / tn);
                        const uint32_t* kn = P.K.nbr + P.K.off[x];
                        const uint32_t* cn = P.C.nbr + P.C.off[x];
                        uint32_t ck = 0, cc = 0;
                        for (uint32_t e = lo + tid; e < hi; e += BLK) {
                            const bool isk = e < kl;
                            const uint32_t y = isk ? kn[e] : cn[e - kl];
                            if (ld<MULTI>(&S.st[y]) == sprev) {
                                if (isk)
                                    ++ck;
                                else
                                    ++cc;
                            }
                        }
#pragma unroll
                        for (int d = W / 2; d > 0; d >>= 1) {
                            ck += (uint32_t)__shfl_xor(ck, d, W);
                            cc += (uint32_t)__shfl_xor(cc, d, W);
                        }
                        if (tid == 0) s_part[0] = s_part[1] = 0;
                        __syncthreads();
                        if (lane == 0 && (ck | cc)) {
                            atomicAdd(&s_part[0], ck);
                            atomicAdd(&s_part[1], cc);
                        }
                        __syncthreads();
                        if (tid == 0) {
                            if (s_part[0]) consume(atomicAdd(&PK[li], s_part[0]));
                            if (s_part[1]) consume(atomicAdd(&PC[li], s_part[1]));
                        }
                        __syncthreads();
                    }
                    if (!barrier()) return;
                    for (uint32_t li = tw * nw + tid / W; li < nLA; li += tn * nw) {
                        const uint32_t x = ld<MULTI>(&LA[li]);
                        bool claim = false;
                        uint32_t len = 0;
                        if (ld<MULTI>(&lval[x]) == -1 && lane == 0) {
                            // read and clear the partial counts
                            const uint32_t ck = atomicAnd(&PK[li], 0u);
                            const uint32_t cc = atomicAnd(&PC[li], 0u);
                            if (ck | cc) {
                                const int32_t nk = ld<MULTI>(&S.kd[x]) - (int32_t)ck;
                                const int32_t nc = ld<MULTI>(&S.cd[x]) - (int32_t)cc;
                                claim = nk <= L || (k > 0 && nc < k);
                                if (claim) {
                                    st<MULTI>(&lval[x], L);
                                    poison(S, x, s);
                                    uint32_t kl, cl;
                                    list_lens(P, S, x, kl, cl);
                                    len = kl + cl;
                                } else {
                                    if (ck) {
                                        store_agent(S.kd + x, nk);
                                        mn2 = min(mn2, (uint32_t)nk);
                                    }
                                    if (cc) store_agent(S.cd + x, nc);
                                }
                            }
                        }
                        append_frontier<MULTI>(claim, x, len, G, GL, fl_cap, nxt, &s_etot);
                    }
                }
            } else {
                // Push: frontier items per workgroup -- a team splits even
                // a small frontier over all its members (whole tiles would
                // leave most of them idle while a few walk long edge lists).
                const uint32_t ipw =
                    solo ? (uint32_t)BLK : min((uint32_t)BLK, (nF + tn - 1) / tn);
                // one tile: up to ipw frontier vertices (`has`, v per
                // thread, `items` of them), their lists edge-parallel; of
                // the tile's edges this workgroup takes share `part` of
                // `parts` (a tile of long lists is split over the team)
                auto push_tile = [&](bool has, uint32_t v, uint32_t items, uint32_t part,
                                     uint32_t parts) {
                    __syncthreads();
                    uint32_t kl = 0, cl = 0;
                    if (has) {
                        T.koff[tid] = P.K.off[v];
                        T.coff[tid] = P.C.off[v];
                        list_lens(P, S, v, kl, cl);
                    }
                    T.klen[tid] = kl;
                    if (tid == 0) T.nitems = items;
                    T.scan[tid] = block_excl_scan<BLK>(kl + cl, T.wsum, &T.total);
                    // every thread's prefix is in LDS before any thread searches it
                    __syncthreads();
                    const uint32_t total = T.total;
                    const uint32_t elo = (uint32_t)((uint64_t)total * part / parts);
                    const uint32_t ehi = (uint32_t)((uint64_t)total * (part + 1) / parts);
                    // kUnroll independent edges per thread per step, so their
                    // loads and atomics overlap instead of forming one chain.
                    for (uint32_t eb = elo; eb < ehi; eb += kUnroll * BLK) {
                        uint32_t x[kUnroll];
                        bool ks[kUnroll], ok[kUnroll];
#pragma unroll
                        for (int u = 0; u < kUnroll; ++u) {
                            const uint32_t e = eb + u * BLK + tid;
                            ok[u] = e < ehi;
                            x[u] = 0;
                            ks[u] = false;
                            if (ok[u]) {
                                const uint32_t it = tile_find<BLK>(T, e);
                                const uint32_t t = e - T.scan[it];
                                ks[u] = t < T.klen[it];
                                x[u] = ks[u] ? P.K.nbr[T.koff[it] + t]
                                             : P.C.nbr[T.coff[it] + (t - T.klen[it])];
                            }
                        }
                        visit(x, ks, ok);
                    }
                };
                for (uint32_t base = tw * ipw; base < nF; base += tn * ipw) {
                    const uint32_t i = base + tid;
                    const bool has = tid < ipw && i < nF;
                    push_tile(has, has ? ld<MULTI>(&F[i]) : 0u, min(nF - base, ipw), 0u, 1u);
                }
                // long lists, BLK at a time as one tile whose edges the
                // members split evenly: the team walks many long lists at
                // once instead of one list after another
                for (uint32_t lb = 0; lb < nL; lb += BLK) {
                    const uint32_t i = lb + tid;
                    const bool has = i < nL;
                    push_tile(has, has ? ld<MULTI>(&FL[i]) : 0u, min(nL - lb, (uint32_t)BLK), tw,
                              tn);
                }
            }
            if (COMB != kCombNone && !pull) {
                // apply the combined decrements: one atomic per touched
                // counter; the atomic whose subtraction takes the counter
                // across its threshold claims the vertex
                __syncthreads();
                for (uint32_t j = tid; j < 2 * kCombRanks; j += BLK) {
                    uint32_t c, x;
                    bool cside;
                    if (COMB == kCombByRank) {
                        c = comb[j];
                        cside = j >= kCombRanks;
                        x = cside ? j - kCombRanks : j;
                        if (c) comb[j] = 0;
                    } else {
                        // the first half of the table holds keys, the
                        // second their counts
                        const uint32_t key = j < kCombRanks ? comb[j] : kCombEmpty;
                        c = key != kCombEmpty ? comb[kCombRanks + j] : 0u;
                        cside = key & 1u;
                        x = key >> 1;
                        if (key != kCombEmpty) {
                            comb[j] = kCombEmpty;
                            comb[kCombRanks + j] = 0;
                        }
                    }
                    bool claim = false;
                    if (c) {
                        const int32_t old = atomicSub(cside ? &S.cd[x] : &S.kd[x], (int32_t)c);
                        const int32_t nw = old - (int32_t)c;
                        if (cside) {
                            claim = k > 0 && old >= k && nw < k;
                        } else if (old >= L + 1 && nw <= L) {
                            claim = true;
                        } else if (nw > L) {
                            mn2 = min(mn2, (uint32_t)nw);
                        }
                        if (claim) {
                            claim = atomicCAS(&lval[x], -1, L) == -1;
                            if (claim) poison(S, x, s);
                        }
                    }
                    uint32_t len = 0;
                    if (claim) {
                        uint32_t a, b;
                        list_lens(P, S, x, a, b);
                        len = a + b;
                    }
                    append_frontier<MULTI>(claim, x, len, G, GL, fl_cap, nxt, &s_etot);
                }
            }
            wave_min_to(mn2, &nxt[2]);
            flush_totals(nxt);
            if (!barrier()) return;
            const SlotView sv2 = slot_read(nxt, s_rd);
            nF = sv2.nF;
            nL = min(sv2.nL, fl_cap);
            mnall = min(mnall, sv2.mn);
            fe = sv2.fedges;
            pa -= fe;

            trace(s, false, pull, L, nF + nL, nA);
            {
                uint32_t* t = F;
                F = G;
                G = t;
                t = FL;
                FL = GL;
                GL = t;
            }
            ++s;
            removed += nF + nL;
        }
        if (removed < nm) {
            // alive vertices exist, so some degree was recorded
            mnall_dbg = mnall;
            if (nA == 0 || mnall == UINT_MAX) {
                fail(kErrInconsistent);
                break;
            }
            L = max(L + 1, (int32_t)min(mnall, (uint32_t)INT_MAX));
        }
    }
    if (tw == 0 && tid == 0) {
        if (lmax_out) *lmax_out = nm ? L : 0;
        if (P.stats) {
            const uint64_t dt = wall_clock64() - t_start;
            P.stats[kStatPhases] = s;
            P.stats[kStatLevels] = levels;
            P.stats[kStatTicksLo] = (uint32_t)dt;
            P.stats[kStatTicksHi] = (uint32_t)(dt >> 32);
            P.stats[kStatFailed] = failed ? 1u : 0u;
            P.stats[kStatRemoved] = removed;
            P.stats[kStatAlive] = nA;
            P.stats[kStatLevel] = (uint32_t)L;
        }
    }
    if (P.core) {
        // levels back to original ids (rows hold 0 for non-members already)
        for (uint32_t i = tw * BLK + tid; i < nm; i += tn * BLK)
            P.out_row[P.order[i]] = ld<MULTI>(&S.lv[i]);
    }
    // the whole (narrowed) team is done with the scratch before rank 0
    // hands the team on (team_resync)
    if (MULTI && !solo) barrier(true);
}

// ---- LDS-resident pass for small cores -------------------------------
//
// A row whose (k,0)-core has at most kLdsMaxMembers members runs on one
// workgroup with both live-degree counters of every member packed into one
// LDS word (key degree in the low half, constraint degree in the high
// half) and a removed-bit per member, so its per-edge atomics never leave
// the CU.  Counts inside a core of n members are at most n - 1 < 2^15, and
// a counter is decremented once per removed neighbour (decrements to
// removed members are skipped or, when racing with the removal, still
// bounded by the live count), so neither half ever borrows or overflows.
// Levels, frontiers and alive lists stay in the pass scratch (global).

__device__ __forceinline__ bool lds_removed(const uint32_t* rmv, uint32_t x) {
    return (rmv[x >> 5] >> (x & 31)) & 1u;
}

// One frontier edge of the LDS-resident pass: x loses a neighbour, on the
// constraint side (high half) when ks, else on the key side.  Returns true
// when this call claims x (at most one call per vertex does).
__device__ __forceinline__ bool lds_edge(uint32_t* cnt, uint32_t* rmv, uint32_t x, bool ks,
                                         int32_t k, int32_t L, uint32_t& mn2) {
    if (lds_removed(rmv, x)) return false;
    bool claim = false;
    if (ks) {
        const uint32_t old = atomicSub(&cnt[x], 0x10000u);
        claim = k > 0 && (int32_t)(old >> 16) == k;
    } else {
        const int32_t old = (int32_t)(atomicSub(&cnt[x], 1u) & 0xffffu);
        if (old == L + 1)
            claim = true;
        else if (old - 1 > L)
            mn2 = min(mn2, (uint32_t)(old - 1));
    }
    if (claim) {
        const uint32_t bit = 1u << (x & 31);
        claim = !(atomicOr(&rmv[x >> 5], bit) & bit);
    }
    return claim;
}

template <int BLK>
__device__ void run_pass_lds(const PassSpec& P, const PassScratch& S, uint32_t* slots,
                             TileLds<BLK>& T, uint32_t* cnt, uint32_t* rmv, int32_t* lmax_out,
                             uint32_t* err) {
    const uint32_t tid = threadIdx.x;
    const uint32_t nm = P.nmembers;
    const uint32_t nwords = (nm + 31) / 32;
    const uint64_t t_start = wall_clock64();
    for (uint32_t i = tid; i < nm; i += BLK) {
        const uint32_t rk = lower_bound(P.K.nbr + P.K.off[i], (uint32_t)P.K.deg[i], nm);
        const uint32_t rc = lower_bound(P.C.nbr + P.C.off[i], (uint32_t)P.C.deg[i], nm);
        S.rk[i] = rk;
        S.rc[i] = rc;
        cnt[i] = rk | (rc << 16);
        S.lv[i] = -1;
    }
    for (uint32_t i = tid; i < nwords; i += BLK) rmv[i] = 0;
    if (tid < 2) {  // both phase slots: no appends yet, no minimum
        slots[kSlotStride * tid + 0] = 0;
        slots[kSlotStride * tid + 1] = 0;
        slots[kSlotStride * tid + 2] = UINT_MAX;
        *slot_u64(slots + kSlotStride * tid, 4) = 0;
    }
    __syncthreads();

    const uint32_t* A = nullptr;
    uint32_t *B = S.a0, *Bn = S.a1, *F = S.f0, *G = S.f1;
    uint32_t nA = nm, removed = 0, s = 0, levels = 0;
    int32_t L = 0;
    const int32_t k = P.k;
    const int32_t* const ub = P.ub;
    bool failed = false;
    uint64_t t_prev = t_start;
    auto trace = [&](bool seed, uint32_t nfr) {
        if (!P.trace || tid != 0) return;
        const uint64_t now = wall_clock64();
        if (s < P.trace_cap) {
            uint32_t* e = P.trace + kTraceWords * (1 + (size_t)s);
            e[kTraceLevel] = ((uint32_t)L & kTraceLevelMask) | (seed ? kTraceSeedBit : 0u);
            e[kTraceFrontier] = nfr;
            e[kTraceAlive] = nA;
            e[kTraceTicks] = (uint32_t)min<uint64_t>(now - t_prev, 0xffffffffull);
        }
        P.trace[0] = min(s + 1, P.trace_cap);
        t_prev = now;
    };
    while (removed < nm) {
        if (++levels > 2 * nm + 2) {
            failed = true;
            break;
        }
        uint32_t* cur = slots + kSlotStride * (s & 1);
        uint32_t* other = slots + kSlotStride * ((s + 1) & 1);
        if (tid == 0) {
            other[0] = 0;
            other[1] = 0;
            other[2] = UINT_MAX;
            *slot_u64(other, 4) = 0;
        }
        uint32_t mn = UINT_MAX;
        // Alive-list entries carry the member's bound in their high half
        // (members are < 2^15; kLdsUbNone: no bound below 2^16 - 1), so
        // later levels read it with the entry instead of from the bound
        // array.
        for (uint32_t base = 0; base < nA; base += BLK * kSeedUnroll) {
            uint32_t v[kSeedUnroll], b[kSeedUnroll];
            bool has[kSeedUnroll];
#pragma unroll
            for (int u = 0; u < kSeedUnroll; ++u) {
                const uint32_t i = base + u * BLK + tid;
                has[u] = i < nA;
                uint32_t e = 0;
                if (has[u]) {
                    if (A)
                        e = A[i];
                    else
                        e = i | (ub ? (uint32_t)min(ub[i], (int32_t)kLdsUbNone) : kLdsUbNone) << 16;
                }
                v[u] = e & 0xffffu;
                b[u] = e >> 16;
            }
#pragma unroll
            for (int u = 0; u < kSeedUnroll; ++u) {
                const bool alive = has[u] && !lds_removed(rmv, v[u]);
                int32_t d = alive ? (int32_t)(cnt[v[u]] & 0xffffu) : 0;
                // the bound acts like a key degree that never drops
                if (b[u] != kLdsUbNone) d = min(d, (int32_t)b[u]);
                const bool seed = alive && d <= L;
                const bool keep = alive && !seed;
                uint32_t len = 0;
                if (seed) {
                    atomicOr(&rmv[v[u] >> 5], 1u << (v[u] & 31));
                    S.lv[v[u]] = L;
                    len = S.rk[v[u]] + S.rc[v[u]];
                }
                if (keep) mn = min(mn, (uint32_t)d);
                wave_append_len(seed, v[u], len, F, &cur[0], slot_u64(cur, 4));
                wave_append(keep, v[u] | (b[u] << 16), B, &cur[1]);
            }
        }
        wave_min_to(mn, &cur[2]);
        __syncthreads();
        uint32_t nF = cur[0];
        nA = cur[1];
        uint32_t mnall = cur[2];
        uint64_t fe = *slot_u64(cur, 4);
        __syncthreads();
        trace(true, nF);
        A = B;
        {
            uint32_t* t = B;
            B = Bn;
            Bn = t;
        }
        ++s;
        removed += nF;

        if (nA == 0) nF = 0;  // the seeds took every alive vertex
        while (nF > 0) {
            uint32_t* nxt = slots + kSlotStride * (s & 1);
            uint32_t* nx2 = slots + kSlotStride * ((s + 1) & 1);
            if (tid == 0) {
                nx2[0] = 0;
                nx2[1] = 0;
                nx2[2] = UINT_MAX;
                *slot_u64(nx2, 4) = 0;
            }
            uint32_t mn2 = UINT_MAX;
            if (fe >= 64ull * nF) {
                // long lists (a dense core): one wavefront per frontier
                // vertex, coalesced list reads, no tile search
                const uint32_t lane = (uint32_t)lane_id();
                for (uint32_t i = tid / W; i < nF; i += BLK / W) {
                    const uint32_t v = F[i];
                    const uint32_t kl = S.rk[v], tot = kl + S.rc[v];
                    const uint32_t* kn = P.K.nbr + P.K.off[v];
                    const uint32_t* cn = P.C.nbr + P.C.off[v];
                    for (uint32_t jb = 0; jb < tot; jb += kLdsUnroll * W) {
                        uint32_t x[kLdsUnroll];
                        bool ks[kLdsUnroll], claim[kLdsUnroll];
#pragma unroll
                        for (int u = 0; u < kLdsUnroll; ++u) {
                            const uint32_t e = jb + u * W + lane;
                            ks[u] = e < kl;
                            x[u] = e < tot ? (ks[u] ? kn[e] : cn[e - kl]) : 0u;
                            claim[u] = false;
                        }
#pragma unroll
                        for (int u = 0; u < kLdsUnroll; ++u)
                            if (jb + u * W + lane < tot)
                                claim[u] = lds_edge(cnt, rmv, x[u], ks[u], k, L, mn2);
#pragma unroll
                        for (int u = 0; u < kLdsUnroll; ++u) {
                            uint32_t len = 0;
                            if (claim[u]) {
                                S.lv[x[u]] = L;
                                len = S.rk[x[u]] + S.rc[x[u]];
                            }
                            wave_append_len(claim[u], x[u], len, G, &nxt[0], slot_u64(nxt, 4));
                        }
                    }
                }
            } else {
                for (uint32_t base = 0; base < nF; base += BLK) {
                    const uint32_t i = base + tid;
                    const bool has = i < nF;
                    const uint32_t v = has ? F[i] : 0u;
                    __syncthreads();
                    uint32_t kl = 0, cl = 0;
                    if (has) {
                        T.koff[tid] = P.K.off[v];
                        T.coff[tid] = P.C.off[v];
                        kl = S.rk[v];
                        cl = S.rc[v];
                    }
                    T.klen[tid] = kl;
                    if (tid == 0) T.nitems = min(nF - base, (uint32_t)BLK);
                    T.scan[tid] = block_excl_scan<BLK>(kl + cl, T.wsum, &T.total);
                    __syncthreads();
                    const uint32_t total = T.total;
                    for (uint32_t eb = 0; eb < total; eb += kLdsUnroll * BLK) {
                        uint32_t x[kLdsUnroll];
                        bool ks[kLdsUnroll], claim[kLdsUnroll];
#pragma unroll
                        for (int u = 0; u < kLdsUnroll; ++u) {
                            const uint32_t e = eb + u * BLK + tid;
                            x[u] = 0;
                            ks[u] = false;
                            claim[u] = false;
                            if (e < total) {
                                const uint32_t it = tile_find<BLK>(T, e);
                                const uint32_t t = e - T.scan[it];
                                ks[u] = t < T.klen[it];
                                x[u] = ks[u] ? P.K.nbr[T.koff[it] + t]
                                             : P.C.nbr[T.coff[it] + (t - T.klen[it])];
                            }
                        }
#pragma unroll
                        for (int u = 0; u < kLdsUnroll; ++u)
                            if (eb + u * BLK + tid < total)
                                claim[u] = lds_edge(cnt, rmv, x[u], ks[u], k, L, mn2);
#pragma unroll
                        for (int u = 0; u < kLdsUnroll; ++u) {
                            uint32_t len = 0;
                            if (claim[u]) {
                                S.lv[x[u]] = L;
                                len = S.rk[x[u]] + S.rc[x[u]];
                            }
                            wave_append_len(claim[u], x[u], len, G, &nxt[0], slot_u64(nxt, 4));
                        }
                    }
                }
            }
            wave_min_to(mn2, &nxt[2]);
            __syncthreads();
            nF = nxt[0];
            mnall = min(mnall, nxt[2]);
            fe = *slot_u64(nxt, 4);
            __syncthreads();
            trace(false, nF);
            {
                uint32_t* t = F;
                F = G;
                G = t;
            }
            ++s;
            removed += nF;
        }
        if (removed < nm) {
            if (nA == 0 || mnall == UINT_MAX) {
                failed = true;
                break;
            }
            L = max(L + 1, (int32_t)min(mnall, (uint32_t)INT_MAX));
        }
    }
    if (tid == 0) {
        if (lmax_out) *lmax_out = nm ? L : 0;
        if (failed) consume(atomicMax(err, (uint32_t)kErrInconsistent));
        if (P.stats) {
            const uint64_t dt = wall_clock64() - t_start;
            P.stats[kStatPhases] = s;
            P.stats[kStatLevels] = levels;
            P.stats[kStatTicksLo] = (uint32_t)dt;
            P.stats[kStatTicksHi] = (uint32_t)(dt >> 32);
            P.stats[kStatFailed] = failed ? 1u : 0u;
            P.stats[kStatRemoved] = removed;
            P.stats[kStatAlive] = nA;
            P.stats[kStatLevel] = (uint32_t)L;
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < nm; i += BLK) P.out_row[P.order[i]] = S.lv[i];
}

// Small rows: each workgroup takes the next row from a queue (longest
// first) and runs it LDS-resident.  Dynamic LDS: cap counter words, then
// cap / 32 removed-bit words.
constexpr int kLdsBlk = 1024;
__global__ __launch_bounds__(kLdsBlk) void k_rows_lds(PassSpec base, const uint32_t* vk,
                                                      int32_t* rows, int32_t* lmax_rows,
                                                      const int32_t* ks, uint32_t nks,
                                                      uint32_t* queue, void* scratch,
                                                      uint32_t cap, uint32_t* err) {
    extern __shared__ uint32_t dyn[];
    __shared__ TileLds<kLdsBlk> T;
    __shared__ __align__(16) uint32_t lslots[2 * kSlotStride];
    __shared__ uint32_t s_next;
    const PassScratch S = scratch_at(
        static_cast<uint32_t*>(scratch) + (size_t)blockIdx.x * kScratchArrays * cap, cap);
    for (;;) {
        if (threadIdx.x == 0) s_next = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t idx = s_next;
        __syncthreads();
        if (idx >= nks) break;
        const int32_t k = ks[idx];
        PassSpec P = base;
        P.k = k;
        P.nmembers = vk[k];
        P.out_row = rows + (size_t)k * base.V;
        P.stats = base.stats ? base.stats + (size_t)kStatWords * k : nullptr;
        if (k != base.trace_k) P.trace = nullptr;
        run_pass_lds<kLdsBlk>(P, S, lslots, T, dyn, dyn + cap, lmax_rows + k, err);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlk, 4) void k_pass_grid(PassSpec P, PassScratch S, uint32_t* slots,
                                                    int32_t* lmax_out, GridSync gs,
                                                    uint32_t shrink_at) {
    __shared__ TileLds<kBlk> T;
    gs.rank = blockIdx.x;
    gs.epoch = 0;
    run_pass<kBlk, true, kCombNone>(P, S, slots, T, lmax_out, &gs, nullptr, shrink_at, nullptr);
}

// Two independent grid passes in one launch: blocks [0, g0.nblocks) run
// pass 0, the rest pass 1, each half with its own barrier words, slots and
// scratch (the kmax pass beside row 0, which needs no kmaxes).
struct PassPair {
    PassSpec P[2];
    PassScratch S[2];
    uint32_t* slots[2];
    int32_t* lmax_out[2];
    GridSync gs[2];
};

__global__ __launch_bounds__(kBlk, 4) void k_pass_pair(PassPair pp, uint32_t shrink_at) {
    __shared__ TileLds<kBlk> T;
    const int h = blockIdx.x >= pp.gs[0].nblocks ? 1 : 0;
    GridSync gs = pp.gs[h];
    gs.rank = h ? blockIdx.x - pp.gs[0].nblocks : blockIdx.x;
    gs.epoch = 0;
    run_pass<kBlk, true, kCombNone>(pp.P[h], pp.S[h], pp.slots[h], T, pp.lmax_out[h], &gs,
                                    nullptr, shrink_at, nullptr);
}

// out[r] = in[order[r]]: a row by rank.
__global__ void k_gather_rank(const int32_t* in, const uint32_t* order, uint32_t V,
                              int32_t* out) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < V; r += gridDim.x * blockDim.x)
        out[r] = in[order[r]];
}

// Persistent rows kernel: block b belongs to team block_team[b]; a team
// runs its passes one after another, re-synchronising between them.  A
// pass is either a peel of row team_ks[p] (team_kb[p] < 0) or the rows
// team_ks[p]+1 .. team_kb[p] derived from row team_ks[p] (chain_device.hpp);
// a team runs all its peels before its chains, so every anchor's values are
// complete when its chain starts (row 0's before the launch).
// Two workgroups of kBlk threads must be resident per CU (the launch packs
// two per CU and a team's members wait for each other): 4 waves per SIMD.
static_assert(chain::kCBlk == kBlk, "chains run in the rows kernel's workgroups");
__global__ __launch_bounds__(kBlk, 4) void k_rows_teams(PassSpec base, const uint32_t* vk,
                                                     int32_t* rows, int32_t* lmax_rows,
                                                     const TeamDesc* teams,
                                                     const uint32_t* block_team,
                                                     const int32_t* team_ks,
                                                     const int32_t* team_kb, void* scratch,
                                                     uint32_t* sync_pool, uint32_t* slot_pool,
                                                     uint64_t timeout_ticks, uint32_t* err,
                                                     uint32_t shrink_at, uint32_t chain0_preset) {
    __shared__ TileLds<kBlk> T;
    __shared__ uint32_t comb[2 * kCombRanks];
    const uint32_t team = block_team[blockIdx.x];
    if (team == ~0u) return;  // an idle block (its XCD had no team for it)
    for (uint32_t j = threadIdx.x; j < 2 * kCombRanks; j += kBlk) comb[j] = 0;
    __syncthreads();
    const TeamDesc td = teams[team];
    uint32_t* const team_scratch = static_cast<uint32_t*>(scratch) + td.scratch_off;
    const PassScratch S = scratch_at(team_scratch, td.scratch_cap);
    GridSync gs;
    gs.w = sync_pool + td.sync_off;
    gs.nblocks = td.size;
    gs.rank = block_team[gridDim.x + blockIdx.x];
    gs.timeout_ticks = timeout_ticks;
    gs.epoch = 0;
    // the team's peels first, then its chains (every anchor is peeled by
    // then); two loops keep the two kinds of pass apart in registers
    uint32_t seq = 0;
    for (uint32_t p = td.pass_begin; p < td.pass_end; ++p) {
        if (team_kb[p] >= 0) continue;
        const int32_t k = team_ks[p];
        gs.nblocks = td.size;  // a peel pass may have narrowed the team
        PassSpec P = base;
        P.k = k;
        P.nmembers = vk[k];
        P.out_row = rows + (size_t)k * base.V;
        P.stats = base.stats ? base.stats + (size_t)kStatWords * k : nullptr;
        if (k != base.trace_k) P.trace = nullptr;
        // one code path for every team size (a team of one runs the same
        // pass with workgroup barriers), which keeps the kernel within the
        // registers of two resident workgroups per CU
        run_pass<kBlk, true, kCombByRank>(P, S, slot_pool + td.slot_off, T, lmax_rows + k, &gs,
                                          err, shrink_at, comb);
        if (!team_resync(gs, ++seq)) return;
    }
    gs.nblocks = td.size;
    uint32_t tag = 0;  // round tags of the team's chains
    for (uint32_t p = td.pass_begin; p < td.pass_end; ++p) {
        if (team_kb[p] < 0) continue;
        chain::ChainArgs A;
        A.V = base.V;
        A.K = base.K;
        A.C = base.C;
        A.order = base.order;
        A.vk = vk;
        A.rows = rows;
        A.lmax_rows = lmax_rows;
        A.stats = base.stats;
        A.err = err;
        const int32_t a = team_ks[p], b = team_kb[p] & ~kChainFromRow0;
        const int32_t src = (team_kb[p] & kChainFromRow0) ? 0 : a;
        const bool preset = a == 0 && chain0_preset != 0;
        if (!chain::run_chain(A, chain::chain_state(team_scratch, td.scratch_cap),
                              slot_pool + td.slot_off, gs, a, b, src, preset, tag))
            return;
        if (!team_resync(gs, ++seq)) return;
    }
}

// ---- ordering of vertices by kmax (descending), counting sort ----------

constexpr int kHistLds = 4096;

__global__ void k_kmax_hist(const int32_t* kmaxes, uint32_t V, int32_t kmax, uint32_t* hist) {
    __shared__ uint32_t h[kHistLds];
    const bool lds = kmax + 1 <= kHistLds;
    if (lds)
        for (int i = threadIdx.x; i <= kmax; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const int32_t b = kmaxes[v];
        if (lds)
            atomicAdd(&h[b], 1u);
        else
            atomicAdd(&hist[b], 1u);
    }
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i <= kmax; i += blockDim.x)
            if (h[i]) atomicAdd(&hist[i], h[i]);
}

// vk[j] = #{v : kmaxes[v] >= j} for j = 0..kmax+1; hist becomes the start
// cursor of bucket j in descending order (= vk[j+1]).  Single workgroup.
__global__ void k_kmax_suffix(uint32_t* hist, uint32_t* vk, int32_t kmax) {
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        vk[kmax + 1] = 0;
        for (int32_t j = kmax; j >= 0; --j) {
            const uint32_t c = hist[j];
            hist[j] = acc;
            acc += c;
            vk[j] = acc;
        }
    }
}

__global__ void k_kmax_scatter(const int32_t* kmaxes, uint32_t V, int32_t kmax, uint32_t* cursor,
                               uint32_t* order) {
    __shared__ uint32_t h[kHistLds];
    __shared__ uint32_t base[kHistLds];
    const bool lds = kmax + 1 <= kHistLds;
    const uint32_t tile0 = blockIdx.x * blockDim.x;
    for (uint32_t t0 = tile0; t0 < V; t0 += gridDim.x * blockDim.x) {
        const uint32_t v = t0 + threadIdx.x;
        if (!lds) {
            if (v < V) order[atomicAdd(&cursor[kmaxes[v]], 1u)] = v;
            continue;
        }
        for (int i = threadIdx.x; i <= kmax; i += blockDim.x) h[i] = 0;
        __syncthreads();
        uint32_t rank = 0;
        int32_t b = 0;
        if (v < V) {
            b = kmaxes[v];
            rank = atomicAdd(&h[b], 1u);
        }
        __syncthreads();
        for (int i = threadIdx.x; i <= kmax; i += blockDim.x)
            if (h[i]) base[i] = atomicAdd(&cursor[i], h[i]);
        __syncthreads();
        if (v < V) order[base[b] + rank] = v;
        __syncthreads();
    }
}

// hist[m] += edges whose endpoints' smaller kmax is m, counted by 64-bit
// LDS atomics per workgroup and added to hist at the end.  A list of more
// than kCoreEdgesLong entries is cut into slices of that many, queued
// (v << 32 | first entry) for k_core_edges_long, where a workgroup takes
// each slice.
constexpr int kCoreEdgesUnroll = 4;
constexpr uint32_t kCoreEdgesLong = 4096;

__device__ __forceinline__ void core_edges_add(unsigned long long* h, unsigned long long* hist,
                                               bool lds, int32_t m, unsigned long long n) {
    if (lds)
        atomicAdd(&h[m], n);
    else
        atomicAdd(&hist[m], n);
}

// Tiles of kCoreEdgesTile consecutive vertices: their list lengths are
// prefix-summed in LDS and every thread takes kCoreEdgesUnroll edges of the
// tile per step (owner found by binary search), so short lists do not leave
// most lanes idle; lists of more than kCoreEdgesLong entries are queued.
constexpr int kCoreEdgesTile = 1024;
__global__ __launch_bounds__(kCoreEdgesTile) void k_core_edges(const int32_t* kmaxes, Csr out,
                                                               uint32_t V, int32_t kmax,
                                                               unsigned long long* hist,
                                                               unsigned long long* longq,
                                                               uint32_t* longn) {
    __shared__ unsigned long long h[kHistLds];
    __shared__ uint32_t s_off[kCoreEdgesTile], s_scan[kCoreEdgesTile];
    __shared__ int32_t s_kv[kCoreEdgesTile];
    __shared__ uint32_t s_wsum[kCoreEdgesTile / W], s_total;
    const bool lds = kmax + 1 <= kHistLds;
    const uint32_t tid = threadIdx.x;
    if (lds)
        for (int i = tid; i <= kmax; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint32_t base = blockIdx.x * kCoreEdgesTile; base < V;
         base += gridDim.x * kCoreEdgesTile) {
        const uint32_t v = base + tid;
        uint32_t n = 0, o = 0;
        int32_t kv = 0;
        if (v < V) {
            n = (uint32_t)out.deg[v];
            o = out.off[v];
            kv = kmaxes[v];
            if (n > kCoreEdgesLong) {
                const uint32_t ns = (n + kCoreEdgesLong - 1) / kCoreEdgesLong;
                const uint32_t at = atomicAdd(longn, ns);
                for (uint32_t i = 0; i < ns; ++i)
                    longq[at + i] = ((unsigned long long)v << 32) | (i * kCoreEdgesLong);
                n = 0;
            }
        }
        s_off[tid] = o;
        s_kv[tid] = kv;
        s_scan[tid] = block_excl_scan<kCoreEdgesTile>(n, s_wsum, &s_total);
        __syncthreads();
        const uint32_t total = s_total;
        const uint32_t items = min(V - base, (uint32_t)kCoreEdgesTile);
        for (uint32_t eb = 0; eb < total; eb += kCoreEdgesUnroll * kCoreEdgesTile) {
            uint32_t u[kCoreEdgesUnroll];
            int32_t ko[kCoreEdgesUnroll];
#pragma unroll
            for (int r = 0; r < kCoreEdgesUnroll; ++r) {
                const uint32_t e = eb + r * kCoreEdgesTile + tid;
                u[r] = ~0u;
                ko[r] = 0;
                if (e < total) {
                    uint32_t lo = 0, hi = items - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi + 1) >> 1;
                        if (s_scan[mid] <= e)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    u[r] = out.nbr[s_off[lo] + (e - s_scan[lo])];
                    ko[r] = s_kv[lo];
                }
            }
#pragma unroll
            for (int r = 0; r < kCoreEdgesUnroll; ++r)
                if (u[r] != ~0u) core_edges_add(h, hist, lds, min(ko[r], kmaxes[u[r]]), 1ull);
        }
        __syncthreads();
    }
    __syncthreads();
    if (lds)
        for (int i = tid; i <= kmax; i += blockDim.x)
            if (h[i]) atomicAdd(&hist[i], h[i]);
}

// The queued slices of long lists, one workgroup per slice.
__global__ __launch_bounds__(1024) void k_core_edges_long(const int32_t* kmaxes, Csr out,
                                                          int32_t kmax, unsigned long long* hist,
                                                          const unsigned long long* longq,
                                                          const uint32_t* longn) {
    __shared__ unsigned long long h[kHistLds];
    const bool lds = kmax + 1 <= kHistLds;
    if (lds)
        for (int i = threadIdx.x; i <= kmax; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint32_t nq = *longn;
    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const unsigned long long e = longq[q];
        const uint32_t v = (uint32_t)(e >> 32), j0 = (uint32_t)e;
        const int32_t kv = kmaxes[v];
        const uint32_t o = out.off[v], n = (uint32_t)out.deg[v];
        const uint32_t j1 = min(n, j0 + kCoreEdgesLong);
        unsigned long long same_n = 0;
        for (uint32_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
            const int32_t m = min(kv, kmaxes[out.nbr[o + j]]);
            if (m == kv)
                ++same_n;
            else
                core_edges_add(h, hist, lds, m, 1ull);
        }
        // per-thread own-bin counts summed per wave
        for (int d = W / 2; d > 0; d >>= 1) same_n += __shfl_xor(same_n, d, W);
        if (lane_id() == 0 && same_n) core_edges_add(h, hist, lds, kv, same_n);
    }
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i <= kmax; i += blockDim.x)
            if (h[i]) atomicAdd(&hist[i], h[i]);
}

// rows[k * V + v] = 0 for k in (knew[v], min(kold[v], rows_old - 1)]: the
// cells a fallen kmax leaves outside the cores.
__global__ void k_clear_stale(const int32_t* knew, const int32_t* kold, uint32_t V,
                              int32_t rows_old, int32_t* rows) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const int32_t a = knew[v], b = min(kold[v], rows_old - 1);
        for (int32_t k = a + 1; k <= b; ++k) rows[(size_t)k * V + v] = 0;
    }
}

__global__ void k_fill_i32(int32_t* p, int32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_max_i32(const int32_t* p, size_t n, int32_t* out) {
    int32_t m = INT_MIN;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        m = max(m, p[i]);
    wave_max_to(m, out);
}

}  // namespace

int rows_teams_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_rows_teams, kBlk, 0) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

int pass_grid_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pass_grid, kBlk, 0) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_pass_grid(const PassSpec& p, const PassScratch& s, uint32_t* slots,
                            int32_t* lmax_out, const GridSync& gs, hipStream_t st) {
    // the sync words must start at zero for every launch (error word kept)
    hipError_t e = hipMemsetAsync(gs.w, 0, kSyncResetWords * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(gs.w + kSyncFlags, 0, gs.nblocks * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    const uint32_t shrink_at = 16 * kBlk;
    hipLaunchKernelGGL(k_pass_grid, dim3(gs.nblocks), dim3(kBlk), 0, st, p, s, slots, lmax_out, gs,
                       shrink_at);
    return hipGetLastError();
}

hipError_t launch_pass_pair(const PassSpec& p0, const PassScratch& s0, uint32_t* slots0,
                            int32_t* lmax0, const GridSync& gs0, const PassSpec& p1,
                            const PassScratch& s1, uint32_t* slots1, int32_t* lmax1,
                            const GridSync& gs1, hipStream_t st) {
    const GridSync* g[2] = {&gs0, &gs1};
    for (const GridSync* x : g) {
        hipError_t e = hipMemsetAsync(x->w, 0, kSyncResetWords * sizeof(uint32_t), st);
        if (e != hipSuccess) return e;
        e = hipMemsetAsync(x->w + kSyncFlags, 0, x->nblocks * sizeof(uint32_t), st);
        if (e != hipSuccess) return e;
    }
    PassPair pp{};
    pp.P[0] = p0;
    pp.P[1] = p1;
    pp.S[0] = s0;
    pp.S[1] = s1;
    pp.slots[0] = slots0;
    pp.slots[1] = slots1;
    pp.lmax_out[0] = lmax0;
    pp.lmax_out[1] = lmax1;
    pp.gs[0] = gs0;
    pp.gs[1] = gs1;
    const uint32_t shrink_at = 16 * kBlk;
    hipLaunchKernelGGL(k_pass_pair, dim3(gs0.nblocks + gs1.nblocks), dim3(kBlk), 0, st, pp,
                       shrink_at);
    return hipGetLastError();
}

hipError_t launch_gather_rank(const int32_t* in, const uint32_t* order, uint32_t V, int32_t* out,
                              hipStream_t st) {
    uint32_t blocks = (V + 255u) / 256u;
    if (blocks > 4096u) blocks = 4096u;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_gather_rank, dim3(blocks), dim3(256), 0, st, in, order, V, out);
    return hipGetLastError();
}

hipError_t launch_rows_teams(const PassSpec& base, const uint32_t* vk, int32_t* rows,
                             int32_t* lmax_rows, const TeamDesc* teams, const uint32_t* block_team,
                             const int32_t* team_ks, const int32_t* team_kb, uint32_t blocks,
                             void* scratch, uint32_t* sync_pool, uint32_t* slot_pool,
                             uint64_t timeout_ticks, uint32_t* err, uint32_t shrink_at,
                             uint32_t chain0_preset, hipStream_t st) {
    hipLaunchKernelGGL(k_rows_teams, dim3(blocks), dim3(kBlk), 0, st, base, vk, rows, lmax_rows,
                       teams, block_team, team_ks, team_kb, scratch, sync_pool, slot_pool,
                       timeout_ticks, err, shrink_at, chain0_preset);
    return hipGetLastError();
}

hipError_t launch_chain_preset(const PassSpec& base, const uint32_t* vk, int32_t* rows,
                               uint32_t* team_scratch, uint64_t cap, uint32_t* slots, uint32_t V,
                               uint64_t e1, hipStream_t st) {
    chain::ChainArgs A;
    A.V = base.V;
    A.K = base.K;
    A.C = base.C;
    A.order = base.order;
    A.vk = vk;
    A.rows = rows;
    A.lmax_rows = nullptr;
    A.stats = nullptr;
    A.err = nullptr;
    const chain::ChainState S = chain::chain_state(team_scratch, cap);
    const uint32_t vb = std::max(1u, std::min(4096u, (V + 255u) / 256u));
    hipLaunchKernelGGL(chain::k_chain_preset_values, dim3(vb), dim3(256), 0, st, A, S);
    // slices of the long lists: at least 4096 entries, and few enough that
    // V_1 + 4 E_1 / slice entries fit the queue's 2 cap words
    const uint64_t want = (4 * e1 + cap - 1) / std::max<uint64_t>(cap, 1);
    const uint32_t slice = (uint32_t)std::min<uint64_t>(
        std::max<uint64_t>(4096, (want + 63) / 64 * 64), 1u << 30);
    hipLaunchKernelGGL(chain::k_chain_preset_supports, dim3(2048), dim3(chain::kCBlk), 0, st, A,
                       S, slots, slice);
    hipLaunchKernelGGL(chain::k_chain_preset_long, dim3(2048), dim3(chain::kCBlk), 0, st, A, S,
                       slots, slice);
    return hipGetLastError();
}

uint32_t rows_lds_static_bytes() {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_rows_lds)) != hipSuccess)
        return 0;
    return (uint32_t)a.sharedSizeBytes;
}

hipError_t launch_rows_lds(const PassSpec& base, const uint32_t* vk, int32_t* rows,
                           int32_t* lmax_rows, const int32_t* ks, uint32_t nks, uint32_t* queue,
                           uint32_t blocks, void* scratch, uint32_t cap, uint32_t* err,
                           hipStream_t st) {
    if (cap > kLdsMaxMembers) return hipErrorInvalidValue;
    const size_t dyn = 4ull * (cap + (cap + 31) / 32);
    hipLaunchKernelGGL(k_rows_lds, dim3(blocks), dim3(kLdsBlk), dyn, st, base, vk, rows,
                       lmax_rows, ks, nks, queue, scratch, cap, err);
    return hipGetLastError();
}

hipError_t launch_order_by_kmax(const int32_t* kmaxes, uint32_t V, int32_t kmax, uint32_t* hist,
                                uint32_t* vk, uint32_t* order, hipStream_t st) {
    hipError_t e = hipMemsetAsync(hist, 0, sizeof(uint32_t) * (size_t)(kmax + 2), st);
    if (e != hipSuccess) return e;
    const int blocks = (int)min((V + 1023u) / 1024u, 2048u);
    hipLaunchKernelGGL(k_kmax_hist, dim3(blocks ? blocks : 1), dim3(1024), 0, st, kmaxes, V, kmax,
                       hist);
    hipLaunchKernelGGL(k_kmax_suffix, dim3(1), dim3(64), 0, st, hist, vk, kmax);
    hipLaunchKernelGGL(k_kmax_scatter, dim3(blocks ? blocks : 1), dim3(1024), 0, st, kmaxes, V,
                       kmax, hist, order);
    return hipGetLastError();
}

hipError_t launch_core_edges(const int32_t* kmaxes, Csr out, uint32_t V, int32_t kmax,
                             uint64_t* hist, uint64_t* longq, hipStream_t st) {
    // hist[kmax+2] then the queue's length word (zeroed together)
    hipError_t e = hipMemsetAsync(hist, 0, sizeof(uint64_t) * (size_t)(kmax + 3), st);
    if (e != hipSuccess) return e;
    uint32_t* longn = reinterpret_cast<uint32_t*>(hist + kmax + 2);
    const unsigned blocks = (unsigned)min((V + 1023u) / 1024u, 4096u);
    hipLaunchKernelGGL(k_core_edges, dim3(blocks ? blocks : 1), dim3(1024), 0, st, kmaxes, out, V,
                       kmax, reinterpret_cast<unsigned long long*>(hist),
                       reinterpret_cast<unsigned long long*>(longq), longn);
    hipLaunchKernelGGL(k_core_edges_long, dim3(1024), dim3(1024), 0, st, kmaxes, out, kmax,
                       reinterpret_cast<unsigned long long*>(hist),
                       reinterpret_cast<const unsigned long long*>(longq), longn);
    return hipGetLastError();
}

hipError_t launch_clear_stale(const int32_t* knew, const int32_t* kold, uint32_t V,
                              int32_t rows_old, int32_t rows_new, int32_t* rows, hipStream_t st) {
    (void)rows_new;
    if (!V || rows_old <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)min((V + 255u) / 256u, 8192u);
    hipLaunchKernelGGL(k_clear_stale, dim3(blocks), dim3(256), 0, st, knew, kold, V, rows_old,
                       rows);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t st) {
    if (!n) return hipSuccess;
    const size_t blocks = min((n + 255) / 256, (size_t)4096);
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_max_i32(const int32_t* p, size_t n, int32_t* out, hipStream_t st) {
    const size_t blocks = min((n + 255) / 256, (size_t)1024);
    hipLaunchKernelGGL(k_max_i32, dim3((unsigned)(blocks ? blocks : 1)), dim3(256), 0, st, p, n,
                       out);
    return hipGetLastError();
}

}  // namespace dcx
