// Rows derived from the row above ("chains") for the batched decomposition.
//
// Row k's values l_k are the greatest fixpoint of the local condition
//   x(v) <= max{ l : |{u in out_k(v) : x(u) >= l}| >= l  and
//                    |{w in in_k(v)  : x(w) >= l}| >= k }
// over the members of the (k,0)-core (ranks [0, V_k) of the core layout;
// out_k / in_k: the member prefixes of the sorted lists).  Since
// D_{k,l} lies inside D_{k-1,l}, row k-1 restricted to V_k is an upper
// bound of row k, and lowering the vertices that violate the condition
// until none does reaches row k exactly (the h-index refinement of the
// reference's kListRefineHIndex family, main.cu:1245-1438, here run from the
// row above instead of from the degrees).  Between consecutive rows only a
// small share of the cells changes, so a chain keeps, per member v, its
// supports at its own level
//   os(v) = |{u in out_k(v) : x(u) >= x(v)}|,  cs(v) = |{w in in_k(v) : x(w) >= x(v)}|
// and visits only vertices whose support falls below the condition
// (os < x or cs < k) and the lists of vertices that change or leave the
// core -- instead of peeling all E_k edges of every row.
//
// One chain: row a (the anchor, already in `rows`) gives rows a+1..b.  Per
// row k:
//   1. the vertices leaving the core (ranks [V_k, V_{k-1})) are taken out
//      of their neighbours' prefixes and supports;
//   2. a scan of [0, V_k) queues every violator (cs < k catches the
//      raised in-degree threshold);
//   3. rounds: each queued vertex is re-evaluated from scratch (its value
//      and both supports at that value; values only fall); then every
//      vertex that fell from old to nv decrements the supports of the
//      neighbours y with nv < x(y) <= old, and a decrement that takes a
//      support below its threshold queues y for the next round;
//   4. the row is written to `rows` (original ids).
// Evaluations read the current values (a neighbour's new value may already
// be visible); such a read can only make a support lower than exact, which
// queues its vertex once more, and re-evaluation is exact.  Values never
// drop below l_k (a vertex's l_k is feasible while every neighbour is at
// or above its own l_k), so the fixpoint reached is l_k.
//
// A chain runs inside the persistent rows launch (k_rows_teams), on the
// team that peeled its anchor or, for row 0's chain, on a team of its own:
// workgroups of one XCD with the bounded team barrier of grid_sync.hpp,
// every shared word accessed agent-coherently.  A vertex's lists are walked
// by one thread when they hold at most kTiny entries together, by one
// wavefront below kLongList entries and by a whole workgroup from there:
// most vertices of the low rows have a handful of member neighbours, and a
// wavefront per vertex would leave such walks bound by the latency of one
// vertex after another.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "dcx_internal.hpp"
#include "grid_sync.hpp"

namespace dcx {
namespace chain {

constexpr int W = 64;
constexpr int kCBlk = 512;
constexpr int kCWaves = kCBlk / W;
constexpr uint32_t kLongList = 2048;
constexpr uint32_t kTiny = 16;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (W - 1); }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x / W; }

// Walk groups: one thread (1), one wavefront (W) or the workgroup (kCBlk);
// gme<G>() is the caller's index inside its group.
template <int G>
__device__ __forceinline__ uint32_t gme() {
    static_assert(G == 1 || G == W || G == kCBlk, "walk group");
    return G == 1 ? 0u : (G == W ? lane_id() : threadIdx.x);
}

template <typename T>
__device__ __forceinline__ T ldc(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void stc(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) x += (uint32_t)__shfl_xor(x, d, W);
    return x;
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) x += (uint64_t)__shfl_xor((unsigned long long)x, d, W);
    return x;
}

// Per lane of the calling wavefront an item (has: valid) whose lists hold
// len < kLongList entries: f(group, item) runs on the lane itself for
// items of at most kTiny entries, and on the whole wavefront, item after
// item, for the others (group: std::integral_constant<int, 1 or W>).
// Called by all lanes of the wavefront.
template <typename F>
__device__ __forceinline__ void by_lane_or_wave(bool has, uint32_t len, uint32_t item, F&& f) {
    const bool tiny = has && len <= kTiny;
    if (tiny) f(std::integral_constant<int, 1>(), item);
    unsigned long long m = __ballot(has && !tiny);
    while (m) {
        const int l = __ffsll(m) - 1;
        m &= m - 1;
        f(std::integral_constant<int, W>(), (uint32_t)__shfl((int)item, l, W));
    }
}

__device__ __forceinline__ uint32_t first_not_below(const uint32_t* a, uint32_t n, uint32_t bound) {
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

// Per-chain state (team scratch, cap words each).
struct ChainState {
    int32_t* x;      // current value per rank
    int32_t* os;     // out-support at own level
    int32_t* cs;     // in-support at own level
    uint32_t* pk;    // member prefix length of the K (out) list
    uint32_t* pc;    // member prefix length of the C (in) list
    uint32_t* mark;  // last round tag the vertex was queued for
    uint32_t* qs0;   // short-list queues (ping-pong by round parity): vertices
    uint32_t* qs1;   //   of at most kTiny entries from the front, the others
                     //   from the back (qmed)
    uint32_t* ql0;   // long-list queues
    uint32_t* ql1;
    int32_t* oqs;    // old value of each short-queue entry (same positions)
    int32_t* oql;    // old value of each long-queue entry
    uint32_t cap;    // words per array
};

__host__ __device__ __forceinline__ ChainState chain_state(uint32_t* p, uint64_t cap) {
    ChainState s;
    s.x = reinterpret_cast<int32_t*>(p);
    s.os = reinterpret_cast<int32_t*>(p + cap);
    s.cs = reinterpret_cast<int32_t*>(p + 2 * cap);
    s.pk = p + 3 * cap;
    s.pc = p + 4 * cap;
    s.mark = p + 5 * cap;
    s.qs0 = p + 6 * cap;
    s.qs1 = p + 7 * cap;
    s.ql0 = p + 8 * cap;
    s.ql1 = p + 9 * cap;
    s.oqs = reinterpret_cast<int32_t*>(p + 10 * cap);
    s.oql = reinterpret_cast<int32_t*>(p + 11 * cap);
    s.cap = (uint32_t)cap;
    return s;
}
static_assert(kScratchArrays >= 12, "a chain needs 12 scratch arrays");
__device__ __forceinline__ uint32_t* qshort(const ChainState& S, uint32_t par) {
    return par ? S.qs1 : S.qs0;
}
__device__ __forceinline__ uint32_t* qlong(const ChainState& S, uint32_t par) {
    return par ? S.ql1 : S.ql0;
}
// Position of the j-th medium entry in a short queue (and in oqs): the
// tiny and medium entries of one round are distinct members, so the two
// ends never meet.
__device__ __forceinline__ uint32_t qmed(const ChainState& S, uint32_t j) { return S.cap - 1 - j; }

// Slot words of a chain team (kSlotWords per team), queue lengths ping-pong
// by round parity: [0..1] tiny entries, [2..3] long entries, [4] row
// maximum, [6..7] medium entries, [8..9] list entries the row's walks read
// (uint64).
enum : uint32_t { kCqTiny = 0, kCqLong = 2, kCqMax = 4, kCqMed = 6, kCqVisited = 8 };
__device__ __forceinline__ unsigned long long* visited_word(uint32_t* slots) {
    return reinterpret_cast<unsigned long long*>(slots + kCqVisited);
}

struct ChainArgs {
    uint32_t V;  // vertices (row stride of `rows`)
    Csr K, C;    // core layout
    const uint32_t* order;
    const uint32_t* vk;
    int32_t* rows;
    int32_t* lmax_rows;
    uint32_t* stats;  // nullable: [kStatWords] of row k at stats + kStatWords * k
    uint32_t* err;
};

// Queues y for round tag `tag` unless it already is.
__device__ __forceinline__ void enqueue(const ChainState& S, uint32_t* slots, uint32_t par,
                                        uint32_t y, uint32_t tag) {
    if (atomicMax(&S.mark[y], tag) >= tag) return;
    const uint32_t len = ldc(&S.pk[y]) + ldc(&S.pc[y]);
    if (len >= kLongList) {
        const uint32_t at = atomicAdd(&slots[kCqLong + par], 1u);
        stc(&qlong(S, par)[at], y);
    } else if (len > kTiny) {
        const uint32_t at = atomicAdd(&slots[kCqMed + par], 1u);
        stc(&qshort(S, par)[qmed(S, at)], y);
    } else {
        const uint32_t at = atomicAdd(&slots[kCqTiny + par], 1u);
        stc(&qshort(S, par)[at], y);
    }
}

// Counts of v's member neighbours with value >= l (K side, C side), summed
// over the calling group G (kCBlk: red is LDS of kCWaves * 2 words).
template <int G>
__device__ __forceinline__ void count_ge(const ChainArgs& A, const ChainState& S, uint32_t v,
                                         int32_t l, uint32_t* red, uint32_t& ck, uint32_t& cc,
                                         uint64_t& vis) {
    const uint32_t nk = ldc(&S.pk[v]), nc = ldc(&S.pc[v]);
    const uint32_t me = gme<G>();
    if (me == 0) vis += nk + nc;  // summed over the wavefront at the row's end
    const uint32_t* kn = A.K.nbr + A.K.off[v];
    const uint32_t* cn = A.C.nbr + A.C.off[v];
    uint32_t a = 0, b = 0;
    for (uint32_t j = me; j < nk; j += G) a += ldc(&S.x[kn[j]]) >= l ? 1u : 0u;
    for (uint32_t j = me; j < nc; j += G) b += ldc(&S.x[cn[j]]) >= l ? 1u : 0u;
    if (G >= W) {
        a = wave_sum(a);
        b = wave_sum(b);
    }
    if (G == kCBlk) {
        __syncthreads();
        if (lane_id() == 0) {
            red[2 * wave_id()] = a;
            red[2 * wave_id() + 1] = b;
        }
        __syncthreads();
        a = b = 0;
        for (int w = 0; w < kCWaves; ++w) {
            a += red[2 * w];
            b += red[2 * w + 1];
        }
    }
    ck = a;
    cc = b;
}

// Thresholds tested per pass over a vertex's lists when it is re-evaluated.
constexpr int kProbe = 4;

// For thresholds t[0..kProbe) (decreasing; an entry <= `below` is unused):
// counts of v's member neighbours with value >= t[j], K side ck[j] and C
// side cc[j], summed over the calling group G (kCBlk: red holds
// 2 * kProbe * kCWaves words).
template <int G>
__device__ __forceinline__ void count_multi(const ChainArgs& A, const ChainState& S, uint32_t v,
                                            const int32_t (&t)[kProbe], uint32_t* red,
                                            uint32_t (&ck)[kProbe], uint32_t (&cc)[kProbe],
                                            uint64_t& vis) {
    const uint32_t nk = ldc(&S.pk[v]), nc = ldc(&S.pc[v]);
    const uint32_t me = gme<G>();
    if (me == 0) vis += nk + nc;  // summed over the wavefront at the row's end
    const uint32_t* kn = A.K.nbr + A.K.off[v];
    const uint32_t* cn = A.C.nbr + A.C.off[v];
    // per-lane counts of both sides packed in one word (K low, C high half:
    // a lane sees at most 2^16 - 1 entries of one list -- lists walked by
    // a wavefront or workgroup are shorter than 2^16 times its width, by a
    // thread at most kTiny long), unpacked before the group sums
    uint32_t acc[kProbe];
#pragma unroll
    for (int j = 0; j < kProbe; ++j) acc[j] = 0;
    for (uint32_t i = me; i < nk; i += G) {
        const int32_t xv = ldc(&S.x[kn[i]]);
#pragma unroll
        for (int j = 0; j < kProbe; ++j) acc[j] += xv >= t[j] ? 1u : 0u;
    }
    for (uint32_t i = me; i < nc; i += G) {
        const int32_t xv = ldc(&S.x[cn[i]]);
#pragma unroll
        for (int j = 0; j < kProbe; ++j) acc[j] += xv >= t[j] ? 0x10000u : 0u;
    }
#pragma unroll
    for (int j = 0; j < kProbe; ++j) {
        ck[j] = acc[j] & 0xffffu;
        cc[j] = acc[j] >> 16;
        if (G >= W) {
            ck[j] = wave_sum(ck[j]);
            cc[j] = wave_sum(cc[j]);
        }
    }
    if (G == kCBlk) {
        __syncthreads();
        if (lane_id() == 0) {
#pragma unroll
            for (int j = 0; j < kProbe; ++j) {
                red[(2 * j) * kCWaves + wave_id()] = ck[j];
                red[(2 * j + 1) * kCWaves + wave_id()] = cc[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kProbe; ++j) {
            uint32_t a = 0, b = 0;
            for (int w = 0; w < kCWaves; ++w) {
                a += red[(2 * j) * kCWaves + w];
                b += red[(2 * j + 1) * kCWaves + w];
            }
            ck[j] = a;
            cc[j] = b;
        }
    }
}

// Re-evaluates v (value old): its new value nv (at most old) and both
// supports at it.  A value l is feasible when at least l K-neighbours and
// at least k C-neighbours have values >= l; feasible values are closed
// downwards and 0 always is (its supports are the prefix lengths), so nv is
// the largest feasible l <= old.  Every pass over the lists tests kProbe
// thresholds: the first old, old-1, old-2, old-4 (values most
// often fall by one or two), the next ones spread over the interval left,
// until the largest feasible value is a tested one.  The supports kept are
// the counts of the pass that found nv feasible: a fresh count could read
// neighbours that fell since and leave a support below its threshold that
// no decrement crosses.  Returns nv; the caller's group has the same view.
template <int G>
__device__ __forceinline__ int32_t evaluate(const ChainArgs& A, const ChainState& S, uint32_t v,
                                            int32_t old, int32_t k, uint32_t* red, uint32_t& os,
                                            uint32_t& cs, uint64_t& vis) {
    int32_t lo = 0, hi = old;  // lo feasible with supports (lk, lc); values above hi are not
    uint32_t lk = ldc(&S.pk[v]), lc = ldc(&S.pc[v]);
    bool first = true;
    while (lo < hi) {
        int32_t t[kProbe];
        if (first) {
            t[0] = old;
#pragma unroll
            for (int j = 1; j < kProbe; ++j) t[j] = max(old - (1 << (j - 1)), lo);
        } else {
            const int32_t step = (hi - lo + kProbe - 1) / kProbe;
#pragma unroll
            for (int j = 0; j < kProbe; ++j) t[j] = max(hi - j * step, lo);
        }
        first = false;
        uint32_t ck[kProbe], cc[kProbe];
        count_multi<G>(A, S, v, t, red, ck, cc, vis);
        // the largest tested threshold above lo that is feasible
        int jf = -1;
        int32_t below = hi + 1;  // the smallest tested value above lo, if none is feasible
#pragma unroll
        for (int j = 0; j < kProbe; ++j) {
            if (jf >= 0 || t[j] <= lo) continue;
            if ((int32_t)ck[j] >= t[j] && (int32_t)cc[j] >= k) {
                jf = j;
            } else {
                below = t[j];
            }
        }
        if (jf >= 0) {
            int32_t tf = 0;
            uint32_t fk = 0, fc = 0;
#pragma unroll
            for (int j = 0; j < kProbe; ++j)
                if (j == jf) {
                    tf = t[j];
                    fk = ck[j];
                    fc = cc[j];
                }
            lo = tf;
            lk = fk;
            lc = fc;
        }
        hi = below - 1;
        if (hi < lo) hi = lo;
    }
    os = lk;
    cs = lc;
    return lo;
}

// v fell from old to nv: neighbours y with nv < x(y) <= old lose v from a
// support; a support that crosses below its threshold queues y.
template <int G>
__device__ __forceinline__ void notify(const ChainArgs& A, const ChainState& S, uint32_t* slots,
                                       uint32_t par, uint32_t v, int32_t old, int32_t nv,
                                       int32_t k, uint32_t tag, uint64_t& vis) {
    const uint32_t nk = ldc(&S.pk[v]), nc = ldc(&S.pc[v]);
    const uint32_t me = gme<G>();
    if (me == 0) vis += nk + nc;  // summed over the wavefront at the row's end
    const uint32_t* kn = A.K.nbr + A.K.off[v];
    const uint32_t* cn = A.C.nbr + A.C.off[v];
    // u in C-list of v: v is in u's K (out) list -> os(u)
    for (uint32_t j = me; j < nc; j += G) {
        const uint32_t u = cn[j];
        const int32_t xu = ldc(&S.x[u]);
        if (xu > nv && xu <= old) {
            const int32_t s = atomicSub(&S.os[u], 1);
            if (s == xu) enqueue(S, slots, par, u, tag);
        }
    }
    // w in K-list of v: v is in w's C (in) list -> cs(w)
    for (uint32_t j = me; j < nk; j += G) {
        const uint32_t w = kn[j];
        const int32_t xw = ldc(&S.x[w]);
        if (xw > nv && xw <= old) {
            const int32_t s = atomicSub(&S.cs[w], 1);
            if (s == k) enqueue(S, slots, par, w, tag);
        }
    }
}

// r (value xr) leaves the core at row k (new member bound V): its member
// neighbours drop it from their prefixes and supports.
template <int G>
__device__ __forceinline__ void leave(const ChainArgs& A, const ChainState& S, uint32_t r,
                                      uint32_t V, uint64_t& vis) {
    const int32_t xr = ldc(&S.x[r]);
    const uint32_t nk = ldc(&S.pk[r]), nc = ldc(&S.pc[r]);
    const uint32_t me = gme<G>();
    if (me == 0) vis += nk + nc;  // summed over the wavefront at the row's end
    const uint32_t* kn = A.K.nbr + A.K.off[r];
    const uint32_t* cn = A.C.nbr + A.C.off[r];
    for (uint32_t j = me; j < nc; j += G) {
        const uint32_t u = cn[j];
        if (u >= V) continue;
        consume(atomicSub(&S.pk[u], 1u));
        if (xr >= ldc(&S.x[u])) consume((uint32_t)atomicSub(&S.os[u], 1));
    }
    for (uint32_t j = me; j < nk; j += G) {
        const uint32_t w = kn[j];
        if (w >= V) continue;
        consume(atomicSub(&S.pc[w], 1u));
        if (xr >= ldc(&S.x[w])) consume((uint32_t)atomicSub(&S.cs[w], 1));
    }
}


// ---- set-up of row 0's chain on the whole grid, before the rows launch
// (row 0 is ready by then; a team would do the same work on a sixteenth of
// the chip).  S is the chain's team scratch, slots its team slots (zeroed).

// x, marks and member prefixes for row 1 (one thread per member rank).
__global__ __launch_bounds__(256) void k_chain_preset_values(ChainArgs A, ChainState S) {
    const uint32_t V1 = A.vk[1];
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < V1;
         r += gridDim.x * blockDim.x) {
        const uint32_t pk = first_not_below(A.K.nbr + A.K.off[r], (uint32_t)A.K.deg[r], V1);
        S.x[r] = min(A.rows[A.order[r]], (int32_t)pk);  // (both bound row 1)
        S.mark[r] = 0u;
        S.pk[r] = pk;
        S.pc[r] = first_not_below(A.C.nbr + A.C.off[r], (uint32_t)A.C.deg[r], V1);
    }
}

// Supports at own level for ranks [0, V_1): one thread or wavefront per
// vertex; the lists of a vertex with long lists are cut into slices of
// `slice` entries (both lists in a row), queued as (rank, first entry) for
// k_chain_preset_long, its supports zeroed for the slices' counts.  The
// queue's two words live in ql0..ql1 and qs0..qs1, each 2 cap words long
// (contiguous in the chain state): the host picks slice >= 4 E_1 / cap, so
// the slices (at most V_1 + 4 E_1 / slice) fit.
__global__ __launch_bounds__(kCBlk) void k_chain_preset_supports(ChainArgs A, ChainState S,
                                                                 uint32_t* slots, uint32_t slice) {
    const uint32_t V1 = A.vk[1];
    uint64_t vis = 0;  // (not logged: row 1's log counts its own walks only)
    for (uint32_t base = (blockIdx.x * kCWaves + wave_id()) * W; base < V1;
         base += gridDim.x * kCBlk) {
        const uint32_t r = base + lane_id();
        const bool has = r < V1;
        const uint32_t len = has ? S.pk[r] + S.pc[r] : 0u;
        const bool lng = len >= kLongList;
        if (lng) {
            S.os[r] = 0;
            S.cs[r] = 0;
            const uint32_t ns = (len + slice - 1) / slice;
            const uint32_t at = atomicAdd(&slots[kCqLong], ns);
            for (uint32_t i = 0; i < ns; ++i) {
                S.ql0[at + i] = r;
                S.qs0[at + i] = i * slice;
            }
        }
        by_lane_or_wave(has && !lng, len, r, [&](auto g, uint32_t rr) {
            constexpr int G = decltype(g)::value;
            uint32_t ck, cc;
            count_ge<G>(A, S, rr, S.x[rr], nullptr, ck, cc, vis);
            if (gme<G>() == 0) {
                S.os[rr] = (int32_t)ck;
                S.cs[rr] = (int32_t)cc;
            }
        });
    }
}

// The queued slices of long lists, one workgroup per slice: counts of the
// slice's member neighbours with value >= x(r), added to r's supports.
__global__ __launch_bounds__(kCBlk) void k_chain_preset_long(ChainArgs A, ChainState S,
                                                             uint32_t* slots, uint32_t slice) {
    __shared__ uint32_t red[2 * kCWaves];
    const uint32_t nq = slots[kCqLong];
    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const uint32_t r = S.ql0[q], j0 = S.qs0[q];
        const uint32_t nk = S.pk[r], nc = S.pc[r];
        const uint32_t j1 = min(nk + nc, j0 + slice);
        const int32_t l = S.x[r];
        const uint32_t* kn = A.K.nbr + A.K.off[r];
        const uint32_t* cn = A.C.nbr + A.C.off[r];
        uint32_t a = 0, b = 0;
        for (uint32_t j = j0 + threadIdx.x; j < j1; j += kCBlk) {
            if (j < nk)
                a += S.x[kn[j]] >= l ? 1u : 0u;
            else
                b += S.x[cn[j - nk]] >= l ? 1u : 0u;
        }
        a = wave_sum(a);
        b = wave_sum(b);
        if (lane_id() == 0) {
            red[2 * wave_id()] = a;
            red[2 * wave_id() + 1] = b;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t ta = 0, tb = 0;
            for (int w = 0; w < kCWaves; ++w) {
                ta += red[2 * w];
                tb += red[2 * w + 1];
            }
            if (ta) atomicAdd(&S.os[r], (int32_t)ta);
            if (tb) atomicAdd(&S.cs[r], (int32_t)tb);
        }
        __syncthreads();
    }
}

// Derives rows a+1..b on the calling team (gs: its barrier words, rank and
// size), starting from row src <= a restricted to the (a+1,0)-core (an upper
// bound of row a+1, since D_{a+1,l} lies inside D_{src,l}; normally src = a,
// the row above).  With `preset`, the chain's values, prefixes and
// supports for row a+1 are already in S (row 0's chain, set up by the
// k_chain_preset_* kernels).  `tag` carries the round tags from chain to
// chain of the team (they only grow).  Returns false when the team stopped
// on a timeout or on inconsistent supports (reported through A.err).  The
// pass-log entry of each derived row gets its rounds as phases, levels = 1
// for the first row of a chain whose time includes the supports' set-up
// from the row above (not preset), 2 for such a row started from an earlier
// row (src < a), and 0 for the others.  Every start value is also capped by
// the vertex's member out-degree (its prefix pk): v keeps at least l_{a+1}(v)
// member out-neighbours in D_{a+1,l}, so that is an upper bound too.
__device__ inline bool run_chain(const ChainArgs& A, const ChainState& S, uint32_t* slots,
                                 GridSync& gs, int32_t a, int32_t b, int32_t src, bool preset,
                                 uint32_t& tag) {
    __shared__ uint32_t red[2 * kProbe * kCWaves];
    __shared__ uint32_t s_n[3];
    const uint32_t tw = gs.rank, tn = gs.nblocks;
    const uint32_t tid = threadIdx.x;
    const uint32_t gw = tw * kCWaves + wave_id(), nw = tn * kCWaves;  // wavefronts
    const uint32_t gt = tw * kCBlk + tid, nth = tn * kCBlk;            // threads
    bool ok = true;
    auto barrier = [&]() {
        if (!ok) return;
        if (tn > 1)
            ok = grid_barrier(gs, false);
        else
            __syncthreads();
    };
    // the queue lengths of parity p (tiny, medium, long), read by every
    // workgroup after the barrier that ends their filling
    auto counts = [&](uint32_t p, uint32_t& nt, uint32_t& nm, uint32_t& nl) {
        if (tid == 0) {
            s_n[0] = coherent_read(&slots[kCqTiny + p]);
            s_n[1] = coherent_read(&slots[kCqMed + p]);
            s_n[2] = coherent_read(&slots[kCqLong + p]);
        }
        __syncthreads();
        nt = s_n[0];
        nm = s_n[1];
        nl = s_n[2];
        __syncthreads();
    };
    // a rank whose walk does not fit one thread goes to a queue of parity p
    auto defer = [&](uint32_t p, uint32_t r, uint32_t len) {
        if (len < kLongList)
            stc(&qshort(S, p)[qmed(S, atomicAdd(&slots[kCqMed + p], 1u))], r);
        else
            stc(&qlong(S, p)[atomicAdd(&slots[kCqLong + p], 1u)], r);
    };
    const uint64_t t_chain = wall_clock64();
    // list entries read by this thread's walks (a walk is counted by its
    // first thread; summed per wavefront at the end of each row; the
    // supports' set-up counts towards the first row)
    uint64_t vis = 0;
    uint32_t V = A.vk[a + 1];
    if (preset) {
        if (tw == 0 && tid < 8) coherent_write(slots + tid, 0u);
        barrier();
    } else {
        // ---- start from row src: values, prefixes and supports for row a+1
        // (members of row a+1 only: the scratch holds V_{a+1} words at least)
        for (uint32_t r = gt; r < V; r += nth) {
            const uint32_t pk = first_not_below(A.K.nbr + A.K.off[r], (uint32_t)A.K.deg[r], V);
            stc(&S.x[r], min(A.rows[(size_t)src * A.V + A.order[r]], (int32_t)pk));
            stc(&S.mark[r], 0u);
            stc(&S.pk[r], pk);
            stc(&S.pc[r], first_not_below(A.C.nbr + A.C.off[r], (uint32_t)A.C.deg[r], V));
        }
        if (tw == 0 && tid < 8) coherent_write(slots + tid, 0u);
        if (tw == 0 && tid == 0) consume((uint32_t)atomicExch(visited_word(slots), 0ull));
        barrier();
        // supports: tiny lists counted by their own thread, the others
        // queued (parity 0) for wavefronts and workgroups
        for (uint32_t r = gt; r < V && ok; r += nth) {
            const uint32_t len = ldc(&S.pk[r]) + ldc(&S.pc[r]);
            if (len > kTiny) {
                defer(0, r, len);
                continue;
            }
            uint32_t ck, cc;
            count_ge<1>(A, S, r, ldc(&S.x[r]), red, ck, cc, vis);
            stc(&S.os[r], (int32_t)ck);
            stc(&S.cs[r], (int32_t)cc);
        }
        barrier();
        if (!ok) {
            if (tid == 0) consume(atomicMax(A.err, (uint32_t)kErrTimeout));
            return false;
        }
        {
            uint32_t nt, nm, nl;
            counts(0, nt, nm, nl);
            for (uint32_t i = tw; i < nl; i += tn) {
                const uint32_t r = ldc(&qlong(S, 0)[i]);
                uint32_t ck, cc;
                count_ge<kCBlk>(A, S, r, ldc(&S.x[r]), red, ck, cc, vis);
                if (tid == 0) {
                    stc(&S.os[r], (int32_t)ck);
                    stc(&S.cs[r], (int32_t)cc);
                }
            }
            for (uint32_t i = gw; i < nm; i += nw) {
                const uint32_t r = ldc(&qshort(S, 0)[qmed(S, i)]);
                uint32_t ck, cc;
                count_ge<W>(A, S, r, ldc(&S.x[r]), red, ck, cc, vis);
                if (lane_id() == 0) {
                    stc(&S.os[r], (int32_t)ck);
                    stc(&S.cs[r], (int32_t)cc);
                }
            }
        }
        // (parity 0's counters are cleared before their next use, after
        // the barriers that every member passes only once it has read them)
        barrier();
    }
    // ---- rows a+1 .. b
    for (int32_t k = a + 1; k <= b && ok; ++k) {
        const uint64_t t_row = k == a + 1 ? t_chain : wall_clock64();
        V = A.vk[k];
        const uint32_t Vp = k > a + 1 ? A.vk[k - 1] : V;
        if (Vp > V) {
            // 1. leavers [V, Vp): tiny lists by their own thread, the others
            // queued for wavefronts and workgroups.  The queues of parity
            // tag & 1 are free and their counters zero: the previous row's
            // rounds ended on them.
            const uint32_t lp = tag & 1u;
            for (uint32_t r = V + gt; r < Vp; r += nth) {
                const uint32_t len = ldc(&S.pk[r]) + ldc(&S.pc[r]);
                if (len > kTiny)
                    defer(lp, r, len);
                else
                    leave<1>(A, S, r, V, vis);
            }
            barrier();
            uint32_t nt, nm, nl;
            counts(lp, nt, nm, nl);
            for (uint32_t i = tw; i < nl; i += tn) leave<kCBlk>(A, S, ldc(&qlong(S, lp)[i]), V, vis);
            for (uint32_t i = gw; i < nm; i += nw)
                leave<W>(A, S, ldc(&qshort(S, lp)[qmed(S, i)]), V, vis);
            barrier();
        }
        // 2. violators of row k
        ++tag;
        uint32_t par = tag & 1u;
        if (tw == 0 && tid == 0) {
            coherent_write(&slots[kCqTiny + par], 0u);
            coherent_write(&slots[kCqMed + par], 0u);
            coherent_write(&slots[kCqLong + par], 0u);
            coherent_write(&slots[kCqMax], 0u);
        }
        barrier();
        for (uint32_t r = gt; r < V; r += nth) {
            const int32_t xr = ldc(&S.x[r]);
            if (ldc(&S.os[r]) < xr || ldc(&S.cs[r]) < k) enqueue(S, slots, par, r, tag);
        }
        barrier();
        // 3. rounds
        uint32_t rounds = 0;
        for (;;) {
            if (!ok) break;
            uint32_t nt, nm, nl;
            counts(par, nt, nm, nl);
            if (nt + nm + nl == 0) break;
            // every round lowers a value (spurious queueing aside, which
            // one exact re-count ends): more rounds than twice the
            // members means the supports are wrong
            if (++rounds > 2 * V + 64) {
                if (tid == 0) consume(atomicMax(A.err, (uint32_t)kErrInconsistent));
                ok = false;
                break;
            }
            const uint32_t ntag = tag + 1, npar = ntag & 1u;
            if (tw == 0 && tid == 0) {
                coherent_write(&slots[kCqTiny + npar], 0u);
                coherent_write(&slots[kCqMed + npar], 0u);
                coherent_write(&slots[kCqLong + npar], 0u);
            }
            // A. evaluate (long lists by workgroups, medium by wavefronts,
            // tiny by threads)
            for (uint32_t i = tw; i < nl; i += tn) {
                const uint32_t v = ldc(&qlong(S, par)[i]);
                const int32_t old = ldc(&S.x[v]);
                uint32_t os, cs;
                const int32_t nv = evaluate<kCBlk>(A, S, v, old, k, red, os, cs, vis);
                if (tid == 0) {
                    stc(&S.os[v], (int32_t)os);
                    stc(&S.cs[v], (int32_t)cs);
                    stc(&S.oql[i], old);
                    stc(&S.x[v], nv);
                }
                __syncthreads();
            }
            for (uint32_t i = gw; i < nm; i += nw) {
                const uint32_t at = qmed(S, i);
                const uint32_t v = ldc(&qshort(S, par)[at]);
                const int32_t old = ldc(&S.x[v]);
                uint32_t os, cs;
                const int32_t nv = evaluate<W>(A, S, v, old, k, red, os, cs, vis);
                if (lane_id() == 0) {
                    stc(&S.os[v], (int32_t)os);
                    stc(&S.cs[v], (int32_t)cs);
                    stc(&S.oqs[at], old);
                    stc(&S.x[v], nv);
                }
            }
            for (uint32_t i = gt; i < nt; i += nth) {
                const uint32_t v = ldc(&qshort(S, par)[i]);
                const int32_t old = ldc(&S.x[v]);
                uint32_t os, cs;
                const int32_t nv = evaluate<1>(A, S, v, old, k, red, os, cs, vis);
                stc(&S.os[v], (int32_t)os);
                stc(&S.cs[v], (int32_t)cs);
                stc(&S.oqs[i], old);
                stc(&S.x[v], nv);
            }
            barrier();
            // B. notify
            for (uint32_t i = tw; i < nl; i += tn) {
                const uint32_t v = ldc(&qlong(S, par)[i]);
                const int32_t old = ldc(&S.oql[i]), nv = ldc(&S.x[v]);
                if (nv < old) notify<kCBlk>(A, S, slots, npar, v, old, nv, k, ntag, vis);
            }
            for (uint32_t i = gw; i < nm; i += nw) {
                const uint32_t at = qmed(S, i);
                const uint32_t v = ldc(&qshort(S, par)[at]);
                const int32_t old = ldc(&S.oqs[at]), nv = ldc(&S.x[v]);
                if (nv < old) notify<W>(A, S, slots, npar, v, old, nv, k, ntag, vis);
            }
            for (uint32_t i = gt; i < nt; i += nth) {
                const uint32_t v = ldc(&qshort(S, par)[i]);
                const int32_t old = ldc(&S.oqs[i]), nv = ldc(&S.x[v]);
                if (nv < old) notify<1>(A, S, slots, npar, v, old, nv, k, ntag, vis);
            }
            barrier();
            tag = ntag;
            par = npar;
        }
        if (!ok) break;
        // 4. the row, and its maximum
        int32_t mx = 0;
        for (uint32_t r = gt; r < V; r += nth) {
            const int32_t xr = ldc(&S.x[r]);
            A.rows[(size_t)k * A.V + A.order[r]] = xr;
            mx = max(mx, xr);
        }
#pragma unroll
        for (int d = W / 2; d > 0; d >>= 1) mx = max(mx, __shfl_xor(mx, d, W));
        if (lane_id() == 0 && mx > 0) consume((uint32_t)atomicMax(&slots[kCqMax], (uint32_t)mx));
        vis = wave_sum64(vis);
        if (lane_id() == 0 && vis) consume((uint32_t)atomicAdd(visited_word(slots), vis));
        vis = 0;
        barrier();
        if (tw == 0 && tid == 0) {
            const int32_t lm = V ? (int32_t)coherent_read(&slots[kCqMax]) : 0;
            const uint64_t seen = atomicExch(visited_word(slots), 0ull);
            A.lmax_rows[k] = lm;
            if (A.stats) {
                uint32_t* st = A.stats + (size_t)kStatWords * k;
                const uint64_t dt = wall_clock64() - t_row;
                st[kStatPhases] = rounds;
                st[kStatLevels] = k == a + 1 && !preset ? (src < a ? 2u : 1u) : 0u;
                st[kStatTicksLo] = (uint32_t)dt;
                st[kStatTicksHi] = (uint32_t)(dt >> 32);
                st[kStatFailed] = 0;
                st[kStatRemoved] = V;
                st[kStatAlive] = (uint32_t)min<uint64_t>(seen, 0xffffffffull);
                st[kStatLevel] = (uint32_t)lm;
            }
        }
    }
    // the team's next pass reuses the scratch
    barrier();
    if (!ok && tid == 0) consume(atomicMax(A.err, (uint32_t)kErrTimeout));
    return ok;
}

}  // namespace chain
}  // namespace dcx
