// Localized maintenance of kmaxes and l-rows after an edge batch: only the
// vertices a batch can change are visited, instead of re-peeling the whole
// graph and recomputing every row k <= M+1 (the reference's maintenance,
// main.cu:1536-1751, recomputes the rows; its kmax maintenance,
// main.cu:668-855, marks candidates and refines them by h-index).
//
// Every task runs on one workgroup (no grid-wide barrier, no co-residency
// assumption) and works on one array of values -- the kmaxes, or row k --
// in place.  A value cell keeps four flag bits above the value while a
// task runs (all cleared before it ends):
//   kQueued  the vertex is on the task's next worklist;
//   kBusy    a wavefront is evaluating the vertex;
//   kRegion  the vertex may rise (insertion tasks);
//   kSeen    the vertex was tested as a candidate, or (kmax deletion) its
//            change was recorded.
// Values stay below 2^28 (the host falls back to recomputation otherwise).
//
// Refinement (all tasks): the h-index operator of the pass,
//     x(v) <- largest h <= x(v) with >= h K-neighbours of value >= h
//             and (rows, k > 0) >= k C-neighbours of value >= h,
// applied to a worklist until nothing changes.  From values that are upper
// bounds of the true ones it ends exactly at the true values (the operator
// is monotone and the true values are its greatest fixpoint); a vertex is
// re-queued when a neighbour it counted drops below its value.  K lists are
// out-lists for rows and in-lists for kmaxes; C lists are in-lists for rows
// k > 0 (non-members of the (k,0)-core hold 0, so they never count for
// h >= 1, and a member always meets the constraint at h = 0).
//
// Task types:
//   kmax deletion  seeds: targets of deleted edges (their in-degree fell);
//                  the old kmaxes are upper bounds.  Records every vertex
//                  whose kmax fell, with its old value.
//   row deletion   row k <= M_del: vertices that left the (k,0)-core get 0
//                  and queue their neighbours; endpoints of deleted edges
//                  inside the old core are seeds.
//   core insertion the kmaxes (in-core, K = in-lists) or row 0 (out-core, K
//                  = out-lists) after insertions, in rounds: round r makes
//                  the r-th inserted entry of every list active, so no
//                  vertex gains more than one K-entry per round and no value
//                  rises by more than one per round.  A round's candidates
//                  are found from the owners whose new entry can count
//                  (other endpoint's value >= the owner's), through vertices
//                  of the same value whose count of K-neighbours of at least
//                  that value (ED, main.cu:404-421) exceeds it; they get the
//                  upper bound value + 1 (kmaxFindUpperBound, main.cu:489-545)
//                  and are refined (kmaxRefineHIndex, main.cu:589-663).
//   row insertion  rows k = 1, 2, ... in order, after the kmaxes and row 0
//                  are final (each row bounded by the new row k - 1):
//                  a vertex whose value rises from a to a' lies, for every b
//                  in (a, a'], in a weakly connected set of vertices that all
//                  rise through b and that holds an endpoint of an inserted
//                  edge.  So candidates are searched from those endpoints,
//                  each carrying the interval (lo, hi] of levels its path can
//                  rise through (lo >= its old value, hi <= the new row
//                  k - 1); a candidate starts at its hi.
#include <hip/hip_runtime.h>

#include "dcx_internal.hpp"

namespace dcx {
namespace {

constexpr int W = 64;
constexpr int kBlk = 512;
constexpr int kWaves = kBlk / W;
constexpr uint32_t kQueued = 0x80000000u, kRegion = 0x40000000u, kSeen = 0x20000000u,
                   kBusy = 0x10000000u;
constexpr uint32_t kValMask = 0x0fffffffu;

__device__ __forceinline__ int lane() { return (int)(threadIdx.x & (W - 1)); }
__device__ __forceinline__ int wave() { return (int)(threadIdx.x / W); }

__device__ __forceinline__ uint32_t ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t val(const int32_t* x, uint32_t v) {
    return (int32_t)(ld(reinterpret_cast<const uint32_t*>(x) + v) & kValMask);
}
__device__ __forceinline__ uint32_t cell(const int32_t* x, uint32_t v) {
    return ld(reinterpret_cast<const uint32_t*>(x) + v);
}
__device__ __forceinline__ uint32_t* cellp(int32_t* x, uint32_t v) {
    return reinterpret_cast<uint32_t*>(x) + v;
}

// Sets the value bits of a cell, keeping its flags (other wavefronts may set
// flags concurrently).
__device__ __forceinline__ void set_value(int32_t* x, uint32_t v, uint32_t nv) {
    uint32_t* p = cellp(x, v);
    uint32_t c = ld(p);
    for (;;) {
        const uint32_t want = (c & ~kValMask) | nv;
        const uint32_t got = atomicCAS(p, c, want);
        if (got == c) return;
        c = got;
    }
}

// At least `need` of list[0, len) hold a value >= h?  Stops at the first
// 64-entry chunk where the count is reached.  Wave-uniform.
__device__ __forceinline__ bool wave_count_ge(const uint32_t* list, uint32_t len,
                                              const int32_t* x, int32_t h, int32_t need) {
    if (need <= 0) return true;
    int32_t c = 0;
    for (uint32_t base = 0; base < len; base += W) {
        const uint32_t i = base + lane();
        const bool ok = i < len && val(x, list[i]) >= h;
        c += __popcll(__ballot(ok));
        if (c >= need) return true;
    }
    return false;
}

struct Lists {
    const uint32_t* k;  // K list
    uint32_t klen;
    const uint32_t* c;  // C list (rows, k > 0)
    uint32_t clen;
    int32_t need_c;     // k (0: no constraint)
};

__device__ __forceinline__ bool holds(const Lists& L, const int32_t* x, int32_t h) {
    if (h <= 0) return true;
    if (!wave_count_ge(L.k, L.klen, x, h, h)) return false;
    return wave_count_ge(L.c, L.clen, x, h, L.need_c);
}

// The operator at one vertex: largest h <= cap that holds.  Wave-uniform.
__device__ int32_t wave_eval(const Lists& L, const int32_t* x, int32_t cap) {
    if (cap <= 0 || holds(L, x, cap)) return cap > 0 ? cap : 0;
    int32_t lo = 0, hi = cap - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (holds(L, x, mid))
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Per-workgroup worklists in global scratch (slot of kSlotLists * qcap words).
struct Slot {
    uint32_t* cur;   // [qcap]
    uint32_t* nxt;   // [qcap]
    uint32_t* reg;   // [2 * qcap] (vertex, old value) pairs; old = -1: rejected
    uint32_t* cand;  // [qcap]
    uint32_t qcap;
};

struct Shared {
    uint32_t task, ncur, nnxt, nreg, ncand, over, tmp;
};

// Appends x to list (counter in LDS) with one LDS atomic per wavefront.
__device__ __forceinline__ void wave_append(bool take, uint32_t x, uint32_t* list,
                                            uint32_t* count, uint32_t cap, uint32_t* over) {
    const unsigned long long m = __ballot(take);
    if (!m) return;
    uint32_t base = 0;
    const int leader = __ffsll(m) - 1;
    if (lane() == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader, W);
    if (take) {
        const uint32_t at = base + (uint32_t)__popcll(m & ((1ull << lane()) - 1));
        if (at < cap)
            list[at] = x;
        else
            atomicOr(over, 1u);
    }
}

// True (and the task marked as stopped) once the work counters of every
// task type together exceed the budget.  Read by thread 0 only; callers
// broadcast through sh.over after a barrier.
__device__ __forceinline__ void check_budget(const LocalArgs& A, Shared& sh) {
    if (threadIdx.x != 0) return;
    bool over = false;
    if (A.budget && A.counters) {
        unsigned long long used = 0;
        for (int t = 1; t <= 4; ++t)
            used += __hip_atomic_load(A.counters + 4 * t + 2, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) +
                    __hip_atomic_load(A.counters + 4 * t + 3, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        over = used > A.budget;
    }
    if (A.ticks && A.t_start) {
        const unsigned long long now = wall_clock64();
        const unsigned long long prev = atomicCAS(A.t_start, 0ull, now);
        over = over || (prev != 0 && now - prev > A.ticks);
    }
    if (over) {
        sh.over = 1;
        atomicOr(A.fail, 0x80u);
    }
}

// Graph views and the task inputs.
struct TaskCtx {
    const LocalArgs* A;
    int32_t* x;        // the task's values
    int32_t k;         // row, or -1 for kmaxes
    bool kside;        // K = in-lists (kmaxes)
    const int32_t* ins_cnt;  // core insertion rounds: inserted entries per K list
    int32_t round;     // active prefix: deg - cnt + min(cnt, round + 1); -1: whole lists
    bool record;       // kmax deletion: record (vertex, old value) at its first change
    int32_t type;      // kLocal* (for the counters)
};

__device__ __forceinline__ uint32_t klen_of(const TaskCtx& T, uint32_t v, uint32_t deg) {
    if (T.round < 0) return deg;
    const int32_t c = T.ins_cnt[v];
    return deg - (uint32_t)c + (uint32_t)min(c, T.round + 1);
}

__device__ __forceinline__ Lists lists_of(const TaskCtx& T, uint32_t v) {
    const LocalArgs& A = *T.A;
    Lists L;
    const Csr& K = T.kside ? A.in : A.out;
    const uint32_t kd = (uint32_t)K.deg[v];
    L.k = K.nbr + K.off[v];
    L.klen = klen_of(T, v, kd);
    L.c = nullptr;
    L.clen = 0;
    L.need_c = 0;
    if (T.k > 0) {
        L.c = A.in.nbr + A.in.off[v];
        L.clen = (uint32_t)A.in.deg[v];
        L.need_c = T.k;
    }
    return L;
}

__device__ __forceinline__ bool member(const TaskCtx& T, uint32_t v) {
    return T.k <= 0 || val(T.A->kmaxes, v) >= T.k;
}

// Queues the neighbours of v (just lowered from old to nv) that may have
// counted it: members (in the region, when only candidates may move) with
// nv < value <= old -- a neighbour of a higher value did not count v at its
// value -- and also, whatever their value above nv, those being evaluated
// right now (kBusy): such a neighbour may still show its previous, higher
// value while it computes a lower one from v's old value.  Values only
// fall, so a neighbour at most nv stays unaffected.  Lists: for rows,
// in-lists (their K lists hold v) and out-lists (their C lists hold v,
// k > 0); for kmaxes out-lists; for row 0 (k = 0) in-lists.
__device__ void push_neighbours(const TaskCtx& T, uint32_t v, int32_t nv, int32_t old,
                                bool region_only, Slot& S, Shared& sh) {
    const LocalArgs& A = *T.A;
    for (int side = 0; side < 2; ++side) {
        const Csr* R = nullptr;
        if (T.kside) {
            if (side == 0) R = &A.out;
        } else if (side == 0) {
            R = &A.in;
        } else if (T.k > 0) {
            R = &A.out;
        }
        if (!R) continue;
        const uint32_t* lst = R->nbr + R->off[v];
        const uint32_t n = (uint32_t)R->deg[v];
        for (uint32_t base = 0; base < n; base += W) {
            const uint32_t i = base + lane();
            bool take = false;
            uint32_t y = 0;
            if (i < n) {
                y = lst[i];
                const uint32_t c = cell(T.x, y);
                const int32_t vy = (int32_t)(c & kValMask);
                if (vy > nv && (vy <= old || (c & kBusy)) && !(c & kQueued) &&
                    (!region_only || (c & kRegion)) && member(T, y)) {
                    const uint32_t prev = atomicOr(cellp(T.x, y), kQueued);
                    take = !(prev & kQueued);
                }
            }
            wave_append(take, y, S.nxt, &sh.nnxt, S.qcap, &sh.over);
        }
    }
}

// Refines the worklist S.cur[0, sh.ncur) to a fixpoint.  Every entry must
// carry kQueued.  Returns false on worklist overflow.
__device__ bool refine(const TaskCtx& T, bool region_only, Slot& S, Shared& sh) {
    for (;;) {
        check_budget(*T.A, sh);
        __syncthreads();
        const uint32_t n = sh.ncur;
        if (n == 0 || sh.over) break;
        if (threadIdx.x == 0) sh.nnxt = 0;
        __syncthreads();
        for (uint32_t i = wave(); i < n; i += kWaves) {
            const uint32_t v = S.cur[i];
            uint32_t c = 0;
            if (lane() == 0) {
                c = atomicAnd(cellp(T.x, v), ~kQueued);
                atomicOr(cellp(T.x, v), kBusy);
            }
            c = __shfl(c, 0, W);
            const int32_t old = (int32_t)(c & kValMask);
            const Lists L = lists_of(T, v);
            const int32_t nv = wave_eval(L, T.x, old);
            if (lane() == 0 && T.A->counters) {
                atomicAdd(T.A->counters + 4 * T.type, 1ull);
                atomicAdd(T.A->counters + 4 * T.type + 2, (unsigned long long)(L.klen + L.clen));
            }
            if (nv < old) {
                if (lane() == 0) {
                    if (T.record && !(atomicOr(cellp(T.x, v), kSeen) & kSeen)) {
                        const uint32_t at = atomicAdd(T.A->n_changed, 1u);
                        if (at < T.A->changed_cap) {
                            T.A->changed[2 * at] = v;
                            T.A->changed[2 * at + 1] = (uint32_t)old;
                        } else {
                            atomicOr(&sh.over, 1u);
                        }
                    }
                    set_value(T.x, v, (uint32_t)nv);
                    atomicAnd(cellp(T.x, v), ~kBusy);
                }
                push_neighbours(T, v, nv, old, region_only, S, sh);
            } else if (lane() == 0) {
                atomicAnd(cellp(T.x, v), ~kBusy);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            sh.ncur = min(sh.nnxt, S.qcap);
        }
        uint32_t* t = S.cur;
        S.cur = S.nxt;
        S.nxt = t;
    }
    __syncthreads();
    return sh.over == 0;
}

// Queues v for refinement (kQueued) if it is not queued yet.
__device__ __forceinline__ bool try_queue(int32_t* x, uint32_t v) {
    return !(atomicOr(cellp(x, v), kQueued) & kQueued);
}

// ---- deletions -----------------------------------------------------------

__device__ void task_kmax_del(const LocalArgs& A, Slot& S, Shared& sh) {
    TaskCtx T{&A, A.kmaxes, -1, true, nullptr, -1, true, kLocalKmaxDel};
    if (threadIdx.x == 0) sh.ncur = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < A.n_del_targets; i += kBlk) {
        const uint32_t v = A.del_targets[i];
        const bool take = try_queue(A.kmaxes, v);
        if (take) {
            const uint32_t at = atomicAdd(&sh.ncur, 1u);
            if (at < S.qcap)
                S.cur[at] = v;
            else
                atomicOr(&sh.over, 1u);
        }
    }
    // every lowered value is recorded once with its old value (kSeen marks
    // the recorded ones): the row tasks need the vertices leaving each core
    refine(T, false, S, sh);
    const uint32_t n = min(ld(A.n_changed), A.changed_cap);
    for (uint32_t i = threadIdx.x; i < n; i += kBlk)
        atomicAnd(cellp(A.kmaxes, A.changed[2 * i]), ~kSeen);
}

__device__ void task_row_del(const LocalArgs& A, int32_t k, Slot& S, Shared& sh) {
    int32_t* x = A.rows + (size_t)k * A.V;
    TaskCtx T{&A, x, k, false, nullptr, -1, false, kLocalRowDel};
    if (threadIdx.x == 0) sh.ncur = sh.nnxt = 0;
    __syncthreads();
    // vertices that left the (k,0)-core: value 0, their neighbours queued
    const uint32_t nch = min(ld(A.n_changed), A.changed_cap);
    for (uint32_t i = wave(); i < nch; i += kWaves) {
        const uint32_t v = A.changed[2 * i];
        const int32_t oldk = (int32_t)A.changed[2 * i + 1];
        if (!(oldk >= k && val(A.kmaxes, v) < k)) continue;
        uint32_t c = 0;
        if (lane() == 0) c = atomicAnd(cellp(x, v), ~kValMask);  // value 0
        c = __shfl(c, 0, W);
        const int32_t old = (int32_t)(c & kValMask);
        if (old > 0) push_neighbours(T, v, 0, old, false, S, sh);  // into nxt
    }
    __syncthreads();
    // pushes went to nxt: move them to cur
    if (threadIdx.x == 0) {
        sh.ncur = min(sh.nnxt, S.qcap);
        sh.nnxt = 0;
    }
    __syncthreads();
    {
        uint32_t* t = S.cur;
        S.cur = S.nxt;
        S.nxt = t;
    }
    // endpoints of deleted edges that were inside the old core
    for (uint32_t i = threadIdx.x; i < A.n_del; i += kBlk) {
        if (A.del_m[i] < k) continue;
        for (int e = 0; e < 2; ++e) {
            const uint32_t v = A.del_pairs[2 * i + e];
            if (val(A.kmaxes, v) < k) continue;
            if (try_queue(x, v)) {
                const uint32_t at = atomicAdd(&sh.ncur, 1u);
                if (at < S.qcap)
                    S.cur[at] = v;
                else
                    atomicOr(&sh.over, 1u);
            }
        }
    }
    refine(T, false, S, sh);
}


// ---- insertions: kmaxes and row 0, in rounds -----------------------------

// side 0: kmaxes (K = in-lists, grouped by target); side 1: row 0 (K =
// out-lists, grouped by source).
__device__ void task_core_ins(const LocalArgs& A, int side, Slot& S, Shared& sh) {
    const bool kside = side == 0;
    int32_t* x = kside ? A.kmaxes : A.rows;
    const uint32_t* gkey = kside ? A.gin_key : A.gout_key;
    const uint32_t* gcnt = kside ? A.gin_cnt : A.gout_cnt;
    const uint32_t ng = kside ? A.n_gin : A.n_gout;
    const Csr& K = kside ? A.in : A.out;
    const Csr& R = kside ? A.out : A.in;
    TaskCtx T{&A, x, kside ? -1 : 0, kside, kside ? A.ins_cnt_in : A.ins_cnt_out, 0, false,
              kLocalCoreIns};
    if (threadIdx.x == 0) sh.tmp = 0;
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < ng; g += kBlk) atomicMax(&sh.tmp, gcnt[g]);
    __syncthreads();
    const int32_t rounds = (int32_t)sh.tmp;
    for (int32_t r = 0; r < rounds; ++r) {
        T.round = r;
        __syncthreads();
        if (threadIdx.x == 0) sh.ncur = sh.nnxt = sh.nreg = sh.ncand = 0;
        __syncthreads();
        // owners whose r-th inserted entry can count: the other endpoint's
        // value is at least the owner's (otherwise the owner cannot rise by
        // it: the other endpoint rises by at most one this round)
        for (uint32_t g = threadIdx.x; g < ng; g += kBlk) {
            const uint32_t o = gkey[g], c = gcnt[g];
            if (c <= (uint32_t)r) continue;
            const uint32_t other = K.nbr[K.off[o] + (uint32_t)K.deg[o] - c + (uint32_t)r];
            if (val(x, other) < val(x, o)) continue;
            if (!(atomicOr(cellp(x, o), kSeen) & kSeen)) {
                const uint32_t at = atomicAdd(&sh.ncand, 1u);
                if (at < S.qcap)
                    S.cand[at] = o;
                else
                    atomicOr(&sh.over, 1u);
            }
        }
        __syncthreads();
        // breadth-first through vertices of the same value: test the
        // candidates (ED: K-neighbours of at least the value, active
        // entries only, must exceed it), then expand the accepted ones
        for (;;) {
            check_budget(A, sh);
            __syncthreads();
            const uint32_t nc = min(sh.ncand, S.qcap);
            if (nc == 0 || sh.over) break;
            for (uint32_t i = wave(); i < nc; i += kWaves) {
                const uint32_t y = S.cand[i];
                const int32_t vy = val(x, y);
                const Lists L = lists_of(T, y);
                const bool ok = wave_count_ge(L.k, L.klen, x, vy, vy + 1);
                if (lane() == 0) {
                    const uint32_t at = atomicAdd(&sh.nreg, 1u);
                    if (at < S.qcap) {
                        S.reg[2 * at] = y;
                        S.reg[2 * at + 1] = ok ? (uint32_t)vy : 0xffffffffu;
                    } else {
                        atomicOr(&sh.over, 1u);
                    }
                    if (ok) {
                        if (A.counters) {
                            atomicAdd(A.counters + 4 * kLocalCoreIns + 1, 1ull);
                            atomicAdd(A.counters + 4 * kLocalCoreIns + 3,
                                      (unsigned long long)R.deg[y]);
                        }
                        atomicOr(cellp(x, y), kRegion);
                        const uint32_t f = atomicAdd(&sh.nnxt, 1u);
                        if (f < S.qcap)
                            S.nxt[f] = y;
                        else
                            atomicOr(&sh.over, 1u);
                    }
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) sh.ncand = 0;
            __syncthreads();
            const uint32_t nf = min(sh.nnxt, S.qcap);
            for (uint32_t i = wave(); i < nf; i += kWaves) {
                const uint32_t z = S.nxt[i];
                const int32_t vz = val(x, z);
                const uint32_t* lst = R.nbr + R.off[z];
                const uint32_t n = (uint32_t)R.deg[z];
                for (uint32_t base = 0; base < n; base += W) {
                    const uint32_t j = base + lane();
                    bool take = false;
                    uint32_t y = 0;
                    if (j < n) {
                        y = lst[j];
                        const uint32_t cy = cell(x, y);
                        if ((int32_t)(cy & kValMask) == vz && !(cy & kSeen))
                            take = !(atomicOr(cellp(x, y), kSeen) & kSeen);
                    }
                    wave_append(take, y, S.cand, &sh.ncand, S.qcap, &sh.over);
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) sh.nnxt = 0;
            __syncthreads();
        }
        if (sh.over) return;
        // the accepted candidates start at value + 1 and are refined among
        // themselves (everything else keeps its value this round)
        const uint32_t nr = min(sh.nreg, S.qcap);
        for (uint32_t i = threadIdx.x; i < nr; i += kBlk) {
            const uint32_t y = S.reg[2 * i], a = S.reg[2 * i + 1];
            if (a == 0xffffffffu) continue;
            set_value(x, y, a + 1);
            atomicOr(cellp(x, y), kQueued);
            const uint32_t at = atomicAdd(&sh.ncur, 1u);
            S.cur[at] = y;  // at < nr <= qcap
        }
        refine(T, true, S, sh);
        if (sh.over) return;
        for (uint32_t i = threadIdx.x; i < nr; i += kBlk)
            atomicAnd(cellp(x, S.reg[2 * i]), kValMask);
        __syncthreads();
    }
}

// ---- insertions: rows k >= 1 ---------------------------------------------

// Offers y a path interval (lo, h]: lo = the path's old value, h = its
// label so far.  A first claim records (y, old value); a higher label on a
// claimed vertex re-queues it for expansion (with lo = -1, looser and still
// a bound).  Lane-level; returns whether y goes to the next frontier, and
// its frontier lo.
__device__ __forceinline__ bool offer(int32_t* x, const int32_t* U, const int32_t* km, int32_t k,
                                      uint32_t y, int32_t lo, int32_t h, Slot& S, Shared& sh,
                                      int32_t* lo_out, unsigned long long* T_counters) {
    if (val(km, y) < k) return false;
    const int32_t h2 = min(h, val(U, y));
    uint32_t c = cell(x, y);
    for (;;) {
        if (c & kRegion) {
            if ((int32_t)(c & kValMask) >= h2) return false;
            const uint32_t prev = atomicMax(cellp(x, y), kRegion | (uint32_t)h2);
            if ((int32_t)(prev & kValMask) >= h2) return false;
            *lo_out = -1;
            return true;
        }
        const int32_t a = (int32_t)(c & kValMask);
        if (max(lo, a) >= h2) return false;
        const uint32_t got = atomicCAS(cellp(x, y), c, kRegion | (uint32_t)h2);
        if (got == c) {
            const uint32_t at = atomicAdd(&sh.nreg, 1u);
            if (T_counters) atomicAdd(T_counters + 4 * kLocalRowIns + 1, 1ull);
            if (at < S.qcap) {
                S.reg[2 * at] = y;
                S.reg[2 * at + 1] = (uint32_t)a;
            } else {
                atomicOr(&sh.over, 1u);
            }
            *lo_out = a;
            return true;
        }
        c = got;
    }
}

// Frontier entries are (vertex, lo) pairs: cur/nxt hold qcap / 2 of them.
// U bounds every new value of the row: the new row 0 (D_{k,l} is a subset
// of D_{0,l}, so l_k <= l_0), which is final before any row task starts, so
// that the rows k >= 1 run as independent tasks.
__device__ void task_row_ins(const LocalArgs& A, int32_t k, Slot& S, Shared& sh) {
    int32_t* x = A.rows + (size_t)k * A.V;
    const int32_t* U = A.rows;
    TaskCtx T{&A, x, k, false, nullptr, -1, false, kLocalRowIns};
    const uint32_t fcap = S.qcap / 2;
    if (threadIdx.x == 0) sh.ncur = sh.nnxt = sh.nreg = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 2 * A.n_ins; i += kBlk) {
        const uint32_t e = i >> 1;
        const uint32_t u = A.ins_pairs[2 * e], v = A.ins_pairs[2 * e + 1];
        if (val(A.kmaxes, u) < k || val(A.kmaxes, v) < k) continue;
        const uint32_t y = (i & 1) ? v : u;
        int32_t lo = -1;
        if (offer(x, U, A.kmaxes, k, y, -1, (int32_t)kValMask, S, sh, &lo, A.counters)) {
            const uint32_t at = atomicAdd(&sh.ncur, 1u);
            if (at < fcap) {
                S.cur[2 * at] = y;
                S.cur[2 * at + 1] = (uint32_t)lo;
            } else {
                atomicOr(&sh.over, 1u);
            }
        }
    }
    for (;;) {
        check_budget(A, sh);
        __syncthreads();
        const uint32_t nf = min(sh.ncur, fcap);
        if (nf == 0 || sh.over) break;
        if (threadIdx.x == 0) sh.nnxt = 0;
        __syncthreads();
        for (uint32_t i = wave(); i < nf; i += kWaves) {
            const uint32_t z = S.cur[2 * i];
            const int32_t lo = (int32_t)S.cur[2 * i + 1];
            const int32_t hz = val(x, z);
            if (lane() == 0 && A.counters)
                atomicAdd(A.counters + 4 * kLocalRowIns + 3,
                          (unsigned long long)(A.out.deg[z] + A.in.deg[z]));
            for (int side = 0; side < 2; ++side) {
                const Csr& R = side == 0 ? A.out : A.in;
                const uint32_t* lst = R.nbr + R.off[z];
                const uint32_t n = (uint32_t)R.deg[z];
                for (uint32_t base = 0; base < n; base += W) {
                    const uint32_t j = base + lane();
                    bool take = false;
                    uint32_t y = 0;
                    int32_t ylo = -1;
                    if (j < n) {
                        y = lst[j];
                        take = offer(x, U, A.kmaxes, k, y, lo, hz, S, sh, &ylo, A.counters);
                    }
                    const unsigned long long m = __ballot(take);
                    if (m) {
                        uint32_t base2 = 0;
                        const int leader = __ffsll(m) - 1;
                        if (lane() == leader) base2 = atomicAdd(&sh.nnxt, (uint32_t)__popcll(m));
                        base2 = __shfl(base2, leader, W);
                        if (take) {
                            const uint32_t at =
                                base2 + (uint32_t)__popcll(m & ((1ull << lane()) - 1));
                            if (at < fcap) {
                                S.nxt[2 * at] = y;
                                S.nxt[2 * at + 1] = (uint32_t)ylo;
                            } else {
                                atomicOr(&sh.over, 1u);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) sh.ncur = min(sh.nnxt, fcap);
        uint32_t* t = S.cur;
        S.cur = S.nxt;
        S.nxt = t;
    }
    __syncthreads();
    if (sh.over) return;
    // every candidate starts at its label and is refined among the others
    const uint32_t nr = min(sh.nreg, S.qcap);
    if (threadIdx.x == 0) sh.ncur = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nr; i += kBlk) {
        const uint32_t y = S.reg[2 * i];
        atomicOr(cellp(x, y), kQueued);
        const uint32_t at = atomicAdd(&sh.ncur, 1u);
        S.cur[at] = y;  // at < nr <= qcap
    }
    refine(T, true, S, sh);
    if (sh.over) return;
    for (uint32_t i = threadIdx.x; i < nr; i += kBlk)
        atomicAnd(cellp(x, S.reg[2 * i]), kValMask);
}

__global__ __launch_bounds__(kBlk) void k_local_tasks(LocalArgs A) {
    __shared__ Shared sh;
    const size_t words = local_slot_words(A.qcap);
    uint32_t* base = A.scratch + (size_t)blockIdx.x * words;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            sh.task = atomicAdd(A.task_ctr, 1u);
            sh.ncur = sh.nnxt = sh.nreg = sh.ncand = sh.over = sh.tmp = 0;
        }
        __syncthreads();
        const uint32_t t = sh.task;
        if (t >= A.ntasks) break;
        Slot S{base, base + A.qcap, base + 2 * (size_t)A.qcap, base + 4 * (size_t)A.qcap,
               A.qcap};
        const LocalTaskDesc d = A.tasks[t];
        switch (d.type) {
            case kLocalKmaxDel: task_kmax_del(A, S, sh); break;
            case kLocalRowDel: task_row_del(A, d.k, S, sh); break;
            case kLocalCoreIns: task_core_ins(A, d.k, S, sh); break;
            case kLocalRowIns: task_row_ins(A, d.k, S, sh); break;
            default: break;
        }
        __syncthreads();
        if (threadIdx.x == 0 && sh.over) atomicOr(A.fail, 1u << (d.type & 7));
    }
}

// Per-edge min of the endpoints' kmaxes (deletions, before anything changes).
__global__ void k_edge_min_kmax(const int32_t* kmaxes, const uint32_t* pairs, uint32_t n,
                                int32_t* out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = min(kmaxes[pairs[2 * i]], kmaxes[pairs[2 * i + 1]]);
}

// cnt[key[g]] = c[g] (set) or 0 (clear).
__global__ void k_set_counts(int32_t* cnt, const uint32_t* key, const uint32_t* c, uint32_t n,
                             int set) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x)
        cnt[key[g]] = set ? (int32_t)c[g] : 0;
}

// Largest value of x (flags masked).
__global__ void k_max_masked(const int32_t* x, uint32_t n, int32_t* out) {
    int32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = max(m, (int32_t)((uint32_t)x[i] & kValMask));
    for (int d = W / 2; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, W));
    if (lane() == 0) atomicMax(out, m);
}

unsigned small_grid(uint32_t n) { return (unsigned)max(1u, min((n + 255u) / 256u, 4096u)); }

}  // namespace



hipError_t launch_local_tasks(const LocalArgs& a, uint32_t blocks, hipStream_t st) {
    if (a.ntasks == 0 || blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_local_tasks, dim3(blocks), dim3(kBlk), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_edge_min_kmax(const int32_t* kmaxes, const uint32_t* pairs, uint32_t n,
                                int32_t* out, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_edge_min_kmax, dim3(small_grid(n)), dim3(256), 0, st, kmaxes, pairs, n,
                       out);
    return hipGetLastError();
}

hipError_t launch_set_counts(int32_t* cnt, const uint32_t* key, const uint32_t* c, uint32_t n,
                             bool set, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_set_counts, dim3(small_grid(n)), dim3(256), 0, st, cnt, key, c, n,
                       (int)set);
    return hipGetLastError();
}

hipError_t launch_max_masked(const int32_t* x, uint32_t n, int32_t* out, hipStream_t st) {
    hipError_t e = hipMemsetAsync(out, 0, 4, st);
    if (e != hipSuccess || !n) return e;
    hipLaunchKernelGGL(k_max_masked, dim3(small_grid(n)), dim3(256), 0, st, x, n, out);
    return hipGetLastError();
}

}  // namespace dcx
