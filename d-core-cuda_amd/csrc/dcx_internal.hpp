// Internal declarations shared by the dcx HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "dcx.h"
#include "grid_sync.hpp"

namespace dcx {

// One adjacency direction in device memory: the list of v starts at off[v]
// and holds deg[v] entries.
struct Csr {
    const uint32_t* off;
    const int32_t* deg;
    const uint32_t* nbr;
};

// One peel pass (a single k of one graph, or the kmax pass).
//   K lists decide the level: a vertex leaves at level L when fewer than
//     L+1 of its K-neighbours are alive (out-lists for an l-row).
//   C lists decide membership: a vertex leaves as soon as fewer than k of
//     its C-neighbours are alive (in-lists for an l-row).
// The kmax pass is the same with the two directions swapped and k = 0
// (swapInOut, src/common.h:114-120).
//
// Two layouts:
//   full (core == false): original ids, whole lists, every vertex is a
//     member; lval is the output array itself (old ids).
//   core (core == true): ids are ranks in the kmax-descending order, lists
//     sorted by rank, members are [0, nmembers) and the members inside a
//     member's list are its prefix; lval is scratch in rank space and the
//     pass ends by writing out_row[order[i]] = lval[i].
struct PassSpec {
    uint32_t V;
    Csr K, C;
    bool core;
    const uint32_t* order;  // core: rank -> original id
    uint32_t nmembers;
    int32_t k;
    int32_t* out_row;       // core: [V] output row in original ids (non-members hold 0)
    int32_t* lval;          // full: [V] output; core: ignored (scratch used)
    uint32_t* stats;        // nullable; [kStatWords] written by the pass leader at the end
    uint32_t* trace;        // nullable; phase trace: [0] = entries, then kTraceWords each
    uint32_t trace_cap;     // entries the trace can hold
    int32_t trace_k;        // batched launch: the row whose team pass is traced
    uint32_t* fail_rec;     // nullable; [kFailWords] what the first failing member saw
    // A cascade phase pulls (every alive vertex counts its neighbours in the
    // frontier) instead of pushing decrements from the frontier when the
    // alive vertices' lists hold at most pull_ratio times the frontier's
    // edges; 0 = always push.
    uint32_t pull_ratio;
    // Nullable: per vertex (rank in the core layout, id in the full one) an
    // upper bound of the row's value.  A vertex still alive at level L with
    // ub <= L cannot be in the (k, L+1)-core, so it is seeded at L directly
    // instead of waiting for a cascade to reach it; the values are
    // unchanged.
    const int32_t* ub;
};

enum : uint32_t { kFailWords = 16 };

// Phase-counter slots of a peel pass: 3 slots, each in its own 128-byte
// line (kSlotStride words), kSlotWords words per pass or team.
// Words of a slot: [0] frontier appends, [1] alive-list appends, [2] smallest
// key degree above the level, [3] the leader's phase, [4..5] edges of the
// frontier's lists (uint64), [6..7] edges of the alive lists (uint64, first
// level only), [8] frontier vertices with long lists (kept apart), [9] alive
// vertices with long lists (first level only).
enum : uint32_t { kSlotStride = 32, kSlotWords = 3 * kSlotStride };

// Phase trace entry (peel passes run on one team when tracing is on).
enum : uint32_t {
    kTraceLevel = 0,    // level L (low 30 bits), kTraceSeedBit, kTracePullBit
    kTraceFrontier = 1, // vertices removed by the phase
    kTraceAlive = 2,    // alive-list length after the phase
    kTraceTicks = 3,    // wall-clock ticks since the previous phase ended
    kTraceWords = 4
};

constexpr uint32_t kTraceSeedBit = 0x80000000u;  // a seeding phase
constexpr uint32_t kTracePullBit = 0x40000000u;  // a cascade phase run as a pull
constexpr uint32_t kTraceLevelMask = 0x3fffffffu;

// Per-pass log written by the device (one entry per pass).
enum : uint32_t {
    kStatPhases = 0,  // synchronised phases (peel) or sweeps (h-index)
    kStatLevels = 1,  // level iterations (peel); 0 for h-index
    kStatTicksLo = 2, // wall-clock ticks from start to end of the pass
    kStatTicksHi = 3,
    kStatFailed = 4,  // nonzero: the pass stopped on inconsistent counters
    kStatRemoved = 5, // vertices removed when it ended
    kStatAlive = 6,   // alive-list length when it ended
    kStatLevel = 7,   // level when it ended
    kStatWords = 8
};

// Per-pass device scratch, sized for `cap` vertices.
struct PassScratch {
    int32_t* kd;   // live key degree (poisoned once removed)
    int32_t* cd;   // live constraint degree (poisoned once removed)
    int32_t* lv;   // core layout: level per rank (-1 alive)
    uint32_t* rk;  // core layout: length of the member prefix of the K list
    uint32_t* rc;  // core layout: length of the member prefix of the C list
    uint32_t* f0;  // frontier ping
    uint32_t* f1;  // frontier pong
    uint32_t* a0;  // alive list ping
    uint32_t* a1;  // alive list pong
    uint32_t* st;  // phase in which a vertex was removed (~0u while alive)
    uint32_t* fl;  // frontier vertices with long lists, two halves (ping-pong)
    uint32_t* lp;  // alive vertices with long lists, then their partial
                   // pull counts (key side, constraint side): quarters
};
constexpr int kScratchArrays = 12;

__host__ __device__ inline PassScratch scratch_at(void* base, uint64_t cap) {
    uint32_t* p = reinterpret_cast<uint32_t*>(base);
    PassScratch s;
    s.kd = reinterpret_cast<int32_t*>(p);
    s.cd = reinterpret_cast<int32_t*>(p + cap);
    s.lv = reinterpret_cast<int32_t*>(p + 2 * cap);
    s.rk = p + 3 * cap;
    s.rc = p + 4 * cap;
    s.f0 = p + 5 * cap;
    s.f1 = p + 6 * cap;
    s.a0 = p + 7 * cap;
    s.a1 = p + 8 * cap;
    s.st = p + 9 * cap;
    s.fl = p + 10 * cap;
    s.lp = p + 11 * cap;
    return s;
}

// Team of workgroups running a list of passes (batched rows).
struct TeamDesc {
    uint32_t xcd;          // XCD all of the team's workgroups run on
    uint32_t size;         // workgroups in the team
    uint32_t pass_begin;   // passes team_ks[pass_begin .. pass_end)
    uint32_t pass_end;
    uint64_t scratch_off;  // uint32 words into the scratch pool
    uint64_t scratch_cap;  // vertices the team's scratch holds
    uint32_t sync_off;     // words into the sync pool (sync_words(size))
    uint32_t slot_off;     // words into the slot pool (kSlotWords)
};

// Workgroups are dispatched to the XCDs round-robin: workgroup b runs on
// XCD b % kXcds.  A team is built from workgroups of one XCD, so the data
// its members hand each other between phases stays within one L2.
constexpr uint32_t kXcds = 8;

// Default pull_ratio of the peel passes (PassSpec::pull_ratio): push only.
// With removed neighbours skipped, dense-core decrements combined in LDS and
// long lists sliced across the team, pushes beat pulls on every graph
// measured (R-MAT 20: 48.3 ms push only, 49.9 ms at ratio 4; Chung-Lu 2^22:
// 127 ms against 151 ms); pulls stay available as an exact alternative.
constexpr uint32_t kDefaultPullRatio = 0;

// ---- cost-model MLP (mlp_kernels.hip) ----------------------------------
constexpr int kMlpIn = 8, kMlpHidden = 32;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpHidden + 1;
// y[s] = f(X[s]) for n samples of kMlpIn floats.
hipError_t launch_mlp_forward(const float* p, const float* X, uint32_t n, float* y,
                              hipStream_t st);
// grad = sum over samples of d/dp 0.5 (f(X[s]) - T[s])^2; loss = that sum.
hipError_t launch_mlp_backward(const float* p, const float* X, const float* T, uint32_t n,
                               float* grad, float* loss, hipStream_t st);
// One Adam step (beta 0.9 / 0.999) on grad / n, step number t >= 1.
hipError_t launch_mlp_adam(float* p, const float* grad, float* m, float* v, uint32_t n, float lr,
                           int t, hipStream_t st);

// ---- peel kernels (peel_kernels.hip) ----------------------------------
// One pass on the whole grid (bounded team barrier); `slots` 12 words,
// gs.w zeroed by the launcher.
// Two grid passes side by side in one launch (gs0.nblocks + gs1.nblocks
// workgroups, all resident).
hipError_t launch_pass_pair(const PassSpec& p0, const PassScratch& s0, uint32_t* slots0,
                            int32_t* lmax0, const GridSync& gs0, const PassSpec& p1,
                            const PassScratch& s1, uint32_t* slots1, int32_t* lmax1,
                            const GridSync& gs1, hipStream_t st);
// out[r] = in[order[r]] for r < V.
hipError_t launch_gather_rank(const int32_t* in, const uint32_t* order, uint32_t V, int32_t* out,
                              hipStream_t st);
hipError_t launch_pass_grid(const PassSpec& p, const PassScratch& s, uint32_t* slots,
                            int32_t* lmax_out, const GridSync& gs, hipStream_t st);
int pass_grid_blocks_per_cu();
// Workgroups of the team rows kernel that fit one CU at once.
int rows_teams_blocks_per_cu();
// Rows of one graph (core layout): the launch has `blocks` workgroups;
// block_team[b] names the team of block b (~0u: the block has no work) and
// block_team[blocks + b] its rank in the team.  Pass p of a team
// (teams[t].pass_begin <= p < pass_end) peels row team_ks[p] when
// team_kb[p] < 0, and otherwise derives rows team_ks[p]+1 .. team_kb[p] from
// row team_ks[p] (row 0, or an anchor an earlier pass of the team peeled),
// or, with kChainFromRow0 set in team_kb[p], starting from row 0's values
// (an upper bound of every row); a team's scratch then holds 12 arrays of
// scratch_cap >= V_{team_ks[p]} words.  base describes the graph (k, nmembers, out_row set per pass;
// stats: row k's log words at stats + kStatWords * k).
constexpr int32_t kChainFromRow0 = 1 << 30;
hipError_t launch_rows_teams(const PassSpec& base, const uint32_t* vk, int32_t* rows,
                             int32_t* lmax_rows, const TeamDesc* teams, const uint32_t* block_team,
                             const int32_t* team_ks, const int32_t* team_kb, uint32_t blocks,
                             void* scratch, uint32_t* sync_pool, uint32_t* slot_pool,
                             uint64_t timeout_ticks, uint32_t* err, uint32_t shrink_at,
                             uint32_t chain0_preset, hipStream_t st);
// Sets up row 0's chain (rows 1.. derived from row 0) on the whole grid:
// values, member prefixes and supports for row 1 in the chain state of its
// team (team_scratch: 12 arrays of cap >= V words; slots: the team's slot
// words, zeroed; e1: E_1, which sizes the slices of long lists).  Row 0
// must be in rows; the team launch then runs that chain with
// chain0_preset = 1.
hipError_t launch_chain_preset(const PassSpec& base, const uint32_t* vk, int32_t* rows,
                               uint32_t* team_scratch, uint64_t cap, uint32_t* slots, uint32_t V,
                               uint64_t e1, hipStream_t st);
// Rows whose core has at most `cap` (<= kLdsMaxMembers) members, one
// LDS-resident workgroup each, taken longest first from ks[0..nks) through
// the counter *queue (zero at launch).  Scratch: blocks * kScratchArrays *
// cap words.
constexpr uint32_t kLdsMaxMembers = 32767;
hipError_t launch_rows_lds(const PassSpec& base, const uint32_t* vk, int32_t* rows,
                           int32_t* lmax_rows, const int32_t* ks, uint32_t nks, uint32_t* queue,
                           uint32_t blocks, void* scratch, uint32_t cap, uint32_t* err,
                           hipStream_t st);
// Static LDS of the LDS-resident rows kernel (the dynamic part comes on top).
uint32_t rows_lds_static_bytes();
hipError_t launch_order_by_kmax(const int32_t* kmaxes, uint32_t V, int32_t kmax,
                                uint32_t* hist /*[kmax+2]*/, uint32_t* vk /*[kmax+2]*/,
                                uint32_t* order, hipStream_t st);
// Zeroes rows[k * V + v] for k in (knew[v], min(kold[v], rows_old - 1)]
// (rows_new: the new row count; rows from rows_old on are the caller's).
hipError_t launch_clear_stale(const int32_t* knew, const int32_t* kold, uint32_t V,
                              int32_t rows_old, int32_t rows_new, int32_t* rows, hipStream_t st);
// hist[m] = edges whose endpoints' smaller kmax is m, m = 0..kmax (hist has
// kmax + 3 words: the last holds the long-list queue's length); longq: room
// for E / 2048 + 64 slices of long lists (E = the graph's edges).
hipError_t launch_core_edges(const int32_t* kmaxes, Csr out, uint32_t V, int32_t kmax,
                             uint64_t* hist, uint64_t* longq, hipStream_t st);

// ---- h-index fixpoint (hindex_kernels.hip) -----------------------------
// One row in the core layout: initialises lv from prev (the previous row in
// rank space) or the member out-degree, refines in place, writes out_row.
int hindex_blocks_per_cu();
hipError_t launch_hindex_row(const PassSpec& p, const PassScratch& s, const int32_t* prev_lv,
                             uint32_t* flags, int32_t* lmax_out, const GridSync& gs,
                             hipStream_t st);

// ---- core-ordered graph (core_graph.hip, sort_kernels.hip) ------------
size_t scan_tmp_words(uint64_t n);
hipError_t launch_excl_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp,
                                hipStream_t st);
size_t radix_tmp_words(uint64_t n);
// Sorts by key bits [lo_bit, bits); being stable, it keeps the input order
// of keys equal in those bits.
hipError_t launch_radix_sort_u64(uint64_t* keys, uint64_t* alt, uint64_t n, int bits,
                                 uint32_t* table, uint64_t** result, hipStream_t st,
                                 int lo_bit = 0);
size_t core_graph_tmp_bytes(uint32_t V, uint64_t E, int32_t kmax);
// Adds sum(deg) to out[0] and the count of list entries in [0, deg) that
// are not vertex ids (< V) to out[1].
// nlen: entries allocated for side.nbr (lists reaching past it are counted
// bad, not read); E: the graph's edge count; tmp: audit_tmp_bytes(E) bytes.
size_t audit_tmp_bytes(uint64_t E);
hipError_t launch_audit(uint32_t V, Csr side, uint64_t nlen, uint64_t E, unsigned long long* out,
                        void* tmp, hipStream_t st);
// Core-ordered graph (core_graph.hip): rank[v] = new id of v (position in
// order); lists of new owner i = the lists of order[i] in new ids, ordered
// by the neighbours' kmax, descending.  With ekh (and vk, the V_k), the
// build also writes the E_k histogram: ekh[m] = out-edges whose endpoints'
// smaller kmax is m, m = 0..kmax (kmax < 4096; else ekh is left alone).
// With st2 and two events, the in side is built on st2 beside the out side
// (st waits for it before returning).
hipError_t build_core_graph(uint32_t V, uint64_t E, const uint32_t* order, const int32_t* kmaxes,
                            int32_t kmax, Csr out, Csr in, uint32_t* rank, uint32_t* cout_off,
                            int32_t* cout_deg, uint32_t* cout_nbr, uint32_t* cin_off,
                            int32_t* cin_deg, uint32_t* cin_nbr, void* tmp, const uint32_t* vk,
                            unsigned long long* ekh, hipStream_t st, hipStream_t st2 = nullptr,
                            hipEvent_t fork = nullptr, hipEvent_t join = nullptr);
// True when build_core_graph can write the E_k histogram for this kmax.
inline bool core_graph_writes_ek(int32_t kmax) { return kmax + 1 <= 4096; }

// ---- graph updates (maint_kernels.hip); pairs are (u, v) interleaved ---
// found[i] = 1 if pair i is an edge (looked up on the shorter side).
hipError_t launch_find_edges(Csr out, Csr in, const uint32_t* pairs, uint32_t n, uint32_t* found,
                             hipStream_t st);
hipError_t launch_check_room(const int32_t* deg, const uint32_t* capv, const uint32_t* gkey,
                             const uint32_t* gcnt, uint32_t ngroups, uint32_t* overflow,
                             hipStream_t st);
hipError_t launch_relocate(uint32_t* off, const int32_t* deg, uint32_t* capv, uint32_t* nbr,
                           const uint32_t* rkey, const uint32_t* new_off, const uint32_t* new_cap,
                           uint32_t nr, hipStream_t st);
hipError_t launch_apply_inserts(const uint32_t* off, int32_t* deg, uint32_t* nbr,
                                const uint32_t* gkey, const uint32_t* gstart, const uint32_t* gcnt,
                                const uint32_t* gval, uint32_t ngroups, hipStream_t st);
hipError_t launch_apply_deletes(const uint32_t* off, int32_t* deg, uint32_t* nbr,
                                const uint32_t* gkey, const uint32_t* gstart, const uint32_t* gcnt,
                                const uint32_t* gval, uint32_t ngroups, uint32_t* missing,
                                hipStream_t st);
hipError_t launch_mvalue(const int32_t* kmaxes, const uint32_t* pairs, uint32_t n, int32_t* m,
                         hipStream_t st);
hipError_t launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t st);
hipError_t launch_max_i32(const int32_t* p, size_t n, int32_t* out, hipStream_t st);

// ---- edge batches by root kmax (batch_form.hip) --------------------------
size_t edge_batch_tmp_bytes(uint32_t n, uint32_t nbuckets);
hipError_t launch_edge_batches(const int32_t* kmaxes, uint32_t V, const uint32_t* pairs, uint32_t n,
                               int mode, uint32_t nbuckets, uint32_t* batch, uint32_t* nbatch,
                               void* tmp, hipStream_t st);

// ---- allocation-free single row (noalloc.hip) ----------------------------
// Most vertices dcx_klist_noalloc accepts: its pass scratch, kScratchArrays
// words per vertex, is a static device array of that size (12 MiB).
constexpr uint32_t kNoAllocMaxVertices = 1u << 18;
// Address of the static scratch array on the current device.
hipError_t noalloc_scratch_address(void** ptr);


// ---- localized maintenance (local_kernels.hip) --------------------------
enum : int32_t { kLocalKmaxDel = 1, kLocalRowDel = 2, kLocalCoreIns = 3, kLocalRowIns = 4 };
struct LocalTaskDesc {
    int32_t type;  // kLocal*
    int32_t k;     // row (row tasks); 0 kmaxes / 1 row 0 (core insertion)
};
// Inputs of one launch of maintenance tasks (one workgroup per task at a
// time, tasks taken from *task_ctr).  Graph lists in the slack layout.
struct LocalArgs {
    uint32_t V;
    Csr out, in;
    int32_t* kmaxes;
    int32_t* rows;  // row k at rows + k * V
    const LocalTaskDesc* tasks;
    uint32_t ntasks;
    uint32_t* task_ctr;  // zero at launch
    uint32_t* scratch;   // gridDim.x slots of local_slot_words(qcap) words
    uint32_t qcap;       // worklist capacity per slot
    uint32_t* fail;      // bit (type) set when a task of that type overflowed
    // deletions (applied to the graph already)
    const uint32_t* del_pairs;
    uint32_t n_del;
    const int32_t* del_m;         // per deleted edge: min of its endpoints' old kmaxes
    const uint32_t* del_targets;  // distinct targets of deleted edges
    uint32_t n_del_targets;
    uint32_t* changed;     // (vertex, old kmax) pairs lowered by the deletions
    uint32_t* n_changed;   // their count (zero before the kmax deletion task)
    uint32_t changed_cap;
    // insertions (applied to the graph already)
    const uint32_t* ins_pairs;
    uint32_t n_ins;
    const uint32_t* gin_key;   // insertion groups by target: key, count
    const uint32_t* gin_cnt;
    uint32_t n_gin;
    const uint32_t* gout_key;  // by source
    const uint32_t* gout_cnt;
    uint32_t n_gout;
    const int32_t* ins_cnt_in;   // [V] the batch's entries at the end of each in-list
    const int32_t* ins_cnt_out;  // [V] same for out-lists
    // counters by task type t (1..4): [4 t] vertices evaluated, [4 t + 1]
    // candidates claimed, [4 t + 2] list entries read by the evaluations,
    // [4 t + 3] frontier expansions
    unsigned long long* counters;
    // work budgets: once the list entries read by evaluations and
    // expansions (all task types) exceed `budget`, or `ticks` of the
    // constant wall clock have passed since the first task of the batch
    // checked it (start kept in *t_start, zero before), every task stops
    // (failure bit 0x80) and the batch is recomputed instead; 0 = none
    unsigned long long budget;
    unsigned long long ticks;
    unsigned long long* t_start;
};
__host__ __device__ inline size_t local_slot_words(uint32_t qcap) { return 5ull * qcap; }
hipError_t launch_local_tasks(const LocalArgs& a, uint32_t blocks, hipStream_t st);
hipError_t launch_edge_min_kmax(const int32_t* kmaxes, const uint32_t* pairs, uint32_t n,
                                int32_t* out, hipStream_t st);
hipError_t launch_set_counts(int32_t* cnt, const uint32_t* key, const uint32_t* c, uint32_t n,
                             bool set, hipStream_t st);
hipError_t launch_max_masked(const int32_t* x, uint32_t n, int32_t* out, hipStream_t st);

}  // namespace dcx
