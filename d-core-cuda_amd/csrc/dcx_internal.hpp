// Internal declarations shared by the dcx HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "dcx.h"
#include "grid_sync.hpp"

namespace dcx {

// One adjacency direction in device memory, reference layout
// (src/common.h:81-91): off[V+1] are capacity offsets (degree + slack),
// deg[V] the used prefix of each list.
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
struct PassSpec {
    uint32_t V;
    Csr K, C;
    const int32_t* kmaxes;  // membership filter kmaxes[v] >= k (nullptr: all)
    const uint32_t* order;  // members first (nullptr: identity order)
    uint32_t nmembers;      // |{v : kmaxes[v] >= k}| (V without filter)
    int32_t k;
    int32_t* lval;          // [V] output row; non-members must already hold 0
    // Levels above this are not peeled: vertices still alive afterwards keep
    // lval = -1 (bounded recompute after an update).
    int32_t level_cap = INT32_MAX;
};

// Per-pass device scratch (one slot per concurrently running pass).
struct PassScratch {
    int32_t* kd;   // [V] live key degree
    int32_t* cd;   // [V] live constraint degree
    uint32_t* f0;  // [V] frontier ping
    uint32_t* f1;  // [V] frontier pong
    uint32_t* a0;  // [V] alive list ping
    uint32_t* a1;  // [V] alive list pong
};
constexpr int kScratchArrays = 6;

__host__ __device__ inline PassScratch scratch_slot(void* base, uint32_t V, int slot) {
    uint32_t* p = reinterpret_cast<uint32_t*>(base) + (size_t)slot * kScratchArrays * V;
    PassScratch s;
    s.kd = reinterpret_cast<int32_t*>(p);
    s.cd = reinterpret_cast<int32_t*>(p + (size_t)V);
    s.f0 = p + 2ull * V;
    s.f1 = p + 3ull * V;
    s.a0 = p + 4ull * V;
    s.a1 = p + 5ull * V;
    return s;
}

// Kernel launchers (peel_kernels.hip).  All enqueue on `st` and return the
// HIP status of the launch.
// One pass on the whole grid: cooperative launch (co-residency) with the
// bounded barrier of grid_sync.hpp; gs.w is zeroed by the launcher.
hipError_t launch_pass_grid(const PassSpec& p, const PassScratch& s, uint32_t* ctrs,
                            int32_t* lmax_out, const GridSync& gs, hipStream_t st);
int pass_grid_blocks_per_cu();
// Rows k = 0..kmax of one graph, one workgroup per pass pulled from a queue;
// a pass whose counters do not add up raises *err (kErrInconsistent).
hipError_t launch_rows_batched(const PassSpec& base, void* scratch, int slots,
                               const uint32_t* vk, int32_t kmax, int32_t* rows,
                               int32_t* lmax_rows, uint32_t* queue_ctr, uint32_t* err,
                               hipStream_t st);
hipError_t launch_order_by_kmax(const int32_t* kmaxes, uint32_t V, int32_t kmax,
                                uint32_t* hist /*[kmax+2]*/, uint32_t* vk /*[kmax+2]*/,
                                uint32_t* order, hipStream_t st);
// h-index fixpoint for one row (hindex_kernels.hip): initialises the row
// from prev_row (or the out-degree) on the members, then refines in place.
int hindex_blocks_per_cu();
hipError_t launch_hindex_row(const PassSpec& p, const int32_t* prev_row, uint32_t* flags,
                             int32_t* lmax_out, int32_t* sweeps_out, const GridSync& gs,
                             hipStream_t st);
// Graph updates (maint_kernels.hip).  pairs are (u, v) interleaved.
hipError_t launch_find_edges(const uint32_t* off, const int32_t* deg, const uint32_t* nbr,
                             const uint32_t* pairs, uint32_t n, uint32_t* found, hipStream_t st);
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
hipError_t launch_capped_fixup(int32_t* kmaxes, const int32_t* old, uint32_t V, int32_t cap,
                               hipStream_t st);
hipError_t launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t st);
hipError_t launch_max_i32(const int32_t* p, size_t n, int32_t* out, hipStream_t st);

}  // namespace dcx
