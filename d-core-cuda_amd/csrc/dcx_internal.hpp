// Internal declarations shared by the dcx HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "dcx.h"

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
hipError_t launch_pass_grid(const PassSpec& p, const PassScratch& s, uint32_t* ctrs,
                            int32_t* lmax_out, int grid, hipStream_t st);
int pass_grid_blocks_per_cu();
// Rows k = 0..kmax of one graph, one workgroup per pass pulled from a queue.
hipError_t launch_rows_batched(const PassSpec& base, void* scratch, int slots,
                               const uint32_t* vk, int32_t kmax, int32_t* rows,
                               int32_t* lmax_rows, uint32_t* queue_ctr, hipStream_t st);
hipError_t launch_order_by_kmax(const int32_t* kmaxes, uint32_t V, int32_t kmax,
                                uint32_t* hist /*[kmax+2]*/, uint32_t* vk /*[kmax+2]*/,
                                uint32_t* order, hipStream_t st);
hipError_t launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t st);
hipError_t launch_max_i32(const int32_t* p, size_t n, int32_t* out, hipStream_t st);

}  // namespace dcx
