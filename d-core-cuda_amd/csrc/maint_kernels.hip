// Device-side graph updates for D-core maintenance.
//
// The graph lives in HBM as in/out adjacency with per-vertex capacity
// (capv) and a bump-allocated tail region for relocated lists.  Updates are
// applied in place, one wavefront per vertex group, so two edits of the same
// list never race:
//   * insert: append after the used prefix (GraphGPU::insertEdges /
//     graphInsertEdges, src/graph_gpu.cu:49-74, but with a capacity check
//     done beforehand and relocation of lists that would overflow);
//   * delete: swap the edge with the last used slot (the reference has no
//     working delete: src/graph_gpu.h:18-19, src/graph.cpp:73-108).
// Plus the small reductions the maintenance driver needs: edge lookup and
// the M value of a batch (findMValue, main.cu:344-364).
#include <hip/hip_runtime.h>

#include <climits>

#include "dcx_internal.hpp"

namespace dcx {
namespace {

constexpr int W = 64;
constexpr uint32_t kFree = 0xffffffffu;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (W - 1)); }

// Index of the first occurrence of x in list[0..n), or -1.  Wave-uniform.
// FRESH: read past the CU's vector cache (the list was just edited).
template <bool FRESH = false>
__device__ __forceinline__ int64_t wave_find(const uint32_t* list, uint32_t n, uint32_t x) {
    for (uint32_t base = 0; base < n; base += W) {
        const uint32_t i = base + lane_id();
        uint32_t y = kFree;
        if (i < n)
            y = FRESH ? __hip_atomic_load(list + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                      : list[i];
        const bool hit = i < n && y == x;
        const unsigned long long m = __ballot(hit);
        if (m) return (int64_t)base + (__ffsll(m) - 1);
    }
    return -1;
}

// found[i] = 1 if pair i (u = pairs[2i], v = pairs[2i+1]) is an edge,
// looked up in the shorter of u's out-list and v's in-list (both hold it;
// an R-MAT edge usually has a hub at one end).
__global__ void k_find_edges(Csr out, Csr in, const uint32_t* pairs, uint32_t n,
                             uint32_t* found) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nw = gridDim.x * blockDim.x / W;
    for (uint32_t i = wave; i < n; i += nw) {
        const uint32_t u = pairs[2 * i], v = pairs[2 * i + 1];
        const uint32_t du = (uint32_t)out.deg[u], dv = (uint32_t)in.deg[v];
        const int64_t p = du <= dv ? wave_find(out.nbr + out.off[u], du, v)
                                   : wave_find(in.nbr + in.off[v], dv, u);
        if (lane_id() == 0) found[i] = p >= 0;
    }
}

// overflow[g] = deg + 1 if key gkey[g] has fewer free slots than gcnt[g],
// else 0 (the host sizes the relocated list from the degree).
__global__ void k_check_room(const int32_t* deg, const uint32_t* capv, const uint32_t* gkey,
                             const uint32_t* gcnt, uint32_t ngroups, uint32_t* overflow) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ngroups;
         g += gridDim.x * blockDim.x) {
        const uint32_t key = gkey[g];
        const uint32_t d = (uint32_t)deg[key];
        overflow[g] = capv[key] - d < gcnt[g] ? d + 1 : 0u;
    }
}

// Moves the lists of rkey[r] to new_off[r] with capacity new_cap[r].
__global__ void k_relocate(uint32_t* off, const int32_t* deg, uint32_t* capv, uint32_t* nbr,
                           const uint32_t* rkey, const uint32_t* new_off,
                           const uint32_t* new_cap, uint32_t nr) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nw = gridDim.x * blockDim.x / W;
    for (uint32_t r = wave; r < nr; r += nw) {
        const uint32_t key = rkey[r];
        const uint32_t from = off[key], to = new_off[r], n = (uint32_t)deg[key];
        const uint32_t cap = new_cap[r];
        for (uint32_t i = lane_id(); i < cap; i += W) nbr[to + i] = i < n ? nbr[from + i] : kFree;
        for (uint32_t i = lane_id(); i < n; i += W) nbr[from + i] = kFree;
        if (lane_id() == 0) {
            off[key] = to;
            capv[key] = cap;
        }
    }
}

// Group g: key gkey[g], values gval[gstart[g] .. +gcnt[g]).
__global__ void k_apply_inserts(const uint32_t* off, int32_t* deg, uint32_t* nbr,
                                const uint32_t* gkey, const uint32_t* gstart,
                                const uint32_t* gcnt, const uint32_t* gval, uint32_t ngroups) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nw = gridDim.x * blockDim.x / W;
    for (uint32_t g = wave; g < ngroups; g += nw) {
        const uint32_t key = gkey[g], s = gstart[g], c = gcnt[g];
        const uint32_t d0 = (uint32_t)deg[key];
        uint32_t* list = nbr + off[key];
        for (uint32_t j = lane_id(); j < c; j += W) list[d0 + j] = gval[s + j];
        if (lane_id() == 0) deg[key] = (int32_t)(d0 + c);
    }
}

// Deletes group values one after another (swap with the last used slot).
// Every value is known to be present (checked before the launch).
__global__ void k_apply_deletes(const uint32_t* off, int32_t* deg, uint32_t* nbr,
                                const uint32_t* gkey, const uint32_t* gstart,
                                const uint32_t* gcnt, const uint32_t* gval, uint32_t ngroups,
                                uint32_t* missing) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nw = gridDim.x * blockDim.x / W;
    for (uint32_t g = wave; g < ngroups; g += nw) {
        const uint32_t key = gkey[g], s = gstart[g], c = gcnt[g];
        uint32_t* list = nbr + off[key];
        uint32_t d = (uint32_t)deg[key];
        for (uint32_t j = 0; j < c; ++j) {
            const int64_t p = wave_find<true>(list, d, gval[s + j]);
            if (p < 0) {
                if (lane_id() == 0) atomicAdd(missing, 1u);
                continue;
            }
            if (lane_id() == 0) {
                const uint32_t last =
                    __hip_atomic_load(list + d - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(list + p, last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(list + d - 1, kFree, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            --d;
            // the next search by the other lanes must see lane 0's writes
            __threadfence();
            __builtin_amdgcn_wave_barrier();
        }
        if (lane_id() == 0) deg[key] = (int32_t)d;
    }
}

// M = max over pairs of min(kmaxes[u], kmaxes[v]) (findMValue main.cu:344-357).
__global__ void k_mvalue(const int32_t* kmaxes, const uint32_t* pairs, uint32_t n, int32_t* m) {
    int32_t best = INT_MIN;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        best = max(best, min(kmaxes[pairs[2 * i]], kmaxes[pairs[2 * i + 1]]));
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) best = max(best, __shfl_xor(best, d, W));
    if (lane_id() == 0 && best != INT_MIN) atomicMax(m, best);
}

unsigned blocks_for(uint64_t n, unsigned per_block, unsigned cap = 4096) {
    uint64_t b = (n + per_block - 1) / per_block;
    if (b == 0) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

}  // namespace

hipError_t launch_find_edges(Csr out, Csr in, const uint32_t* pairs, uint32_t n, uint32_t* found,
                             hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_find_edges, dim3(blocks_for(n, 4)), dim3(256), 0, st, out, in, pairs, n,
                       found);
    return hipGetLastError();
}

hipError_t launch_check_room(const int32_t* deg, const uint32_t* capv, const uint32_t* gkey,
                             const uint32_t* gcnt, uint32_t ngroups, uint32_t* overflow,
                             hipStream_t st) {
    if (!ngroups) return hipSuccess;
    hipLaunchKernelGGL(k_check_room, dim3(blocks_for(ngroups, 256)), dim3(256), 0, st, deg, capv,
                       gkey, gcnt, ngroups, overflow);
    return hipGetLastError();
}

hipError_t launch_relocate(uint32_t* off, const int32_t* deg, uint32_t* capv, uint32_t* nbr,
                           const uint32_t* rkey, const uint32_t* new_off, const uint32_t* new_cap,
                           uint32_t nr, hipStream_t st) {
    if (!nr) return hipSuccess;
    hipLaunchKernelGGL(k_relocate, dim3(blocks_for(nr, 4)), dim3(256), 0, st, off, deg, capv, nbr,
                       rkey, new_off, new_cap, nr);
    return hipGetLastError();
}

hipError_t launch_apply_inserts(const uint32_t* off, int32_t* deg, uint32_t* nbr,
                                const uint32_t* gkey, const uint32_t* gstart, const uint32_t* gcnt,
                                const uint32_t* gval, uint32_t ngroups, hipStream_t st) {
    if (!ngroups) return hipSuccess;
    hipLaunchKernelGGL(k_apply_inserts, dim3(blocks_for(ngroups, 4)), dim3(256), 0, st, off, deg,
                       nbr, gkey, gstart, gcnt, gval, ngroups);
    return hipGetLastError();
}

hipError_t launch_apply_deletes(const uint32_t* off, int32_t* deg, uint32_t* nbr,
                                const uint32_t* gkey, const uint32_t* gstart, const uint32_t* gcnt,
                                const uint32_t* gval, uint32_t ngroups, uint32_t* missing,
                                hipStream_t st) {
    if (!ngroups) return hipSuccess;
    hipLaunchKernelGGL(k_apply_deletes, dim3(blocks_for(ngroups, 4)), dim3(256), 0, st, off, deg,
                       nbr, gkey, gstart, gcnt, gval, ngroups, missing);
    return hipGetLastError();
}

hipError_t launch_mvalue(const int32_t* kmaxes, const uint32_t* pairs, uint32_t n, int32_t* m,
                         hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mvalue, dim3(blocks_for(n, 256, 256)), dim3(256), 0, st, kmaxes, pairs, n,
                       m);
    return hipGetLastError();
}

}  // namespace dcx
