// Core-ordered graph: vertices relabelled by kmax (descending), every list
// sorted by the new ids.  Then for every k the (k,0)-core is the id prefix
// [0, V_k), and within a member's list the neighbours inside the core form
// a prefix too -- a k-pass touches exactly the edges of the core (E_k per
// direction) and needs no membership test.  Built once per decomposition,
// after the kmax pass, with one radix sort per direction.
#include <hip/hip_runtime.h>

#include "dcx_internal.hpp"

namespace dcx {
namespace {

constexpr int W = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (W - 1)); }

__global__ void k_rank(const uint32_t* order, uint32_t V, uint32_t* rank) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x)
        rank[order[i]] = i;
}

__global__ void k_new_degrees(const uint32_t* order, uint32_t V, const int32_t* deg,
                              uint32_t* dnew) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x)
        dnew[i] = (uint32_t)deg[order[i]];
}

// keys[off_new[i] + j] = (i << B) | rank[nbr_j] for the list of old vertex
// order[i] (one wavefront per vertex).
__global__ void k_make_keys(const uint32_t* order, const uint32_t* rank, uint32_t V,
                            const uint32_t* off, const int32_t* deg, const uint32_t* nbr,
                            const uint32_t* off_new, int B, uint64_t* keys) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nw = gridDim.x * blockDim.x / W;
    for (uint32_t i = wave; i < V; i += nw) {
        const uint32_t v = order[i];
        const uint32_t o = off[v], n = (uint32_t)deg[v], dst = off_new[i];
        for (uint32_t j = lane_id(); j < n; j += W)
            keys[dst + j] = ((uint64_t)i << B) | rank[nbr[o + j]];
    }
}

__global__ void k_keys_to_nbr(const uint64_t* keys, uint64_t n, uint64_t mask, uint32_t* nbr) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        nbr[i] = (uint32_t)(keys[i] & mask);
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

// out[0] += sum of deg, out[1] += entries of [0, deg) that are not vertex
// ids (one wavefront per vertex).
__global__ void k_audit(uint32_t V, const uint32_t* off, const int32_t* deg, const uint32_t* nbr,
                        unsigned long long* out) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nw = gridDim.x * blockDim.x / W;
    uint64_t sum = 0, bad = 0;
    for (uint32_t v = wave; v < V; v += nw) {
        const int32_t d = deg[v];
        if (lane_id() == 0) sum += d < 0 ? 0 : (uint64_t)d;
        if (d < 0) {
            if (lane_id() == 0) ++bad;
            continue;
        }
        const uint32_t* l = nbr + off[v];
        for (uint32_t j = lane_id(); j < (uint32_t)d; j += W) bad += l[j] >= V;
    }
    for (int dd = W / 2; dd > 0; dd >>= 1) {
        sum += __shfl_xor(sum, dd, W);
        bad += __shfl_xor(bad, dd, W);
    }
    // one pair of atomics per block (same-address atomics serialise)
    __shared__ unsigned long long s_sum, s_bad;
    if (threadIdx.x == 0) s_sum = s_bad = 0;
    __syncthreads();
    if (lane_id() == 0) {
        atomicAdd(&s_sum, (unsigned long long)sum);
        atomicAdd(&s_bad, (unsigned long long)bad);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(out, s_sum);
        atomicAdd(out + 1, s_bad);
    }
}

unsigned grid_for(uint64_t n, unsigned per, unsigned cap = 8192) {
    uint64_t b = (n + per - 1) / per;
    if (!b) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

int bits_for(uint32_t v) {
    int b = 0;
    while (b < 32 && (v >> b)) ++b;
    return b ? b : 1;
}

}  // namespace

size_t core_graph_tmp_bytes(uint32_t V, uint64_t E) {
    // keys + alternate keys + scan words + radix table
    return 16ull * (E ? E : 1) + 4ull * (scan_tmp_words(V) + 1) +
           4ull * (radix_tmp_words(E) + 1) + 256;
}

hipError_t launch_audit(uint32_t V, Csr side, unsigned long long* out, hipStream_t st) {
    if (V == 0) return hipSuccess;
    hipLaunchKernelGGL(k_audit, dim3(grid_for(V, 16, 2048)), dim3(1024), 0, st, V, side.off,
                       side.deg, side.nbr, out);
    return hipGetLastError();
}

hipError_t build_core_graph(uint32_t V, uint64_t E, const uint32_t* order, Csr out, Csr in,
                            uint32_t* rank, uint32_t* cout_off, int32_t* cout_deg,
                            uint32_t* cout_nbr, uint32_t* cin_off, int32_t* cin_deg,
                            uint32_t* cin_nbr, void* tmp, hipStream_t st) {
    if (V == 0) return hipSuccess;
    char* p = static_cast<char*>(tmp);
    uint64_t* keys = reinterpret_cast<uint64_t*>(p);
    uint64_t* alt = keys + (E ? E : 1);
    uint32_t* scan_tmp = reinterpret_cast<uint32_t*>(alt + (E ? E : 1));
    uint32_t* table = scan_tmp + scan_tmp_words(V) + 1;
    const int B = bits_for(V - 1);
    const uint64_t mask = (1ull << B) - 1;
    hipLaunchKernelGGL(k_rank, dim3(grid_for(V, 256)), dim3(256), 0, st, order, V, rank);
    const Csr sides[2] = {out, in};
    uint32_t* offs[2] = {cout_off, cin_off};
    int32_t* degs[2] = {cout_deg, cin_deg};
    uint32_t* nbrs[2] = {cout_nbr, cin_nbr};
    for (int s = 0; s < 2; ++s) {
        uint32_t* dnew = reinterpret_cast<uint32_t*>(degs[s]);
        hipLaunchKernelGGL(k_new_degrees, dim3(grid_for(V, 256)), dim3(256), 0, st, order, V,
                           sides[s].deg, dnew);
        hipError_t e = launch_excl_scan_u32(dnew, offs[s], V, scan_tmp, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, offs[s] + V, (uint32_t)E);
        if (E == 0) continue;
        hipLaunchKernelGGL(k_make_keys, dim3(grid_for(V, 4)), dim3(256), 0, st, order, rank, V,
                           sides[s].off, sides[s].deg, sides[s].nbr, offs[s], B, keys);
        uint64_t* sorted = nullptr;
        e = launch_radix_sort_u64(keys, alt, E, 2 * B, table, &sorted, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_keys_to_nbr, dim3(grid_for(E, 256)), dim3(256), 0, st, sorted, E,
                           mask, nbrs[s]);
    }
    return hipGetLastError();
}

}  // namespace dcx
