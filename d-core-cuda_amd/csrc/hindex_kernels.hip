// h-index fixpoint kernels: the second exact way to compute one l-row.
//
// Restates the l-value iteration designed in the reference's commented-out
// kList* family (kListRefineHIndex main.cu:1164-1242, hOutIndex/hInIndex
// src/cuda_utils.cuh:48-77): for v in the (k,0)-core C_k
//     l'(v) = max h <= l(v) such that
//             #{w in out(v) & C_k : l(w) >= h} >= h   and
//             #{w in in(v)  & C_k : l(w) >= h} >= k,
// iterated to a fixpoint from any upper bound of the true values.  Values
// only decrease and never drop below the true l (the operator is monotone
// and the true row is a fixpoint), so the updates run in place and in any
// order; a sweep that changes nothing ends the row.
//
// Mapping: one 64-lane wavefront per vertex.  The predicate above is
// monotone in h, so each wavefront binary-searches h in [0, l(v)]; every
// probe counts with ballot + popcount over the neighbour values, which
// stay in registers for lists up to 64*kCache entries.
#include <hip/hip_runtime.h>

#include "dcx_internal.hpp"
#include "grid_sync.hpp"

namespace dcx {
namespace {

constexpr int W = 64;
constexpr int kCache = 4;
constexpr int kBlk = 256;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (W - 1)); }

__device__ __forceinline__ int32_t ld_relaxed(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Count of list entries with value >= h; vals[] caches the first
// kCache*64 entries of the list (lane-strided).
__device__ __forceinline__ uint32_t count_ge(const int32_t (&vals)[kCache], const uint32_t* nbr,
                                             uint32_t n, const int32_t* l, int32_t h) {
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < kCache; ++r) c += __popcll(__ballot(vals[r] >= h));
    for (uint32_t base = kCache * W; base < n; base += W) {
        const uint32_t i = base + lane_id();
        const int32_t x = i < n ? ld_relaxed(l + nbr[i]) : -1;
        c += __popcll(__ballot(x >= h));
    }
    return c;
}

__device__ __forceinline__ void load_cache(int32_t (&vals)[kCache], const uint32_t* nbr,
                                           uint32_t n, const int32_t* l) {
#pragma unroll
    for (int r = 0; r < kCache; ++r) {
        const uint32_t i = r * W + lane_id();
        vals[r] = i < n ? ld_relaxed(l + nbr[i]) : -1;
    }
}

// One vertex: returns its new value (<= lv).  Uniform across the wavefront.
__device__ __forceinline__ int32_t refine(const PassSpec& P, uint32_t v, int32_t lv) {
    const uint32_t* on = P.K.nbr + P.K.off[v];
    const uint32_t* in = P.C.nbr + P.C.off[v];
    const uint32_t no = (uint32_t)P.K.deg[v], ni = (uint32_t)P.C.deg[v];
    int32_t ov[kCache], iv[kCache];
    load_cache(ov, on, no, P.lval);
    load_cache(iv, in, ni, P.lval);
    const int32_t k = P.k;
    // fast accept: the current value already satisfies the predicate
    auto pred = [&](int32_t h) -> bool {
        if ((int32_t)count_ge(ov, on, no, P.lval, h) < h) return false;
        return k == 0 || (int32_t)count_ge(iv, in, ni, P.lval, h) >= k;
    };
    if (lv == 0 || pred(lv)) return lv;
    int32_t lo = 0, hi = lv - 1;  // pred(0) holds for members of C_k
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (pred(mid))
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Rows are refined in place; flags[0] counts changes of the current sweep
// (double-buffered by sweep parity, reset one sweep ahead).
__global__ __launch_bounds__(kBlk) void k_hindex_row(PassSpec P, uint32_t* flags,
                                                     int32_t* sweeps_out, GridSync gs) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nwaves = gridDim.x * blockDim.x / W;
    const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
    if (leader) {
        flags[0] = 0;
        flags[1] = 0;
        flags[2] = 0;
    }
    if (!grid_barrier(gs)) return;
    int32_t sweep = 0;
    for (;;) {
        // values strictly decrease on every sweep that changes anything and
        // are bounded below, so the sweep count is bounded by the sum of the
        // initial values; a count past V * (max value) cannot happen
        if (sweep > 1 << 20) {
            if (threadIdx.x == 0) sync_fail(gs, kErrInconsistent);
            return;
        }
        uint32_t* cur = flags + (sweep % 3);
        if (leader) flags[(sweep + 1) % 3] = 0;
        bool any = false;
        for (uint32_t i = wave; i < P.nmembers; i += nwaves) {
            const uint32_t v = P.order ? P.order[i] : i;
            const int32_t lv = ld_relaxed(P.lval + v);
            const int32_t nv = refine(P, v, lv);
            if (nv < lv) {
                if (lane_id() == 0)
                    __hip_atomic_store(P.lval + v, nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                any = true;
            }
        }
        if (any && lane_id() == 0) atomicAdd(cur, 1u);
        if (!grid_barrier(gs)) return;
        ++sweep;
        if (__hip_atomic_load(cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
    }
    if (leader && sweeps_out) *sweeps_out = sweep;
}

// Upper bound for row k: row k-1 on the members of C_k (zero elsewhere is
// already in place), or the full out-degree when there is no previous row.
__global__ void k_hindex_init(PassSpec P, const int32_t* prev_row) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.nmembers;
         i += gridDim.x * blockDim.x) {
        const uint32_t v = P.order ? P.order[i] : i;
        P.lval[v] = prev_row ? prev_row[v] : P.K.deg[v];
    }
}

// lmax of a row: max over members.
__global__ void k_row_max(PassSpec P, int32_t* out) {
    int32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.nmembers;
         i += gridDim.x * blockDim.x) {
        const uint32_t v = P.order ? P.order[i] : i;
        m = max(m, P.lval[v]);
    }
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, W));
    if (lane_id() == 0) atomicMax(out, m);
}

}  // namespace

int hindex_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_hindex_row, kBlk, 0) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_hindex_row(const PassSpec& p, const int32_t* prev_row, uint32_t* flags,
                             int32_t* lmax_out, int32_t* sweeps_out, const GridSync& gs,
                             hipStream_t st) {
    if (p.nmembers == 0) return hipMemsetAsync(lmax_out, 0, sizeof(int32_t), st);
    const unsigned ib = (unsigned)min((p.nmembers + 255u) / 256u, 4096u);
    hipLaunchKernelGGL(k_hindex_init, dim3(ib), dim3(256), 0, st, p, prev_row);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(gs.w, 0, kSyncResetWords * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hindex_row, dim3(gs.nblocks), dim3(kBlk), 0, st, p, flags, sweeps_out,
                       gs);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(lmax_out, 0, sizeof(int32_t), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_row_max, dim3(ib), dim3(256), 0, st, p, lmax_out);
    return hipGetLastError();
}

}  // namespace dcx
