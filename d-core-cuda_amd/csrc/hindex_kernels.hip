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
// Runs on the core-ordered graph (members are ranks [0, V_k), member
// neighbours are list prefixes).  Mapping: one 64-lane wavefront per
// vertex.  The predicate above is monotone in h, so each wavefront
// binary-searches h in [0, l(v)]; every probe counts with ballot + popcount
// over the neighbour values, which stay in registers for lists up to
// 64*kCache entries.
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

// Values of the row being refined: rank space (core layout) or the output
// array itself (full layout, every vertex a member).
__device__ __forceinline__ int32_t* row_values(const PassSpec& P, const PassScratch& S) {
    return P.core ? S.lv : P.lval;
}

// One vertex: returns its new value (<= lv).  Uniform across the wavefront.
__device__ __forceinline__ int32_t refine(const PassSpec& P, const PassScratch& S, uint32_t v,
                                          int32_t lv) {
    const int32_t* l = row_values(P, S);
    const uint32_t* on = P.K.nbr + P.K.off[v];
    const uint32_t* in = P.C.nbr + P.C.off[v];
    const uint32_t no = P.core ? S.rk[v] : (uint32_t)P.K.deg[v];
    const int32_t k = P.k;
    // the C side only matters for k > 0
    const uint32_t ni = k == 0 ? 0u : (P.core ? S.rc[v] : (uint32_t)P.C.deg[v]);
    int32_t ov[kCache], iv[kCache];
    load_cache(ov, on, no, l);
    load_cache(iv, in, ni, l);
    auto pred = [&](int32_t h) -> bool {
        if ((int32_t)count_ge(ov, on, no, l, h) < h) return false;
        return k == 0 || (int32_t)count_ge(iv, in, ni, l, h) >= k;
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

// Upper bound for the row: the previous row (l_k <= l_{k-1}) or an older
// value of the same row that can only have dropped, capped by the member
// K-degree; in the core layout also the member prefix lengths.
__global__ void k_hindex_init(PassSpec P, PassScratch S, const int32_t* prev_lv) {
    const uint32_t nm = P.nmembers;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += gridDim.x * blockDim.x) {
        uint32_t rk;
        if (P.core) {
            rk = lower_bound(P.K.nbr + P.K.off[i], (uint32_t)P.K.deg[i], nm);
            S.rk[i] = rk;
            S.rc[i] = lower_bound(P.C.nbr + P.C.off[i], (uint32_t)P.C.deg[i], nm);
        } else {
            rk = (uint32_t)P.K.deg[i];
        }
        const int32_t u = prev_lv ? min(prev_lv[i], (int32_t)rk) : (int32_t)rk;
        row_values(P, S)[i] = u;
    }
}

// Rows are refined in place; flags count changes per sweep (triple
// buffered, reset one sweep ahead).
__global__ __launch_bounds__(kBlk) void k_hindex_row(PassSpec P, PassScratch S, uint32_t* flags,
                                                     GridSync gs) {
    gs.rank = blockIdx.x;
    gs.epoch = 0;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nwaves = gridDim.x * blockDim.x / W;
    const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
    const uint64_t t_start = wall_clock64();
    __shared__ uint32_t s_changed;
    // change flags kSlotStride words apart (own lines), RMW access only
    if (leader) {
        coherent_write(flags, 0u);
        coherent_write(flags + kSlotStride, 0u);
        coherent_write(flags + 2 * kSlotStride, 0u);
    }
    if (!grid_barrier(gs)) return;
    int32_t sweep = 0;
    for (;;) {
        // every sweep that changes something lowers a value, so the count
        // is finite; this bound only guards against a broken counter
        if (sweep > 1 << 20) {
            if (threadIdx.x == 0) sync_fail(gs, kErrInconsistent);
            return;
        }
        uint32_t* cur = flags + kSlotStride * (sweep % 3);
        if (leader) coherent_write(flags + kSlotStride * ((sweep + 1) % 3), 0u);
        bool any = false;
        int32_t* const l = row_values(P, S);
        for (uint32_t v = wave; v < P.nmembers; v += nwaves) {
            const int32_t lv = ld_relaxed(l + v);
            const int32_t nv = refine(P, S, v, lv);
            if (nv < lv) {
                if (lane_id() == 0)
                    __hip_atomic_store(l + v, nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                any = true;
            }
        }
        if (any && lane_id() == 0) consume(atomicAdd(cur, 1u));
        if (!grid_barrier(gs)) return;
        ++sweep;
        // one coherent read per workgroup: every workgroup must take the
        // same decision
        if (threadIdx.x == 0) s_changed = coherent_read(cur);
        __syncthreads();
        const uint32_t changed = s_changed;
        __syncthreads();
        if (changed == 0) break;
    }
    if (leader && P.stats) {
        const uint64_t dt = wall_clock64() - t_start;
        P.stats[kStatPhases] = (uint32_t)sweep;
        P.stats[kStatLevels] = 0;
        P.stats[kStatTicksLo] = (uint32_t)dt;
        P.stats[kStatTicksHi] = (uint32_t)(dt >> 32);
        P.stats[kStatFailed] = 0;
        P.stats[kStatRemoved] = P.nmembers;
        P.stats[kStatAlive] = 0;
        P.stats[kStatLevel] = 0;
    }
}

// Writes the row back to original ids (core layout) and its maximum.
__global__ void k_hindex_out(PassSpec P, PassScratch S, int32_t* lmax_out) {
    int32_t m = 0;
    const int32_t* l = row_values(P, S);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P.nmembers;
         i += gridDim.x * blockDim.x) {
        const int32_t x = l[i];
        if (P.core) P.out_row[P.order[i]] = x;
        m = max(m, x);
    }
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, W));
    if (lane_id() == 0) atomicMax(lmax_out, m);
}

}  // namespace

int hindex_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_hindex_row, kBlk, 0) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_hindex_row(const PassSpec& p, const PassScratch& s, const int32_t* prev_lv,
                             uint32_t* flags, int32_t* lmax_out, const GridSync& gs,
                             hipStream_t st) {
    hipError_t e = hipMemsetAsync(lmax_out, 0, sizeof(int32_t), st);
    if (e != hipSuccess || p.nmembers == 0) return e;
    const unsigned ib = (unsigned)min((p.nmembers + 255u) / 256u, 4096u);
    hipLaunchKernelGGL(k_hindex_init, dim3(ib), dim3(256), 0, st, p, s, prev_lv);
    e = hipMemsetAsync(gs.w, 0, kSyncResetWords * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(gs.w + kSyncFlags, 0, gs.nblocks * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hindex_row, dim3(gs.nblocks), dim3(kBlk), 0, st, p, s, flags, gs);
    hipLaunchKernelGGL(k_hindex_out, dim3(ib), dim3(256), 0, st, p, s, lmax_out);
    return hipGetLastError();
}

}  // namespace dcx
