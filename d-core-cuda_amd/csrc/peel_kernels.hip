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
//   * membership and live degrees are set up once per pass, restricted to
//     the (k,0)-core (members are a prefix of a kmax-descending order, so a
//     pass over a small core never touches the rest of the graph);
//   * each level seeds from a compacted alive list instead of a scan over
//     all V, and the next level number jumps straight to the smallest live
//     key degree (levels that would remove nothing are skipped);
//   * frontier vertices are expanded edge-parallel: a tile of up to BLK
//     vertices is prefix-summed over its list lengths in LDS and every
//     thread takes one edge (load-balanced search), so hubs do not stall a
//     wavefront;
//   * frontier and alive-list appends are compacted with a 64-lane ballot
//     and popcount, one atomic per wavefront;
//   * phase counters are triple-buffered so each phase needs exactly one
//     barrier;
//   * a pass runs either on one workgroup (batched mode: many passes in one
//     persistent launch, workgroup barriers only) or on the whole grid
//     (cooperative launch, grid barriers).
#include <hip/hip_runtime.h>

#include <climits>

#include "dcx_internal.hpp"
#include "grid_sync.hpp"

namespace dcx {
namespace {

constexpr int W = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (W - 1)); }

// Append `val` to arr[] for every lane with `pred`, one atomic per wavefront.
// Every lane of the wavefront must reach this call.
__device__ __forceinline__ void wave_append(bool pred, uint32_t val, uint32_t* arr,
                                            uint32_t* ctr) {
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return;
    const int leader = __ffsll(m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    base = __shfl(base, leader, W);
    if (pred) arr[base + __popcll(m & ((1ull << lane_id()) - 1ull))] = val;
}

// Wavefront minimum, then one atomicMin per wavefront.
__device__ __forceinline__ void wave_min_to(uint32_t x, uint32_t* dst) {
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) x = min(x, (uint32_t)__shfl_xor(x, d, W));
    if (lane_id() == 0 && x != UINT_MAX) atomicMin(dst, x);
}

template <int BLK>
struct TileLds {
    uint32_t item[BLK];
    uint32_t koff[BLK];
    uint32_t coff[BLK];
    uint32_t klen[BLK];
    uint32_t scan[BLK];
    int32_t cntk[BLK];
    int32_t cntc[BLK];
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

// Loads a tile of frontier/member vertices and prefix-sums their list
// lengths (K list then C list per vertex).
template <int BLK>
__device__ __forceinline__ void tile_prepare(TileLds<BLK>& T, const PassSpec& P, bool has,
                                             uint32_t v, uint32_t n) {
    const uint32_t tid = threadIdx.x;
    uint32_t kl = 0, cl = 0;
    if (has) {
        T.item[tid] = v;
        T.koff[tid] = P.K.off[v];
        T.coff[tid] = P.C.off[v];
        kl = (uint32_t)P.K.deg[v];
        cl = (uint32_t)P.C.deg[v];
    }
    T.klen[tid] = kl;
    if (tid == 0) T.nitems = n;
    T.scan[tid] = block_excl_scan<BLK>(kl + cl, T.wsum, &T.total);
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

__device__ __forceinline__ void slot_reset(uint32_t* s) {
    s[0] = 0;         // frontier appends
    s[1] = 0;         // alive-list appends
    s[2] = UINT_MAX;  // smallest key degree seen above the level
    s[3] = 0;
}

template <bool GRID>
__device__ __forceinline__ uint32_t slot_read(const uint32_t* p) {
    if constexpr (GRID)
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *(volatile const uint32_t*)p;
}

// One full peel pass.  GRID: all workgroups of a cooperative launch work on
// the pass, `slots` is in global memory and `gs` is the bounded grid
// barrier; otherwise the calling workgroup runs the whole pass and `slots`
// is in LDS.  Logic errors (counters that do not add up) are reported into
// *err (and abort a grid launch) instead of looping.
template <int BLK, bool GRID>
__device__ void run_pass(const PassSpec& P, const PassScratch& S, uint32_t* slots,
                         TileLds<BLK>& T, int32_t* lmax_out, const GridSync* gs,
                         uint32_t* err) {
    const uint32_t tid = threadIdx.x;
    const uint32_t wg = GRID ? blockIdx.x : 0u;
    const uint32_t nwg = GRID ? gridDim.x : 1u;
    auto barrier = [&]() -> bool {
        if constexpr (GRID) {
            return grid_barrier(*gs);
        } else {
            __syncthreads();
            return true;
        }
    };
    auto fail = [&](uint32_t code) {
        if (tid == 0) {
            if constexpr (GRID)
                sync_fail(*gs, code);
            else
                atomicMax(err, code);
        }
    };

    if (wg == 0 && tid == 0) {
        slot_reset(slots);
        slot_reset(slots + 4);
        slot_reset(slots + 8);
    }

    // ---- set-up: members, live degrees restricted to members, alive list
    const uint32_t nm = P.nmembers;
    for (uint32_t base = wg * BLK; base < nm; base += nwg * BLK) {
        const uint32_t i = base + tid;
        const bool has = i < nm;
        const uint32_t v = has ? (P.order ? P.order[i] : i) : 0u;
        if (P.kmaxes) {
            __syncthreads();
            T.cntk[tid] = 0;
            T.cntc[tid] = 0;
            tile_prepare<BLK>(T, P, has, v, min(nm - base, (uint32_t)BLK));
            const uint32_t total = T.total;
            for (uint32_t eb = 0; eb < total; eb += BLK) {
                const uint32_t e = eb + tid;
                if (e < total) {
                    const uint32_t it = tile_find<BLK>(T, e);
                    const uint32_t t = e - T.scan[it];
                    const bool kside = t < T.klen[it];
                    const uint32_t x = kside ? P.K.nbr[T.koff[it] + t]
                                             : P.C.nbr[T.coff[it] + (t - T.klen[it])];
                    if (P.kmaxes[x] >= P.k) atomicAdd(kside ? &T.cntk[it] : &T.cntc[it], 1);
                }
            }
            __syncthreads();
            if (has) {
                S.kd[v] = T.cntk[tid];
                S.cd[v] = T.cntc[tid];
            }
        } else if (has) {
            S.kd[v] = P.K.deg[v];
            S.cd[v] = P.C.deg[v];
        }
        if (has) {
            P.lval[v] = -1;
            S.a0[i] = v;
        }
    }
    if (!barrier()) return;

    uint32_t *A = S.a0, *B = S.a1, *F = S.f0, *G = S.f1;
    uint32_t nA = nm, removed = 0, s = 0, levels = 0;
    int32_t L = 0;
    const int32_t k = P.k;
    // Without the membership filter the constraint can fail before any
    // removal (KList semantics on the whole graph, main.cu:144-169).
    const bool check_c = P.kmaxes == nullptr && k > 0;
    while (removed < nm) {
        // Each removal needs at most two level iterations (a level number
        // taken from a stale minimum seeds nothing, the next one seeds the
        // true minimum), so more than 2*nm+2 means the counters are wrong.
        if (++levels > 2 * nm + 2) {
            fail(kErrInconsistent);
            break;
        }
        // ---- seed level L from the alive list; survivors go to B
        uint32_t* cur = slots + 4 * (s % 3);
        if (wg == 0 && tid == 0) slot_reset(slots + 4 * ((s + 1) % 3));
        uint32_t mn = UINT_MAX;
        for (uint32_t base = wg * BLK; base < nA; base += nwg * BLK) {
            const uint32_t i = base + tid;
            const bool has = i < nA;
            const uint32_t v = has ? A[i] : 0u;
            const bool alive = has && P.lval[v] == -1;
            const int32_t d = alive ? S.kd[v] : 0;
            const bool seed = alive && (d <= L || (check_c && S.cd[v] < k));
            const bool keep = alive && !seed;
            if (seed) P.lval[v] = L;
            if (keep) mn = min(mn, (uint32_t)d);
            wave_append(seed, v, F, &cur[0]);
            wave_append(keep, v, B, &cur[1]);
        }
        wave_min_to(mn, &cur[2]);
        if (!barrier()) return;
        uint32_t nF = slot_read<GRID>(&cur[0]);
        nA = slot_read<GRID>(&cur[1]);
        uint32_t mnall = slot_read<GRID>(&cur[2]);
        {
            uint32_t* t = A;
            A = B;
            B = t;
        }
        ++s;
        removed += nF;

        // ---- cascade inside level L until no vertex crosses a threshold
        while (nF > 0) {
            uint32_t* nxt = slots + 4 * (s % 3);
            if (wg == 0 && tid == 0) slot_reset(slots + 4 * ((s + 1) % 3));
            uint32_t mn2 = UINT_MAX;
            for (uint32_t base = wg * BLK; base < nF; base += nwg * BLK) {
                const uint32_t i = base + tid;
                const bool has = i < nF;
                const uint32_t v = has ? F[i] : 0u;
                __syncthreads();
                tile_prepare<BLK>(T, P, has, v, min(nF - base, (uint32_t)BLK));
                const uint32_t total = T.total;
                for (uint32_t eb = 0; eb < total; eb += BLK) {
                    const uint32_t e = eb + tid;
                    bool claim = false;
                    uint32_t x = 0;
                    if (e < total) {
                        const uint32_t it = tile_find<BLK>(T, e);
                        const uint32_t t = e - T.scan[it];
                        const bool kside = t < T.klen[it];
                        x = kside ? P.K.nbr[T.koff[it] + t]
                                  : P.C.nbr[T.coff[it] + (t - T.klen[it])];
                        if (P.lval[x] == -1) {
                            if (kside) {
                                // x loses a C-neighbour: constraint degree
                                const int32_t old = atomicSub(&S.cd[x], 1);
                                claim = k > 0 && old == k;
                            } else {
                                // x loses a K-neighbour: key degree
                                const int32_t old = atomicSub(&S.kd[x], 1);
                                if (old == L + 1)
                                    claim = true;
                                else if (old - 1 > L)
                                    mn2 = min(mn2, (uint32_t)(old - 1));
                            }
                            if (claim) claim = atomicCAS(&P.lval[x], -1, L) == -1;
                        }
                    }
                    wave_append(claim, x, G, &nxt[0]);
                }
            }
            wave_min_to(mn2, &nxt[2]);
            if (!barrier()) return;
            nF = slot_read<GRID>(&nxt[0]);
            mnall = min(mnall, slot_read<GRID>(&nxt[2]));
            {
                uint32_t* t = F;
                F = G;
                G = t;
            }
            ++s;
            removed += nF;
        }
        if (removed < nm) {
            // alive vertices exist, so some degree was recorded
            if (nA == 0 || mnall == UINT_MAX) {
                fail(kErrInconsistent);
                break;
            }
            L = max(L + 1, (int32_t)min(mnall, (uint32_t)INT_MAX));
            if (L > P.level_cap) break;  // survivors keep lval = -1
        }
    }
    if (wg == 0 && tid == 0 && lmax_out) *lmax_out = nm ? L : 0;
}

template <int BLK>
__global__ __launch_bounds__(BLK) void k_pass_grid(PassSpec P, PassScratch S, uint32_t* slots,
                                                   int32_t* lmax_out, GridSync gs) {
    __shared__ TileLds<BLK> T;
    run_pass<BLK, true>(P, S, slots, T, lmax_out, &gs, nullptr);
}

// Persistent rows kernel: workgroup b owns scratch slot b and pulls k
// values from a device counter (largest passes first) until k > kmax.
template <int BLK>
__global__ __launch_bounds__(BLK) void k_rows_batched(PassSpec base, void* scratch,
                                                      const uint32_t* vk, int32_t kmax,
                                                      int32_t* rows, int32_t* lmax_rows,
                                                      uint32_t* qctr, uint32_t* err) {
    __shared__ TileLds<BLK> T;
    __shared__ uint32_t slots[12];
    __shared__ int32_t s_k;
    const PassScratch S = scratch_slot(scratch, base.V, (int)blockIdx.x);
    for (;;) {
        if (threadIdx.x == 0) s_k = (int32_t)atomicAdd(qctr, 1u);
        __syncthreads();
        const int32_t k = s_k;
        __syncthreads();
        if (k > kmax) break;
        PassSpec P = base;
        P.k = k;
        P.nmembers = vk[k];
        P.lval = rows + (size_t)k * base.V;
        run_pass<BLK, false>(P, S, slots, T, lmax_rows + k, nullptr, err);
        __syncthreads();
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
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, W));
    if (lane_id() == 0) atomicMax(out, m);
}

constexpr int kGridBlk = 256;
constexpr int kRowsBlk = 512;

}  // namespace

int pass_grid_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_pass_grid<kGridBlk>, kGridBlk, 0) !=
        hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

hipError_t launch_pass_grid(const PassSpec& p, const PassScratch& s, uint32_t* ctrs,
                            int32_t* lmax_out, const GridSync& gs, hipStream_t st) {
    // the sync words must start at zero for every launch
    hipError_t e = hipMemsetAsync(gs.w, 0, kSyncResetWords * sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pass_grid<kGridBlk>, dim3(gs.nblocks), dim3(kGridBlk), 0, st, p, s, ctrs,
                       lmax_out, gs);
    return hipGetLastError();
}

hipError_t launch_rows_batched(const PassSpec& base, void* scratch, int slots, const uint32_t* vk,
                               int32_t kmax, int32_t* rows, int32_t* lmax_rows, uint32_t* qctr,
                               uint32_t* err, hipStream_t st) {
    hipLaunchKernelGGL(k_rows_batched<kRowsBlk>, dim3(slots), dim3(kRowsBlk), 0, st, base,
                       scratch, vk, kmax, rows, lmax_rows, qctr, err);
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
