// Core-ordered graph: vertices relabelled by kmax (descending), every list
// ordered by its neighbours' kmax, descending, and holding new ids.  Then
// for every k the (k,0)-core is the id prefix [0, V_k), and within a
// member's list the neighbours inside the core form a prefix too (new id <
// V_k exactly when kmax >= k) -- a k-pass touches exactly the edges of the
// core (E_k per direction) and needs no membership test.  Built once per
// decomposition, after the kmax pass: the lists are moved to their new
// owners' places and ordered there.  A list of at most 64 entries is
// sorted by new id on one wavefront (new-id order is kmax order, so the
// neighbour's kmax need not be read); a longer one by a counting sort on
// kmax in one workgroup (ties in no fixed order), from one 8-byte gather
// per entry of a packed (kmax, new id) word.  Graphs whose kmax exceeds the
// LDS histogram sort every list fully by new id with one radix sort per
// direction instead.
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

// rank[v] = new id of v and kr[v] = (kmax(v) << 32) | new id.
__global__ void k_rank_packed(const uint32_t* order, const int32_t* kmaxes, uint32_t V,
                              uint32_t* rank, uint64_t* kr) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
        const uint32_t v = order[i];
        rank[v] = i;
        kr[v] = ((uint64_t)(uint32_t)kmaxes[v] << 32) | i;
    }
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

constexpr uint32_t kShortList = 64;
constexpr uint32_t kTinyList = 16;
constexpr int kListBins = 4096;  // kmax + 1 at most, for the counting sort
constexpr int kLongBlk = 256;

// kmax of the vertex with new id r: the largest k with V_k > r (svk: V_k
// for k = 0..kmax, non-increasing, V_0 = V > r; kmax is non-increasing in
// the new id, so min(kmax of ids i and r) is the kmax of id max(i, r)).
__device__ __forceinline__ int32_t kmax_of_id(const uint32_t* svk, int32_t kmax, uint32_t r) {
    int32_t lo = 0, hi = kmax;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (svk[mid] > r)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Loads V_k into LDS and clears the block's E_k bins (EK only).
template <bool EK, typename H>
__device__ __forceinline__ void ek_begin(const uint32_t* vk, int32_t kmax, uint32_t* svk, H* seh) {
    if (!EK) return;
    for (int32_t b = threadIdx.x; b <= kmax; b += blockDim.x) {
        svk[b] = vk[b];
        seh[b] = 0;
    }
    __syncthreads();
}

// Adds the block's E_k bins to ekh (EK only).
template <bool EK, typename H>
__device__ __forceinline__ void ek_end(int32_t kmax, const H* seh, unsigned long long* ekh) {
    if (!EK) return;
    __syncthreads();
    for (int32_t b = threadIdx.x; b <= kmax; b += blockDim.x)
        if (seh[b]) atomicAdd(&ekh[b], (unsigned long long)seh[b]);
}

// Place of a list entry with new id r held by lane l of a wavefront
// holding the list's n entries: #{j : r_j < r} + #{j < l : r_j == r}
// (ascending, repeated entries in list order).
__device__ __forceinline__ uint32_t list_place(uint32_t r, uint32_t l, uint32_t n) {
    uint32_t pos = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t rj = (uint32_t)__shfl((int)r, (int)j, W);
        pos += (rj < r || (rj == r && j < l)) ? 1u : 0u;
    }
    return pos;
}

// Lists of at most kShortList entries: a wavefront takes four consecutive
// new owners at a time, one per 16-lane quarter when all four lists have at
// most kTinyList entries, else one list after another on all 64 lanes.
// Longer lists are queued (new owner id) for k_lists_long.  EK (out side):
// every entry also adds 1 to ekh[min of its endpoints' kmax], the E_k
// histogram (E_k = edges inside the (k,0)-core = its suffix sums).
template <bool EK>
__global__ __launch_bounds__(1024) void k_lists_short(const uint32_t* order, const uint32_t* rank,
                                                      uint32_t V, const uint32_t* off,
                                                      const int32_t* deg, const uint32_t* nbr,
                                                      const uint32_t* off_new, uint32_t* out,
                                                      uint32_t* longq, uint32_t* longn,
                                                      const uint32_t* vk, int32_t kmax,
                                                      unsigned long long* ekh) {
    __shared__ uint32_t svk[EK ? kListBins : 1];
    __shared__ unsigned long long seh[EK ? kListBins : 1];
    ek_begin<EK>(vk, kmax, svk, seh);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / W;
    const uint32_t nw = gridDim.x * blockDim.x / W;
    const uint32_t lane = (uint32_t)lane_id();
    for (uint32_t i0 = wave * 4; i0 < V; i0 += nw * 4) {
        // the four owners' lists: quarter q looks up owner i0 + q
        const uint32_t q = lane / kTinyList, l = lane % kTinyList;
        const uint32_t iq = i0 + q;
        const uint32_t vq = iq < V ? order[iq] : 0u;
        const uint32_t nq = iq < V ? (uint32_t)deg[vq] : 0u;
        const bool tiny = __all(nq <= kTinyList);
        if (tiny) {
            const bool has = l < nq;
            const uint32_t u = has ? nbr[off[vq] + l] : 0u;
            const uint32_t r = has ? rank[u] : 0u;
            // every quarter runs the longest list's count (entries past a
            // shorter list's end are not counted, nor written)
            const uint32_t nmax = __reduce_max_sync(~0ull, nq);
            uint32_t pos = 0;
            for (uint32_t j = 0; j < nmax; ++j) {
                const uint32_t rj = (uint32_t)__shfl((int)r, (int)j, (int)kTinyList);
                pos += (j < nq && (rj < r || (rj == r && j < l))) ? 1u : 0u;
            }
            if (has) {
                out[off_new[iq] + pos] = r;
                if (EK) atomicAdd(&seh[kmax_of_id(svk, kmax, max(iq, r))], 1ull);
            }
            continue;
        }
        for (uint32_t t = 0; t < 4 && i0 + t < V; ++t) {
            const uint32_t i = i0 + t;
            const uint32_t v = __shfl(vq, (int)(t * kTinyList), W);
            const uint32_t n = __shfl(nq, (int)(t * kTinyList), W);
            if (n > kShortList) {
                if (lane == 0) longq[atomicAdd(longn, 1u)] = i;
                continue;
            }
            const bool has = lane < n;
            const uint32_t u = has ? nbr[off[v] + lane] : 0u;
            const uint32_t r = has ? rank[u] : 0u;
            const uint32_t pos = list_place(r, lane, n);
            if (has) {
                out[off_new[i] + pos] = r;
                if (EK) atomicAdd(&seh[kmax_of_id(svk, kmax, max(i, r))], 1ull);
            }
        }
    }
    ek_end<EK>(kmax, seh, ekh);
}

// Entries of a long list each thread keeps in registers between the
// histogram and the placement (longer lists gather the rest twice).
constexpr int kLongRegs = 8;

// The queued long lists, one workgroup each: LDS histogram of the
// neighbours' kmax (and its range), start of every bin of the range in
// descending order (kLongBlk bins per step), then each entry placed at its
// bin's cursor; the range's bins are cleared for the next list.  kr: the
// packed (kmax, new id) word of every vertex.  EK (out side): the list's
// entries are also added to the E_k histogram ekh, from the kmax bins:
// bin b < kmax(v) gets the entries of kmax b, bin kmax(v) the rest.
template <bool EK>
__global__ __launch_bounds__(kLongBlk) void k_lists_long(const uint32_t* order, const uint64_t* kr,
                                                         const uint32_t* off, const int32_t* deg,
                                                         const uint32_t* nbr,
                                                         const uint32_t* off_new, uint32_t* out,
                                                         const uint32_t* longq,
                                                         const uint32_t* longn, int32_t kmax,
                                                         unsigned long long* ekh) {
    __shared__ uint32_t cnt[kListBins];
    __shared__ uint32_t seh[EK ? kListBins : 1];
    if (EK)
        for (int b = threadIdx.x; b <= kmax; b += kLongBlk) seh[b] = 0;
    __shared__ uint32_t w[kLongBlk / W];
    __shared__ int32_t s_lo, s_hi;
    __shared__ uint32_t s_carry;
    const uint32_t nq = *longn;
    const uint32_t tid = threadIdx.x;
    for (int b = tid; b < kListBins; b += kLongBlk) cnt[b] = 0;
    if (tid == 0) {
        s_lo = kListBins;
        s_hi = -1;
    }
    __syncthreads();
    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const uint32_t i = longq[q];
        const uint32_t v = order[i];
        const uint32_t n = (uint32_t)deg[v];
        const uint32_t* l = nbr + off[v];
        uint32_t* dst = out + off_new[i];
        int32_t lo = kListBins, hi = -1;
        uint64_t held[kLongRegs];
#pragma unroll
        for (int h = 0; h < kLongRegs; ++h) {
            const uint32_t j = tid + (uint32_t)h * kLongBlk;
            held[h] = j < n ? kr[l[j]] : 0ull;
        }
#pragma unroll
        for (int h = 0; h < kLongRegs; ++h) {
            if (tid + (uint32_t)h * kLongBlk >= n) break;
            const int32_t k = (int32_t)(held[h] >> 32);
            atomicAdd(&cnt[k], 1u);
            lo = min(lo, k);
            hi = max(hi, k);
        }
        for (uint32_t j = tid + kLongRegs * kLongBlk; j < n; j += kLongBlk) {
            const int32_t k = (int32_t)(kr[l[j]] >> 32);
            atomicAdd(&cnt[k], 1u);
            lo = min(lo, k);
            hi = max(hi, k);
        }
        atomicMin(&s_lo, lo);
        atomicMax(&s_hi, hi);
        if (tid == 0) s_carry = 0;
        __syncthreads();
        lo = s_lo;
        hi = s_hi;
        if (EK) {
            // (read before the scan below turns the counts into cursors)
            const int32_t kv = (int32_t)(kr[v] >> 32);
            uint32_t above = 0;
            for (int b = lo + (int)tid; b <= hi; b += kLongBlk) {
                const uint32_t c = cnt[b];
                if (b < kv)
                    seh[b] += c;
                else
                    above += c;
            }
            if (above) atomicAdd(&seh[kv], above);
            __syncthreads();
        }
        // starts in descending kmax, kLongBlk bins per step
        for (int top = hi; top >= lo; top -= kLongBlk) {
            const int b = top - (int)tid;
            const uint32_t c = b >= lo ? cnt[b] : 0u;
            // exclusive scan over the step (wave shuffles, then waves)
            const int ln = lane_id(), wv = (int)(tid / W);
            uint32_t inc = c;
            for (int d = 1; d < W; d <<= 1) {
                const uint32_t y = __shfl_up(inc, d, W);
                if (ln >= d) inc += y;
            }
            if (ln == W - 1) w[wv] = inc;
            __syncthreads();
            uint32_t before = s_carry;
            for (int x = 0; x < wv; ++x) before += w[x];
            if (b >= lo) cnt[b] = before + inc - c;
            __syncthreads();
            if (tid == kLongBlk - 1) s_carry = before + inc;
            __syncthreads();
        }
#pragma unroll
        for (int h = 0; h < kLongRegs; ++h) {
            if (tid + (uint32_t)h * kLongBlk >= n) break;
            dst[atomicAdd(&cnt[(int32_t)(held[h] >> 32)], 1u)] = (uint32_t)held[h];
        }
        for (uint32_t j = tid + kLongRegs * kLongBlk; j < n; j += kLongBlk) {
            const uint64_t w = kr[l[j]];
            dst[atomicAdd(&cnt[(int32_t)(w >> 32)], 1u)] = (uint32_t)w;
        }
        __syncthreads();
        for (int b = lo + (int)tid; b <= hi; b += kLongBlk) cnt[b] = 0;
        if (tid == 0) {
            s_lo = kListBins;
            s_hi = -1;
        }
        __syncthreads();
    }
    ek_end<EK>(kmax, seh, ekh);
}

// out[0] += sum of deg, out[1] += entries of [0, deg) that are not vertex
// ids, plus the negative degrees.  Tiles of kAuditTile consecutive vertices:
// their list lengths are prefix-summed in LDS and the tile's entries are
// dealt over all threads (owner found by binary search), so a tile of short
// lists keeps every lane busy.  Lists of more than kAuditLong entries are
// cut into slices of kAuditSlice, queued (v << 32 | first entry) for
// k_audit_long, one wavefront per slice: R-MAT's hubs and dense vertices
// have the lowest ids, and a tile of them would leave one block with a
// large share of all edges.
// A list that would reach past the nbr array (nlen entries), or whose
// slices do not fit the queue (qcap), is counted bad and not read: the
// audit itself stays in bounds whatever the degrees hold.
constexpr int kAuditTile = 1024;
constexpr uint32_t kAuditLong = 256;
constexpr uint32_t kAuditSlice = 2048;
__global__ __launch_bounds__(kAuditTile) void k_audit(uint32_t V, const uint32_t* off,
                                                      const int32_t* deg, const uint32_t* nbr,
                                                      uint64_t nlen, unsigned long long* out,
                                                      unsigned long long* longq, uint32_t* longn,
                                                      uint32_t qcap) {
    __shared__ uint32_t s_off[kAuditTile], s_scan[kAuditTile];
    __shared__ uint32_t s_wsum[kAuditTile / W];
    __shared__ unsigned long long s_sum, s_bad;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) s_sum = s_bad = 0;
    __syncthreads();
    uint64_t sum = 0, bad = 0;
    for (uint32_t base = blockIdx.x * kAuditTile; base < V; base += gridDim.x * kAuditTile) {
        const uint32_t v = base + tid;
        uint32_t n = 0;
        if (v < V) {
            const int32_t d = deg[v];
            const uint32_t o = off[v];
            if (d < 0) {
                ++bad;
            } else {
                sum += (uint32_t)d;
                if ((uint64_t)o + (uint32_t)d > nlen)
                    ++bad;  // not read
                else
                    n = (uint32_t)d;
            }
            s_off[tid] = o;
            if (n > kAuditLong) {
                const uint32_t ns = (n + kAuditSlice - 1) / kAuditSlice;
                const uint32_t at = atomicAdd(longn, ns);
                if ((uint64_t)at + ns <= qcap) {
                    for (uint32_t i = 0; i < ns; ++i)
                        longq[at + i] = ((unsigned long long)v << 32) | (i * kAuditSlice);
                } else {
                    ++bad;  // not read
                }
                n = 0;
            }
        }
        // block-wide exclusive scan of n (lengths < 2^32 per tile: a tile's
        // lists fit the 32-bit list offsets)
        uint32_t inc = n;
        for (int d = 1; d < W; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d, W);
            if (lane_id() >= d) inc += y;
        }
        if (lane_id() == W - 1) s_wsum[tid / W] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < kAuditTile / W; ++w) {
            const uint32_t x = s_wsum[w];
            if (w < (int)(tid / W)) before += x;
            total += x;
        }
        s_scan[tid] = before + inc - n;
        __syncthreads();
        const uint32_t items = min(V - base, (uint32_t)kAuditTile);
        for (uint32_t e = tid; e < total; e += kAuditTile) {
            uint32_t lo = 0, hi = items - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (s_scan[mid] <= e)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            bad += nbr[s_off[lo] + (e - s_scan[lo])] >= V ? 1u : 0u;
        }
        __syncthreads();
    }
    for (int dd = W / 2; dd > 0; dd >>= 1) {
        sum += __shfl_xor(sum, dd, W);
        bad += __shfl_xor(bad, dd, W);
    }
    // one pair of atomics per block (same-address atomics serialise)
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

// The queued slices of long lists, one wavefront each: out[1] += their
// entries >= V.
__global__ __launch_bounds__(kAuditTile) void k_audit_long(uint32_t V, const uint32_t* off,
                                                           const int32_t* deg, const uint32_t* nbr,
                                                           unsigned long long* out,
                                                           const unsigned long long* longq,
                                                           const uint32_t* longn, uint32_t qcap) {
    __shared__ unsigned long long s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    const uint32_t nq = min(*longn, qcap);  // entries past qcap were not written
    uint64_t bad = 0;
    const uint32_t nw = gridDim.x * (kAuditTile / W);
    for (uint32_t q = blockIdx.x * (kAuditTile / W) + threadIdx.x / W; q < nq; q += nw) {
        const unsigned long long e = longq[q];
        const uint32_t v = (uint32_t)(e >> 32), j0 = (uint32_t)e;
        const uint32_t* l = nbr + off[v];
        const uint32_t j1 = min((uint32_t)deg[v], j0 + kAuditSlice);
        for (uint32_t j = j0 + lane_id(); j < j1; j += W) bad += l[j] >= V ? 1u : 0u;
    }
    for (int dd = W / 2; dd > 0; dd >>= 1) bad += __shfl_xor(bad, dd, W);
    if (lane_id() == 0 && bad) atomicAdd(&s_bad, (unsigned long long)bad);
    __syncthreads();
    if (threadIdx.x == 0 && s_bad) atomicAdd(out + 1, s_bad);
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

// Scratch words of one side's list build: scan words, the long-list
// queue's length (64 words, keeping the queue 256-byte aligned) and queue.
size_t side_tmp_words(uint32_t V, uint64_t E) {
    return scan_tmp_words(V) + 1 + 64 + ((E ? E : 1) / kShortList + 64);
}

size_t core_graph_tmp_bytes(uint32_t V, uint64_t E, int32_t kmax) {
    if (kmax + 1 <= kListBins)  // packed words; per side scan words, long-list queue, its length
        return 8ull * V + 2 * 4ull * (side_tmp_words(V, E)) + 256;
    // keys + alternate keys + scan words + radix table
    return 16ull * (E ? E : 1) + 4ull * (scan_tmp_words(V) + 1) +
           4ull * (radix_tmp_words(E) + 1) + 256;
}

// Slices of a valid side: at most E / kAuditLong lists are longer than
// kAuditLong, each cut into at most 1 + n / kAuditSlice slices.
static uint32_t audit_qcap(uint64_t E) {
    return (uint32_t)std::min<uint64_t>(E / kAuditLong + E / kAuditSlice + 64, 0xffffffffu);
}

size_t audit_tmp_bytes(uint64_t E) { return 8ull * audit_qcap(E) + 256; }

hipError_t launch_audit(uint32_t V, Csr side, uint64_t nlen, uint64_t E, unsigned long long* out,
                        void* tmp, hipStream_t st) {
    if (V == 0) return hipSuccess;
    uint32_t* longn = static_cast<uint32_t*>(tmp);
    unsigned long long* longq = reinterpret_cast<unsigned long long*>(static_cast<char*>(tmp) + 256);
    const uint32_t qcap = audit_qcap(E);
    hipError_t e = hipMemsetAsync(longn, 0, 4, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_audit, dim3(grid_for(V, kAuditTile, 2048)), dim3(kAuditTile), 0, st, V,
                       side.off, side.deg, side.nbr, nlen, out, longq, longn, qcap);
    hipLaunchKernelGGL(k_audit_long, dim3(2048), dim3(kAuditTile), 0, st, V, side.off, side.deg,
                       side.nbr, out, longq, longn, qcap);
    return hipGetLastError();
}

hipError_t build_core_graph(uint32_t V, uint64_t E, const uint32_t* order, const int32_t* kmaxes,
                            int32_t kmax, Csr out, Csr in, uint32_t* rank, uint32_t* cout_off,
                            int32_t* cout_deg, uint32_t* cout_nbr, uint32_t* cin_off,
                            int32_t* cin_deg, uint32_t* cin_nbr, void* tmp, const uint32_t* vk,
                            unsigned long long* ekh, hipStream_t st, hipStream_t st2,
                            hipEvent_t fork, hipEvent_t join) {
    if (V == 0) return hipSuccess;
    char* p = static_cast<char*>(tmp);
    if (kmax + 1 <= kListBins) {
        uint64_t* kr = reinterpret_cast<uint64_t*>(p);
        hipLaunchKernelGGL(k_rank_packed, dim3(grid_for(V, 256)), dim3(256), 0, st, order, kmaxes,
                           V, rank, kr);
        if (ekh) {
            const hipError_t me =
                hipMemsetAsync(ekh, 0, ((size_t)kmax + 1) * sizeof(unsigned long long), st);
            if (me != hipSuccess) return me;
        }
        // the in side on st2 (when given) beside the out side: both are
        // bound by their random gathers, and each side's long lists end
        // on a few workgroups that the other side's kernels fill in beside
        const bool two = st2 && fork && join;
        if (two) {
            hipError_t e = hipEventRecord(fork, st);
            if (e == hipSuccess) e = hipStreamWaitEvent(st2, fork, 0);
            if (e != hipSuccess) return e;
        }
        const Csr sides[2] = {out, in};
        uint32_t* offs[2] = {cout_off, cin_off};
        int32_t* degs[2] = {cout_deg, cin_deg};
        uint32_t* nbrs[2] = {cout_nbr, cin_nbr};
        for (int s = 0; s < 2; ++s) {
            const hipStream_t ss = s == 1 && two ? st2 : st;
            uint32_t* scan_tmp = reinterpret_cast<uint32_t*>(kr + V) + s * side_tmp_words(V, E);
            uint32_t* longn = scan_tmp + scan_tmp_words(V) + 1;
            uint32_t* longq = longn + 64;  // 256-byte aligned from longn
            uint32_t* dnew = reinterpret_cast<uint32_t*>(degs[s]);
            hipLaunchKernelGGL(k_new_degrees, dim3(grid_for(V, 256)), dim3(256), 0, ss, order, V,
                               sides[s].deg, dnew);
            hipError_t e = launch_excl_scan_u32(dnew, offs[s], V, scan_tmp, ss);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, ss, offs[s] + V, (uint32_t)E);
            if (E == 0) continue;
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, ss, longn, 0u);
            if (s == 0 && ekh) {
                hipLaunchKernelGGL(k_lists_short<true>, dim3(grid_for(V, 64, 2048)), dim3(1024),
                                   0, ss, order, rank, V, sides[s].off, sides[s].deg,
                                   sides[s].nbr, offs[s], nbrs[s], longq, longn, vk, kmax, ekh);
                hipLaunchKernelGGL(k_lists_long<true>, dim3(4096), dim3(kLongBlk), 0, ss, order,
                                   kr, sides[s].off, sides[s].deg, sides[s].nbr, offs[s], nbrs[s],
                                   longq, longn, kmax, ekh);
            } else {
                hipLaunchKernelGGL(k_lists_short<false>, dim3(grid_for(V, 64)), dim3(1024), 0, ss,
                                   order, rank, V, sides[s].off, sides[s].deg, sides[s].nbr,
                                   offs[s], nbrs[s], longq, longn, vk, kmax, nullptr);
                hipLaunchKernelGGL(k_lists_long<false>, dim3(4096), dim3(kLongBlk), 0, ss, order,
                                   kr, sides[s].off, sides[s].deg, sides[s].nbr, offs[s], nbrs[s],
                                   longq, longn, kmax, nullptr);
            }
        }
        if (two) {
            hipError_t e = hipEventRecord(join, st2);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, join, 0);
            if (e != hipSuccess) return e;
        }
        return hipGetLastError();
    }
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
