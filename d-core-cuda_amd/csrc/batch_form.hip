// Edge batches by root kmax on the device: getKmaxEdgeBatches
// (main.cu:1441-1525) computed from the resident kmaxes, without a host
// round trip of the kmaxes.
//
// Root of an edge: the endpoint with the larger kmax, the target on ties
// (main.cu:1449-1451); its kmax is the edge's bucket.
//   KEDGE (mode 0, the reference's active mode): the reference takes, in
//     repeated passes over a bucket's edges in input order, every edge
//     whose root is not yet in the pass's batch -- so batch t of a bucket
//     holds the t-th edge (input order) of every root that has one.  Here:
//     edges sorted by (root, input index) (stable radix sort), the rank of
//     each within its root's run, and batch = (batches of the lower
//     buckets) + rank, the lower buckets' batch counts being an exclusive
//     scan of each bucket's largest rank + 1.
//   KMAX (mode 1): round after round, the smallest non-empty bucket gives
//     its next edge, and so does every other non-empty bucket whose kmax
//     differs from it by more than one.  One workgroup plays the rounds
//     over the edges sorted by (bucket, input index).
#include <hip/hip_runtime.h>

#include "dcx_internal.hpp"

namespace dcx {
namespace {

constexpr int kBlk = 256;
constexpr int kRoundBlk = 1024;

unsigned grid_for(uint64_t n) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlk - 1) / kBlk, 4096));
}

__global__ void k_eb_keys(const int32_t* kmaxes, const uint32_t* pairs, uint32_t n, int by_root,
                          uint64_t* keys, int32_t* bucket, int32_t* maxb) {
    int32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t u = pairs[2 * i], v = pairs[2 * i + 1];
        const uint32_t root = kmaxes[u] > kmaxes[v] ? u : v;
        const int32_t b = kmaxes[root];
        bucket[i] = b;
        m = max(m, b);
        keys[i] = ((uint64_t)(by_root ? root : (uint32_t)b) << 32) | i;
    }
    atomicMax(maxb, m);
}

// flags[j] = 1 where a run of equal high words starts
__global__ void k_eb_flags(const uint64_t* keys, uint32_t n, uint32_t* flags) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        flags[j] = (j == 0 || (keys[j] >> 32) != (keys[j - 1] >> 32)) ? 1u : 0u;
}

__global__ void k_eb_starts(const uint32_t* flags, const uint32_t* gidx, uint32_t n,
                            uint32_t* start) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        if (flags[j]) start[gidx[j]] = j;
}

// rank of each edge within its root's run; per bucket the number of its
// batches (largest rank + 1)
__global__ void k_eb_rank(const uint64_t* keys, const uint32_t* flags, const uint32_t* gidx,
                          const uint32_t* start, const int32_t* bucket, uint32_t n,
                          uint32_t* rank, uint32_t* nb) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t g = gidx[j] + flags[j] - 1;
        const uint32_t i = (uint32_t)keys[j];
        const uint32_t r = j - start[g];
        rank[i] = r;
        atomicMax(&nb[bucket[i]], r + 1);
    }
}

__global__ void k_eb_final(const uint32_t* rank, const int32_t* bucket, const uint32_t* boff,
                           uint32_t n, uint32_t* batch) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        batch[i] = boff[bucket[i]] + rank[i];
}

__global__ void k_eb_count(const int32_t* bucket, uint32_t n, uint32_t* cnt) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&cnt[bucket[i]], 1u);
}

// KMAX rounds on one workgroup.  head[b] counts the edges bucket b gave.
__global__ __launch_bounds__(kRoundBlk) void k_eb_kmax_rounds(const uint64_t* keys,
                                                               const uint32_t* bstart,
                                                               const uint32_t* cnt,
                                                               uint32_t* head, uint32_t nbuck,
                                                               uint32_t* batch,
                                                               uint32_t* rounds) {
    __shared__ uint32_t s_min;
    for (uint32_t round = 0;; ++round) {
        if (threadIdx.x == 0) s_min = 0xffffffffu;
        __syncthreads();
        uint32_t m = 0xffffffffu;
        for (uint32_t b = threadIdx.x; b < nbuck; b += kRoundBlk)
            if (head[b] < cnt[b]) {
                m = b;
                break;
            }
        if (m != 0xffffffffu) atomicMin(&s_min, m);
        __syncthreads();
        const uint32_t j0 = s_min;
        if (j0 == 0xffffffffu) {
            if (threadIdx.x == 0) *rounds = round;
            return;
        }
        for (uint32_t b = threadIdx.x; b < nbuck; b += kRoundBlk) {
            if (head[b] >= cnt[b]) continue;
            const uint32_t d = b > j0 ? b - j0 : j0 - b;
            if (b != j0 && d <= 1) continue;
            batch[(uint32_t)keys[bstart[b] + head[b]]] = round;
            head[b] += 1;
        }
        __syncthreads();
    }
}

}  // namespace

size_t edge_batch_tmp_bytes(uint32_t n, uint32_t nbuckets) {
    // keys, alt (2 words each), bucket, flags, gidx, rank, start
    const size_t w = 2ull * n + 2ull * n + 5ull * n + 3ull * (nbuckets + 1) +
                     radix_tmp_words(n) + scan_tmp_words(std::max<uint64_t>(n, nbuckets + 1)) + 64;
    return 4 * w;
}

// Device pairs[2n] -> batch[n] (device) and the batch count (device word).
// kmaxes with at most nbuckets - 1 as largest value; tmp of
// edge_batch_tmp_bytes(n, nbuckets).
hipError_t launch_edge_batches(const int32_t* kmaxes, uint32_t V, const uint32_t* pairs, uint32_t n,
                               int mode, uint32_t nbuckets, uint32_t* batch, uint32_t* nbatch,
                               void* tmp, hipStream_t st) {
    if (n == 0) return hipMemsetAsync(nbatch, 0, 4, st);
    uint32_t* w = static_cast<uint32_t*>(tmp);
    uint64_t* keys = reinterpret_cast<uint64_t*>(w);
    uint64_t* alt = keys + n;
    int32_t* bucket = reinterpret_cast<int32_t*>(alt + n);
    uint32_t* flags = reinterpret_cast<uint32_t*>(bucket + n);
    uint32_t* gidx = flags + n;
    uint32_t* rank = gidx + n;
    uint32_t* start = rank + n;
    uint32_t* nb = start + n;           // [nbuckets + 1]
    uint32_t* boff = nb + nbuckets + 1; // [nbuckets + 1]
    uint32_t* head = boff + nbuckets + 1;
    uint32_t* table = head + nbuckets + 1;
    uint32_t* scan_tmp = table + radix_tmp_words(n);
    int32_t* maxb = reinterpret_cast<int32_t*>(scan_tmp + scan_tmp_words(
                                                               std::max<uint64_t>(n, nbuckets + 1)));
    hipError_t e = hipMemsetAsync(nb, 0, 3ull * (nbuckets + 1) * 4, st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(maxb, 0, 4, st);
    if (e != hipSuccess) return e;
    const int by_root = mode == 0;
    hipLaunchKernelGGL(k_eb_keys, dim3(grid_for(n)), dim3(kBlk), 0, st, kmaxes, pairs, n, by_root,
                       keys, bucket, maxb);
    int hi_bits = 0;
    for (uint64_t x = by_root ? (uint64_t)V : (uint64_t)nbuckets; x > 1; x = (x + 1) / 2) ++hi_bits;
    hi_bits = std::max(hi_bits, 1);
    uint64_t* sorted = nullptr;
    e = launch_radix_sort_u64(keys, alt, n, 32 + hi_bits, table, &sorted, st, 32);
    if (e != hipSuccess) return e;
    if (by_root) {
        hipLaunchKernelGGL(k_eb_flags, dim3(grid_for(n)), dim3(kBlk), 0, st, sorted, n, flags);
        e = launch_excl_scan_u32(flags, gidx, n, scan_tmp, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_eb_starts, dim3(grid_for(n)), dim3(kBlk), 0, st, flags, gidx, n,
                           start);
        hipLaunchKernelGGL(k_eb_rank, dim3(grid_for(n)), dim3(kBlk), 0, st, sorted, flags, gidx,
                           start, bucket, n, rank, nb);
        // boff[b] = batches of the buckets below b; boff[nbuckets] = all
        e = launch_excl_scan_u32(nb, boff, nbuckets + 1, scan_tmp, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_eb_final, dim3(grid_for(n)), dim3(kBlk), 0, st, rank, bucket, boff,
                           n, batch);
        e = hipMemcpyAsync(nbatch, boff + nbuckets, 4, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL(k_eb_count, dim3(grid_for(n)), dim3(kBlk), 0, st, bucket, n, nb);
        e = launch_excl_scan_u32(nb, boff, nbuckets + 1, scan_tmp, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_eb_kmax_rounds, dim3(1), dim3(kRoundBlk), 0, st, sorted, boff, nb,
                           head, nbuckets, batch, nbatch);
    }
    return hipGetLastError();
}

}  // namespace dcx
