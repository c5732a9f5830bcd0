// Device primitives for building the core-ordered graph: exclusive scan of
// uint32 arrays and a stable LSD radix sort of uint64 keys.
//
// Radix sort: 4-bit digits.  Per pass, every block owns a tile of
// kTile = kSortBlk * kPerThread consecutive keys: (1) it histograms the
// digit of its tile, (2) one block scans the digit-major table of tile
// histograms, (3) every block re-reads its tile and writes each key to
// table[digit][tile] + (rank of the key among equal digits before it in
// the tile) -- ranks come from a per-thread count table scanned in LDS,
// which keeps the sort stable.
#include <hip/hip_runtime.h>

#include "dcx_internal.hpp"

namespace dcx {
namespace {

constexpr int kSortBlk = 256;
constexpr int kPerThread = 8;
constexpr int kTile = kSortBlk * kPerThread;
constexpr int kDigits = 16;
constexpr int W = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (W - 1)); }

// In-LDS exclusive scan of n <= 4096 uint32 values by one 256-thread block.
__device__ void lds_excl_scan(uint32_t* a, int n) {
    __shared__ uint32_t part[kSortBlk];
    const int t = threadIdx.x;
    const int per = (n + kSortBlk - 1) / kSortBlk;
    const int lo = t * per, hi = min(n, lo + per);
    uint32_t s = 0;
    for (int i = lo; i < hi; ++i) s += a[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < kSortBlk; ++i) {
            const uint32_t x = part[i];
            part[i] = acc;
            acc += x;
        }
    }
    __syncthreads();
    uint32_t acc = part[t];
    for (int i = lo; i < hi; ++i) {
        const uint32_t x = a[i];
        a[i] = acc;
        acc += x;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kSortBlk) void k_radix_hist(const uint64_t* keys, uint64_t n,
                                                          int shift, uint32_t ntiles,
                                                          uint32_t* table) {
    __shared__ uint32_t h[kDigits];
    if (threadIdx.x < kDigits) h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    for (int j = 0; j < kPerThread; ++j) {
        const uint64_t i = base + (uint64_t)j * kSortBlk + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kDigits - 1)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kDigits) table[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the digit-major table (kDigits * ntiles entries) by one
// block: each thread scans a contiguous chunk, then chunk sums are scanned.
__global__ __launch_bounds__(kSortBlk) void k_radix_scan(uint32_t* table, uint64_t m) {
    __shared__ uint32_t part[kSortBlk];
    const int t = threadIdx.x;
    const uint64_t per = (m + kSortBlk - 1) / kSortBlk;
    const uint64_t lo = t * per, hi = min(m, lo + per);
    uint32_t s = 0;
    for (uint64_t i = lo; i < hi; ++i) s += table[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < kSortBlk; ++i) {
            const uint32_t x = part[i];
            part[i] = acc;
            acc += x;
        }
    }
    __syncthreads();
    uint32_t acc = part[t];
    for (uint64_t i = lo; i < hi; ++i) {
        const uint32_t x = table[i];
        table[i] = acc;
        acc += x;
    }
}

__global__ __launch_bounds__(kSortBlk) void k_radix_scatter(const uint64_t* in, uint64_t* out,
                                                             uint64_t n, int shift, uint32_t ntiles,
                                                             const uint32_t* table) {
    // cnt[d * kSortBlk + t]: keys of digit d held by thread t (thread t owns
    // the tile's items t*kPerThread .. +kPerThread, so thread order is key
    // order and the scan below gives stable ranks)
    __shared__ uint32_t cnt[kDigits * kSortBlk];
    const int t = threadIdx.x;
    for (int d = 0; d < kDigits; ++d) cnt[d * kSortBlk + t] = 0;
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)t * kPerThread;
    uint64_t k[kPerThread];
    for (int j = 0; j < kPerThread; ++j) {
        const uint64_t i = base + j;
        k[j] = i < n ? in[i] : 0;
        if (i < n) cnt[((k[j] >> shift) & (kDigits - 1)) * kSortBlk + t]++;
    }
    __syncthreads();
    lds_excl_scan(cnt, kDigits * kSortBlk);
    // cnt[d][t] = keys of the tile with a smaller digit + keys of digit d in
    // threads before t; the digit's own start is cnt[d][0]
    __shared__ uint32_t dstart[kDigits];
    if (t < kDigits) dstart[t] = cnt[t * kSortBlk];
    __syncthreads();
    for (int j = 0; j < kPerThread; ++j) {
        const uint64_t i = base + j;
        if (i < n) {
            const int d = (int)((k[j] >> shift) & (kDigits - 1));
            const uint32_t r = cnt[d * kSortBlk + t]++ - dstart[d];
            out[(uint64_t)table[(uint64_t)d * ntiles + blockIdx.x] + r] = k[j];
        }
    }
}

// ---- exclusive scan of uint32 arrays (3 phases) -------------------------

__global__ __launch_bounds__(kSortBlk) void k_scan_tiles(const uint32_t* in, uint32_t* out,
                                                          uint64_t n, uint32_t* sums) {
    __shared__ uint32_t a[kTile];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    for (int j = 0; j < kPerThread; ++j) {
        const int li = j * kSortBlk + threadIdx.x;
        const uint64_t i = base + li;
        a[li] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    // thread t scans its kPerThread consecutive values
    const int t = threadIdx.x;
    uint32_t s = 0;
    for (int j = 0; j < kPerThread; ++j) s += a[t * kPerThread + j];
    __shared__ uint32_t part[kSortBlk];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < kSortBlk; ++i) {
            const uint32_t x = part[i];
            part[i] = acc;
            acc += x;
        }
        sums[blockIdx.x] = acc;
    }
    __syncthreads();
    uint32_t acc = part[t];
    for (int j = 0; j < kPerThread; ++j) {
        const uint32_t x = a[t * kPerThread + j];
        a[t * kPerThread + j] = acc;
        acc += x;
    }
    __syncthreads();
    for (int j = 0; j < kPerThread; ++j) {
        const int li = j * kSortBlk + threadIdx.x;
        const uint64_t i = base + li;
        if (i < n) out[i] = a[li];
    }
}

__global__ __launch_bounds__(kSortBlk) void k_scan_add(uint32_t* out, uint64_t n,
                                                        const uint32_t* sums) {
    const uint32_t add = sums[blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    for (int j = 0; j < kPerThread; ++j) {
        const uint64_t i = base + (uint64_t)j * kSortBlk + threadIdx.x;
        if (i < n) out[i] += add;
    }
}

uint32_t tiles_for(uint64_t n) { return (uint32_t)((n + kTile - 1) / kTile); }

}  // namespace

size_t scan_tmp_words(uint64_t n) { return (size_t)tiles_for(n) + 1; }

hipError_t launch_excl_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp,
                                hipStream_t st) {
    if (!n) return hipSuccess;
    const uint32_t nt = tiles_for(n);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kSortBlk), 0, st, in, out, n, tmp);
    hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(kSortBlk), 0, st, tmp, (uint64_t)nt);
    hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(kSortBlk), 0, st, out, n, tmp);
    return hipGetLastError();
}

// digit table (kDigits * tiles) + 1, then the tile sums of its scan
size_t radix_tmp_words(uint64_t n) {
    const uint64_t m = (uint64_t)kDigits * tiles_for(n);
    return (size_t)m + 1 + scan_tmp_words(m) + 1;
}

hipError_t launch_radix_sort_u64(uint64_t* keys, uint64_t* alt, uint64_t n, int bits,
                                 uint32_t* table, uint64_t** result, hipStream_t st) {
    *result = keys;
    if (n < 2) return hipSuccess;
    const uint32_t nt = tiles_for(n);
    uint64_t* a = keys;
    uint64_t* b = alt;
    const uint64_t m = (uint64_t)kDigits * nt;
    uint32_t* sums = table + m + 1;
    for (int shift = 0; shift < bits; shift += 4) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nt), dim3(kSortBlk), 0, st, a, n, shift, nt, table);
        // the digit table scanned by tiles (one block per tile, in place)
        hipError_t e = launch_excl_scan_u32(table, table, m, sums, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(nt), dim3(kSortBlk), 0, st, a, b, n, shift, nt,
                           table);
        uint64_t* t = a;
        a = b;
        b = t;
    }
    *result = a;
    return hipGetLastError();
}

}  // namespace dcx
