// Device primitives for building the core-ordered graph: exclusive scan of
// uint32 arrays and a stable LSD radix sort of uint64 keys.
//
// Radix sort: 8-bit digits.  Per pass, every block owns a tile of
// kTile = kSortBlk * kPerThread consecutive keys: (1) it histograms the
// digit of its tile, (2) the digit-major table of tile histograms is
// scanned, (3) every block re-reads its tile in kPerThread rounds of
// kSortBlk consecutive keys and writes each key to table[digit][tile] +
// (keys of that digit in earlier rounds) + (in earlier wavefronts of the
// round) + (in earlier lanes of its wavefront: a popcount of the lanes
// whose digit matches, found with one ballot per digit bit).  Rounds,
// wavefronts and lanes follow the input order, so the sort is stable.
#include <hip/hip_runtime.h>

#include "dcx_internal.hpp"

namespace dcx {
namespace {

constexpr int kSortBlk = 256;
constexpr int kPerThread = 8;
constexpr int kTile = kSortBlk * kPerThread;
constexpr int kDigitBits = 8;
constexpr int kDigits = 1 << kDigitBits;
constexpr int kSortWaves = kSortBlk / 64;
static_assert(kDigits == kSortBlk, "one thread per digit in the scatter's prefix step");
constexpr int W = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (W - 1)); }

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
    __shared__ uint32_t run[kDigits];              // keys of each digit placed so far
    __shared__ uint32_t wc[kSortWaves][kDigits];   // this round: per wavefront, then prefix
    __shared__ uint32_t gbase[kDigits];            // the tile's start in each digit's bucket
    const int t = threadIdx.x, w = t / W, ln = lane_id();
    run[t] = 0;
    gbase[t] = table[(uint64_t)t * ntiles + blockIdx.x];
    const uint64_t tbase = (uint64_t)blockIdx.x * kTile;
    for (int j = 0; j < kPerThread; ++j) {
        const uint64_t i = tbase + (uint64_t)j * kSortBlk + t;
        const bool has = i < n;
        const uint64_t key = has ? in[i] : 0ull;
        const uint32_t d = (uint32_t)((key >> shift) & (kDigits - 1));
        // lanes of this wavefront holding a key of the same digit
        unsigned long long peers = __ballot(has);
#pragma unroll
        for (int b = 0; b < kDigitBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << ln) - 1ull));
#pragma unroll
        for (int v = 0; v < kSortWaves; ++v) wc[v][t] = 0;
        __syncthreads();
        if (has && rank + 1 == (uint32_t)__popcll(peers)) wc[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            // thread t: digit t, prefix over the wavefronts
            uint32_t acc = run[t];
#pragma unroll
            for (int v = 0; v < kSortWaves; ++v) {
                const uint32_t c = wc[v][t];
                wc[v][t] = acc;
                acc += c;
            }
            run[t] = acc;
        }
        __syncthreads();
        if (has) out[(uint64_t)gbase[d] + wc[w][d] + rank] = key;
        __syncthreads();  // wc is cleared at the start of the next round
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
                                 uint32_t* table, uint64_t** result, hipStream_t st, int lo_bit) {
    *result = keys;
    if (n < 2) return hipSuccess;
    const uint32_t nt = tiles_for(n);
    uint64_t* a = keys;
    uint64_t* b = alt;
    const uint64_t m = (uint64_t)kDigits * nt;
    uint32_t* sums = table + m + 1;
    for (int shift = lo_bit; shift < bits; shift += kDigitBits) {
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
