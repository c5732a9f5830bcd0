// Counter-based random numbers for the seeded graph generators.
//
// splitmix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): the output of draw i of stream s depends only on
// (seed, s, i), so generation parallelises over samples with identical
// results for any thread count.
#pragma once
#include <cstdint>

// The same functions serve the host generators and the device sampler of
// dcx_gen_rmat_device (compiled by hipcc), so both draw identical samples.
#if defined(__HIPCC__)
#define DCX_HD __host__ __device__
#else
#define DCX_HD
#endif

namespace dcx {

DCX_HD inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

// Sequential splitmix64 stream.
struct SplitMix {
    uint64_t s;
    DCX_HD explicit SplitMix(uint64_t seed) : s(seed) {}
    DCX_HD uint64_t next() { return mix64(s += kGolden); }
    // uniform in [0, 1) with 53 random bits
    DCX_HD double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // uniform integer in [0, n) by multiply-high (bias < 2^-32 * n)
    DCX_HD uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

// Stream for sample `i` of a generator seeded with `seed`.
DCX_HD inline SplitMix stream_for(uint64_t seed, uint64_t i) {
    return SplitMix(mix64(seed * kGolden + 0x632be59bd9b4e019ULL) ^ mix64(i + 1));
}

}  // namespace dcx
