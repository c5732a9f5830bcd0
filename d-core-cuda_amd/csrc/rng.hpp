// Counter-based random numbers for the seeded graph generators.
//
// splitmix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): the output of draw i of stream s depends only on
// (seed, s, i), so generation parallelises over samples with identical
// results for any thread count.
#pragma once
#include <cstdint>

namespace dcx {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

// Sequential splitmix64 stream.
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() { return mix64(s += kGolden); }
    // uniform in [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // uniform integer in [0, n) by multiply-high (bias < 2^-32 * n)
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

// Stream for sample `i` of a generator seeded with `seed`.
inline SplitMix stream_for(uint64_t seed, uint64_t i) {
    return SplitMix(mix64(seed * kGolden + 0x632be59bd9b4e019ULL) ^ mix64(i + 1));
}

}  // namespace dcx
