// splitmix.hpp -- SplitMix64 (Steele, Lea & Flood 2014) keyed by (seed, stream, cell index):
// the same stream gaukuk_amd.problems.uniform_pm1 draws, so setups built here and the Python
// generators produce the same perturbations.
#pragma once
#include <cstdint>

namespace Gaukuk {

inline uint64_t SplitMix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [-1, 1)
inline double UniformPm1(uint64_t seed, uint64_t idx, uint64_t stream = 0) {
    const uint64_t key = (seed << 32) + (stream << 28) + idx;
    const double u = (double)(SplitMix64(key) >> 11) * (1.0 / 9007199254740992.0);
    return 2.0 * u - 1.0;
}

}  // namespace Gaukuk
