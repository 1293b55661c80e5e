// The Monte-Carlo sample stream of the reference's cell averaging:
// boost::mt19937 seeded with randomSeed (+1000*partNo) and mapped through
// boost::uniform_real(-0.5, 0.5) (src/CGAL_triangulation/averaged_interpolation_2.cc:96-102).
//
// The MT19937 state cannot be jumped cheaply, so the host runs the bare
// twist recurrence once per (seed, stream length, launch partition) and keeps
// the 624-word state at the first MT block each workgroup needs ("plan").
// A workgroup then regenerates its own words in LDS.
#pragma once

#include <cstdint>
#include <vector>

namespace dtfe {

constexpr int MT_N = 624;
constexpr int MT_M = 397;
constexpr uint32_t MT_UPPER = 0x80000000u;
constexpr uint32_t MT_LOWER = 0x7fffffffu;
constexpr uint32_t MT_MATRIX = 0x9908b0dfu;

inline void mt_seed(uint32_t seed, uint32_t *st) {
    st[0] = seed;
    for (int i = 1; i < MT_N; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
}

// One twist in place: the 624 untempered words of the next MT block.
inline void mt_twist(uint32_t *st) {
    for (int i = 0; i < MT_N; ++i) {
        uint32_t y = (st[i] & MT_UPPER) | (st[(i + 1) % MT_N] & MT_LOWER);
        uint32_t v = st[(i + MT_M) % MT_N] ^ (y >> 1);
        st[i] = (y & 1u) ? (v ^ MT_MATRIX) : v;
    }
}

// States needed by a launch whose workgroup b starts at MT block first_block[b]
// (non-decreasing). plan[b] = the untempered words of block first_block[b]-1,
// or the seeded array when first_block[b] == 0.
inline void mt_build_plan(uint32_t seed, const std::vector<int64_t> &first_block,
                          std::vector<uint32_t> &plan) {
    plan.resize(first_block.size() * (size_t)MT_N);
    uint32_t st[MT_N];
    mt_seed(seed, st);
    int64_t current = -1;  // st holds the words of block `current` (-1 = seeded array)
    for (size_t b = 0; b < first_block.size(); ++b) {
        int64_t need = first_block[b] - 1;
        while (current < need) {
            mt_twist(st);
            ++current;
        }
        for (int i = 0; i < MT_N; ++i) plan[b * MT_N + i] = st[i];
    }
}

__host__ __device__ inline uint32_t mt_temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// boost uniform_real(-0.5, 0.5) on a 32-bit engine: x / 2^32 * 1 + (-0.5),
// exact in FP64.
__host__ __device__ inline double mt_uniform(uint32_t tempered) {
    return (double)tempered * (1.0 / 4294967296.0) - 0.5;
}

}  // namespace dtfe
