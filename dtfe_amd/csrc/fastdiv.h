// Unsigned 32-bit division by a divisor fixed at launch time, as a multiply
// and two shifts (Granlund & Montgomery, "Division by invariant integers
// using multiplication", PLDI 1994, Fig. 4.1): for 1 <= d < 2^32 and
// l = ceil(log2 d),
//     m  = floor(2^32 (2^l - d) / d) + 1
//     t  = mulhi(m, n)
//     q  = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0)
// equals floor(n / d) for every 0 <= n < 2^32.
//
// The kernels split flat row indices into grid coordinates with it; a 64-bit
// division by a runtime divisor costs tens of vector instructions per lane.
// DTFE_HD is defined by the includer (__host__ __device__ in HIP sources,
// empty for host-only builds).
#pragma once

#include <cstdint>

#ifndef DTFE_HD
#define DTFE_HD
#endif

namespace dtfe {

struct FastDiv {
    uint32_t d = 1;
    uint32_t m = 1;
    uint32_t s1 = 0;
    uint32_t s2 = 0;

    FastDiv() = default;
    explicit FastDiv(uint32_t divisor) : d(divisor) {
        uint32_t l = 0;
        while (l < 32 && (uint64_t(1) << l) < divisor) ++l;
        m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - divisor)) / divisor + 1);
        s1 = l < 1 ? l : 1;
        s2 = l > 0 ? l - 1 : 0;
    }

    DTFE_HD uint32_t div(uint32_t n) const {
        uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
        return (t + ((n - t) >> s1)) >> s2;
    }
};

}  // namespace dtfe
