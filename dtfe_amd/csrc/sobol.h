// The quasi-random points of averaging method 1: GSL's 3-dimensional Sobol
// generator (gsl_qrng_sobol; GSL's qrng/sobol.c, after Bratley & Fox, ACM
// TOMS 14 (1988), Algorithm 659: 30-bit numerators, dimension 0 the van der
// Corput sequence, dimension 1 the primitive polynomial x + 1 with initial
// value 1, dimension 2 x^2 + x + 1 with initial values 1, 1; the all-zero
// point is not returned) and the reference's filter to the unit simplex
// (quasiRandomSequence, src/CGAL_triangulation/averaged_interpolation_1.cc:31-48).
// Host code.
#pragma once

#include <cstdint>
#include <vector>

namespace dtfe {

class Sobol3 {
  public:
    Sobol3() {
        const int poly[3] = {1, 3, 7}, degree[3] = {0, 1, 2};
        const uint32_t first[2][3] = {{0, 1, 1}, {0, 0, 1}};  // leading direction values per dimension
        for (int k = 0; k < kBits; ++k) v_[k][0] = 1;
        for (int d = 1; d < 3; ++d) {
            const int deg = degree[d];
            bool taps[8];
            int bits = poly[d];
            for (int k = deg - 1; k >= 0; --k) {
                taps[k] = bits & 1;
                bits >>= 1;
            }
            for (int j = 0; j < deg; ++j) v_[j][d] = first[j][d];
            for (int j = deg; j < kBits; ++j) {
                uint32_t x = v_[j - deg][d];
                uint32_t scale = 1;
                for (int k = 0; k < deg; ++k) {
                    scale <<= 1;
                    if (taps[k]) x ^= scale * v_[j - k - 1][d];
                }
                v_[j][d] = x;
            }
        }
        uint32_t scale = 1;
        for (int j = kBits - 2; j >= 0; --j) {
            scale <<= 1;
            for (int d = 0; d < 3; ++d) v_[j][d] *= scale;
        }
        inv_ = 1.0 / (2.0 * scale);
    }

    // The next point of the sequence.
    void next(double out[3]) {
        int bit = 0;
        for (uint64_t c = count_; c & 1; c >>= 1) ++bit;  // lowest zero bit of the count
        for (int d = 0; d < 3; ++d) {
            num_[d] ^= v_[bit][d];
            out[d] = num_[d] * inv_;
        }
        ++count_;
    }

  private:
    static constexpr int kBits = 30;
    uint32_t v_[kBits][3];
    uint32_t num_[3] = {0, 0, 0};
    uint64_t count_ = 0;
    double inv_ = 0.0;
};

// The first n points (x, y, z) of the sequence with x + y + z <= 1.
inline std::vector<double> sobol_simplex(int64_t n) {
    std::vector<double> out(3 * (size_t)n);
    Sobol3 q;
    for (int64_t i = 0; i < n;) {
        double r[3];
        q.next(r);
        if (r[0] + r[1] + r[2] <= 1.0) {
            for (int j = 0; j < 3; ++j) out[3 * i + j] = r[j];
            ++i;
        }
    }
    return out;
}

}  // namespace dtfe
