// Fixed-size vector with the interface of the reference's Pvector
// (src/Pvector.h:45-152): contiguous T x[N], element access, arithmetic.
#ifndef DTFE_DROPIN_PVECTOR_H
#define DTFE_DROPIN_PVECTOR_H

#include <cstddef>

template <typename T, size_t N>
struct Pvector {
    T x[N];

    Pvector() {
        for (size_t i = 0; i < N; ++i) x[i] = T(0);
    }
    Pvector(T a, T b, T c) {
        static_assert(N == 3, "three-component constructor");
        x[0] = a;
        x[1] = b;
        x[2] = c;
    }
    explicit Pvector(const T *v) {
        for (size_t i = 0; i < N; ++i) x[i] = v[i];
    }
    T &operator[](size_t i) { return x[i]; }
    const T &operator[](size_t i) const { return x[i]; }
    Pvector operator+(const Pvector &o) const {
        Pvector r;
        for (size_t i = 0; i < N; ++i) r.x[i] = x[i] + o.x[i];
        return r;
    }
    Pvector &operator+=(const Pvector &o) {
        for (size_t i = 0; i < N; ++i) x[i] += o.x[i];
        return *this;
    }
    Pvector operator-(const Pvector &o) const {
        Pvector r;
        for (size_t i = 0; i < N; ++i) r.x[i] = x[i] - o.x[i];
        return r;
    }
    Pvector &operator-=(const Pvector &o) {
        for (size_t i = 0; i < N; ++i) x[i] -= o.x[i];
        return *this;
    }
    Pvector operator*(T s) const {
        Pvector r;
        for (size_t i = 0; i < N; ++i) r.x[i] = x[i] * s;
        return r;
    }
    Pvector &operator*=(T s) {
        for (size_t i = 0; i < N; ++i) x[i] *= s;
        return *this;
    }
    Pvector operator/(T s) const {
        Pvector r;
        for (size_t i = 0; i < N; ++i) r.x[i] = x[i] / s;
        return r;
    }
    Pvector &operator/=(T s) {
        for (size_t i = 0; i < N; ++i) x[i] /= s;
        return *this;
    }
    bool operator==(const Pvector &o) const {
        for (size_t i = 0; i < N; ++i)
            if (x[i] != o.x[i]) return false;
        return true;
    }
    bool operator!=(const Pvector &o) const { return !(*this == o); }
    static Pvector zero() { return Pvector(); }
};

#endif
