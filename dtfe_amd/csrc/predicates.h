// Exact orientation signs on the GPU for DTFE point location.
//
// The reference locates points with CGAL's exact-predicates kernel
// (src/CGAL_triangulation/CGAL_include_3D.h:52-55; locate at
// src/CGAL_triangulation/unaveraged_interpolation.cc:289). Here a sample point
// p is tested against the faces of a tetrahedron with an FP64 determinant
// and a forward error bound (Shewchuk 1997, "Adaptive Precision
// Floating-Point Arithmetic and Fast Robust Geometric Predicates", bound
// (7 + 56 eps) eps * permanent); when the bound cannot certify the sign, the
// determinant is evaluated exactly as a floating-point expansion (FMA
// two-product, two-sum, zero elimination). An exact zero is resolved by
// moving p symbolically by eps*(1, eps, eps^2): the sign then comes from the
// face normal's x, y, z components, so every point has exactly one
// containing tetrahedron.
#pragma once

#include <hip/hip_runtime.h>

namespace dtfe {

struct d3 {
    double x, y, z;
};

__device__ __forceinline__ void ex_two_sum(double a, double b, double &x, double &y) {
#pragma clang fp contract(off)
    x = a + b;
    double bv = x - a;
    double av = x - bv;
    y = (a - av) + (b - bv);
}

__device__ __forceinline__ void ex_two_diff(double a, double b, double &x, double &y) {
#pragma clang fp contract(off)
    x = a - b;
    double bv = a - x;
    double av = x + bv;
    y = (a - av) + (bv - b);
}

__device__ __forceinline__ void ex_two_prod(double a, double b, double &x, double &y) {
#pragma clang fp contract(off)
    x = a * b;
    y = __builtin_fma(a, b, -x);
}

// acc (length n, increasing magnitude, nonoverlapping) += b, in place.
__device__ __forceinline__ int ex_grow(double *acc, int n, double b) {
    double q = b;
    int k = 0;
    for (int i = 0; i < n; ++i) {
        double qn, h;
        ex_two_sum(q, acc[i], qn, h);
        q = qn;
        if (h != 0.0) acc[k++] = h;
    }
    if (q != 0.0 || k == 0) acc[k++] = q;
    return k;
}

// acc += s * x * y * z, exactly (x*y as two terms, each times z as two terms)
__device__ __forceinline__ int ex_add_triple(double *acc, int n, double s, double x, double y, double z) {
    double p1, p0;
    ex_two_prod(x, y, p1, p0);
    double a1, a0, b1, b0;
    ex_two_prod(p1, z, a1, a0);
    ex_two_prod(p0, z, b1, b0);
    n = ex_grow(acc, n, s * b0);
    n = ex_grow(acc, n, s * b1);
    n = ex_grow(acc, n, s * a0);
    n = ex_grow(acc, n, s * a1);
    return n;
}

__device__ __forceinline__ int ex_sign(const double *acc, int n) {
    double t = acc[n - 1];
    return (t > 0.0) - (t < 0.0);
}

// Exact sign of det[a-p; b-p; c-p] (= -orient(a, b, c, p)).
__device__ __noinline__ int exact_det3_sign(d3 a, d3 b, d3 c, d3 p) {
    // exact differences as two-term sums: u = a - p, v = b - p, w = c - p
    double u[3][2], v[3][2], w[3][2];
    ex_two_diff(a.x, p.x, u[0][0], u[0][1]);
    ex_two_diff(a.y, p.y, u[1][0], u[1][1]);
    ex_two_diff(a.z, p.z, u[2][0], u[2][1]);
    ex_two_diff(b.x, p.x, v[0][0], v[0][1]);
    ex_two_diff(b.y, p.y, v[1][0], v[1][1]);
    ex_two_diff(b.z, p.z, v[2][0], v[2][1]);
    ex_two_diff(c.x, p.x, w[0][0], w[0][1]);
    ex_two_diff(c.y, p.y, w[1][0], w[1][1]);
    ex_two_diff(c.z, p.z, w[2][0], w[2][1]);
    double acc[200];
    acc[0] = 0.0;
    int n = 1;
    // det = sum over permutations (i,j,k) of sgn * u_i v_j w_k
    const int perm[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
    for (int q = 0; q < 6; ++q) {
        double s = q < 3 ? 1.0 : -1.0;
        int i = perm[q][0], j = perm[q][1], k = perm[q][2];
        for (int ia = 0; ia < 2; ++ia)
            for (int jb = 0; jb < 2; ++jb)
                for (int kc = 0; kc < 2; ++kc) {
                    double x = u[i][ia], y = v[j][jb], z = w[k][kc];
                    if (x == 0.0 || y == 0.0 || z == 0.0) continue;
                    n = ex_add_triple(acc, n, s, x, y, z);
                }
    }
    return ex_sign(acc, n);
}

// Exact sign of (b-a)_i (c-a)_j - (b-a)_j (c-a)_i.
__device__ __noinline__ int exact_cross_sign(double ai, double aj, double bi, double bj, double ci,
                                             double cj) {
    double u[2][2], v[2][2];
    ex_two_diff(bi, ai, u[0][0], u[0][1]);
    ex_two_diff(bj, aj, u[1][0], u[1][1]);
    ex_two_diff(ci, ai, v[0][0], v[0][1]);
    ex_two_diff(cj, aj, v[1][0], v[1][1]);
    double acc[40];
    acc[0] = 0.0;
    int n = 1;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            double p1, p0;
            ex_two_prod(u[0][a], v[1][b], p1, p0);
            n = ex_grow(acc, n, p0);
            n = ex_grow(acc, n, p1);
            ex_two_prod(u[1][a], v[0][b], p1, p0);
            n = ex_grow(acc, n, -p0);
            n = ex_grow(acc, n, -p1);
        }
    return ex_sign(acc, n);
}

// Sign of orient(a, b, c, p) = det[b-a; c-a; p-a] for the symbolically
// perturbed p; 0 only if a, b, c are collinear.
__device__ __noinline__ int orient_sos_slow(d3 a, d3 b, d3 c, d3 p) {
    int s = -exact_det3_sign(a, b, c, p);
    if (s != 0) return s;
    // gradient of orient w.r.t. p is n = (b-a) x (c-a)
    s = exact_cross_sign(a.y, a.z, b.y, b.z, c.y, c.z);  // n_x
    if (s != 0) return s;
    s = exact_cross_sign(a.z, a.x, b.z, b.x, c.z, c.x);  // n_y
    if (s != 0) return s;
    return exact_cross_sign(a.x, a.y, b.x, b.y, c.x, c.y);  // n_z
}

}  // namespace dtfe
