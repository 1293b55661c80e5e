// TEST INFRASTRUCTURE — part of the CPU oracle. Only tests/, the smoke check
// and the CPU-baseline leg of the benchmark may use anything under oracle/.
//
// Exact geometric signs for the oracle's point location.
//
// The reference locates sample points with CGAL's exact-predicates kernel
// (src/CGAL_triangulation/CGAL_include_3D.h:52-55, locate at
// src/CGAL_triangulation/unaveraged_interpolation.cc:289). CGAL is absent
// here, so the oracle restates an exact orientation sign from the published
// method of floating-point expansions (J. R. Shewchuk, "Adaptive Precision
// Floating-Point Arithmetic and Fast Robust Geometric Predicates", 1997):
// Dekker's split/two-product and Knuth's two-sum, with zero elimination.
// Ties (sample exactly on a face plane) are broken by a symbolic
// perturbation of the *sample point* by eps*(1, eps, eps^2), which gives
// every point a unique containing tetrahedron.
#ifndef DTFE_ORACLE_EXACT_H
#define DTFE_ORACLE_EXACT_H

#include <cmath>

namespace oracle_exact {

static inline void two_sum(double a, double b, double &x, double &y) {
    x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    y = ar + br;
}

static inline void fast_two_sum(double a, double b, double &x, double &y) {
    x = a + b;
    double bv = x - a;
    y = b - bv;
}

static inline void two_diff(double a, double b, double &x, double &y) {
    x = a - b;
    double bv = a - x;
    double av = x + bv;
    double br = bv - b;
    double ar = a - av;
    y = ar + br;
}

// Dekker split: a = hi + lo with hi, lo holding at most 26 significant bits.
static inline void split(double a, double &hi, double &lo) {
    double c = 134217729.0 * a;  // 2^27 + 1
    double abig = c - a;
    hi = c - abig;
    lo = a - hi;
}

static inline void two_product(double a, double b, double &x, double &y) {
    x = a * b;
    double ahi, alo, bhi, blo;
    split(a, ahi, alo);
    split(b, bhi, blo);
    double err1 = x - ahi * bhi;
    double err2 = err1 - alo * bhi;
    double err3 = err2 - ahi * blo;
    y = alo * blo - err3;
}

// h = e + b ; e nonoverlapping, increasing magnitude. Returns length of h.
static inline int grow_expansion(int elen, const double *e, double b, double *h) {
    double q = b;
    int k = 0;
    for (int i = 0; i < elen; ++i) {
        double qn, hh;
        two_sum(q, e[i], qn, hh);
        q = qn;
        if (hh != 0.0) h[k++] = hh;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

// h = e + f (repeated growth; O(len(e)*len(f)), fine for the oracle).
static inline int expansion_sum(int elen, const double *e, int flen, const double *f,
                                double *h, double *tmp) {
    int hlen = elen;
    for (int i = 0; i < elen; ++i) h[i] = e[i];
    for (int j = 0; j < flen; ++j) {
        int n = grow_expansion(hlen, h, f[j], tmp);
        for (int i = 0; i < n; ++i) h[i] = tmp[i];
        hlen = n;
    }
    return hlen;
}

// h = b * e.
static inline int scale_expansion(int elen, const double *e, double b, double *h) {
    int k = 0;
    double q, hh;
    two_product(e[0], b, q, hh);
    if (hh != 0.0) h[k++] = hh;
    for (int i = 1; i < elen; ++i) {
        double p1, p0, sum;
        two_product(e[i], b, p1, p0);
        two_sum(q, p0, sum, hh);
        if (hh != 0.0) h[k++] = hh;
        fast_two_sum(p1, sum, q, hh);
        if (hh != 0.0) h[k++] = hh;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

// h = e * f
static inline int expansion_product(int elen, const double *e, int flen, const double *f,
                                    double *h) {
    double part[512], acc[1024], tmp[1024];
    int alen = 1;
    acc[0] = 0.0;
    for (int j = 0; j < flen; ++j) {
        int plen = scale_expansion(elen, e, f[j], part);
        alen = expansion_sum(alen, acc, plen, part, tmp, h);
        for (int i = 0; i < alen; ++i) acc[i] = tmp[i];
    }
    for (int i = 0; i < alen; ++i) h[i] = acc[i];
    return alen;
}

static inline int sign_of(int n, const double *e) {
    double top = e[n - 1];
    return (top > 0.0) - (top < 0.0);
}

static inline int exact_diff(double a, double b, double *e) {
    double x, y;
    two_diff(a, b, x, y);
    int k = 0;
    if (y != 0.0) e[k++] = y;
    e[k++] = x;
    return k;
}

// exact sign of (u1*v2 - u2*v1) with u_i, v_i given as expansions
static inline int cross_term(int lu1, const double *u1, int lv2, const double *v2,
                             int lu2, const double *u2, int lv1, const double *v1,
                             double *out) {
    double a[64], b[64], tmp[128];
    int la = expansion_product(lu1, u1, lv2, v2, a);
    int lb = expansion_product(lu2, u2, lv1, v1, b);
    for (int i = 0; i < lb; ++i) b[i] = -b[i];
    return expansion_sum(la, a, lb, b, out, tmp);
}

// Exact sign of det[b-a; c-a; d-a].
static inline int orient3d_exact(const double *a, const double *b, const double *c,
                                 const double *d) {
    double u[3][2], v[3][2], w[3][2];
    int lu[3], lv[3], lw[3];
    for (int i = 0; i < 3; ++i) {
        lu[i] = exact_diff(b[i], a[i], u[i]);
        lv[i] = exact_diff(c[i], a[i], v[i]);
        lw[i] = exact_diff(d[i], a[i], w[i]);
    }
    // det = ux (vy wz - vz wy) + uy (vz wx - vx wz) + uz (vx wy - vy wx)
    double m[3][64], t[3][256], s1[512], s2[768], tmp[768];
    int lm0 = cross_term(lv[1], v[1], lw[2], w[2], lv[2], v[2], lw[1], w[1], m[0]);
    int lm1 = cross_term(lv[2], v[2], lw[0], w[0], lv[0], v[0], lw[2], w[2], m[1]);
    int lm2 = cross_term(lv[0], v[0], lw[1], w[1], lv[1], v[1], lw[0], w[0], m[2]);
    int lt0 = expansion_product(lm0, m[0], lu[0], u[0], t[0]);
    int lt1 = expansion_product(lm1, m[1], lu[1], u[1], t[1]);
    int lt2 = expansion_product(lm2, m[2], lu[2], u[2], t[2]);
    int l1 = expansion_sum(lt0, t[0], lt1, t[1], s1, tmp);
    int l2 = expansion_sum(l1, s1, lt2, t[2], s2, tmp);
    return sign_of(l2, s2);
}

// Exact signs of the components of n = (b-a) x (c-a).
static inline void normal_signs_exact(const double *a, const double *b, const double *c,
                                      int *sx, int *sy, int *sz) {
    double u[3][2], v[3][2], out[128];
    int lu[3], lv[3];
    for (int i = 0; i < 3; ++i) {
        lu[i] = exact_diff(b[i], a[i], u[i]);
        lv[i] = exact_diff(c[i], a[i], v[i]);
    }
    int n;
    n = cross_term(lu[1], u[1], lv[2], v[2], lu[2], u[2], lv[1], v[1], out);
    *sx = sign_of(n, out);
    n = cross_term(lu[2], u[2], lv[0], v[0], lu[0], u[0], lv[2], v[2], out);
    *sy = sign_of(n, out);
    n = cross_term(lu[0], u[0], lv[1], v[1], lu[1], u[1], lv[0], v[0], out);
    *sz = sign_of(n, out);
}

// FP64 value of det[b-a; c-a; d-a] plus a certified sign when the
// forward error bound allows it (returns 0 when uncertain).
static inline double orient3d_fast(const double *a, const double *b, const double *c,
                                   const double *d, int *certain_sign) {
    double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    double vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    double wx = d[0] - a[0], wy = d[1] - a[1], wz = d[2] - a[2];
    double m0 = vy * wz - vz * wy;
    double m1 = vz * wx - vx * wz;
    double m2 = vx * wy - vy * wx;
    double det = ux * m0 + uy * m1 + uz * m2;
    double perm = std::fabs(ux) * (std::fabs(vy * wz) + std::fabs(vz * wy)) +
                  std::fabs(uy) * (std::fabs(vz * wx) + std::fabs(vx * wz)) +
                  std::fabs(uz) * (std::fabs(vx * wy) + std::fabs(vy * wx));
    // Shewchuk's first-stage bound for this determinant shape, (7+56eps)eps,
    // rounded up to 8.9e-16 to cover the extra rounding of the differences.
    double bound = 8.9e-16 * perm;
    if (det > bound) *certain_sign = 1;
    else if (-det > bound) *certain_sign = -1;
    else *certain_sign = 0;
    return det;
}

// Sign of orient(a,b,c,p) with p perturbed by eps*(1,eps,eps^2); never 0
// unless (a,b,c) is degenerate (collinear).
static inline int orient_sos(const double *a, const double *b, const double *c,
                             const double *p, double *value) {
    int s;
    *value = orient3d_fast(a, b, c, p, &s);
    if (s != 0) return s;
    s = orient3d_exact(a, b, c, p);
    if (s != 0) return s;
    int sx, sy, sz;
    normal_signs_exact(a, b, c, &sx, &sy, &sz);
    if (sx != 0) return sx;
    if (sy != 0) return sy;
    return sz;
}

}  // namespace oracle_exact

#endif
