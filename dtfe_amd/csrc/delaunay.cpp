// Exact 3D Delaunay tessellation on the host: the stand-in for the
// reference's CGAL Delaunay_triangulation_3 (src/CGAL_triangulation/CGAL_include_3D.h:52-55,
// built in delaunayTriangulation, src/CGAL_triangulation/triangulation.cpp:162-195),
// which is not available in this image.
//
// Incremental Bowyer-Watson insertion with an infinite vertex (hull faces
// belong to "infinite" tetrahedra), points inserted in a biased randomized
// insertion order (rounds of doubling size, each round sorted along a Morton
// curve) and located by a visibility walk from the last new tetrahedron.
// Signs are exact: FP64 with a forward error bound, else floating-point
// expansions (Knuth two-sum, Dekker two-product). Degenerate (cospherical)
// inputs are resolved by symbolically perturbing the lifted coordinate
// |x|^2 of point i by eps^(n-i): the result is the Delaunay tessellation of
// the perturbed points, with no flat tetrahedra. A point on the plane of a
// hull face conflicts with the infinite tetrahedron iff it conflicts with
// the finite one behind that face.
//
// Output: finite tetrahedra (positively oriented: det[x1-x0; x2-x0; x3-x0] > 0)
// with, per face opposite vertex i, the neighbouring finite tetrahedron or -1.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// floating-point expansions (exact sums and products of doubles)
// ---------------------------------------------------------------------------
struct Ex {
    std::vector<double> t;  // nonoverlapping, increasing magnitude, no zeros (empty = 0)
};

inline void two_sum(double a, double b, double &x, double &y) {
    x = a + b;
    double bv = x - a, av = x - bv;
    y = (a - av) + (b - bv);
}

inline void split(double a, double &hi, double &lo) {
    double c = 134217729.0 * a;
    double ab = c - a;
    hi = c - ab;
    lo = a - hi;
}

inline void two_prod(double a, double b, double &x, double &y) {
    x = a * b;
    double ah, al, bh, bl;
    split(a, ah, al);
    split(b, bh, bl);
    y = al * bl - (((x - ah * bh) - al * bh) - ah * bl);
}

Ex ex_diff(double a, double b) {
    double x = a - b;
    double bv = a - x, av = x + bv;
    double y = (a - av) + (bv - b);
    Ex e;
    if (y != 0.0) e.t.push_back(y);
    if (x != 0.0) e.t.push_back(x);
    return e;
}

void grow(std::vector<double> &e, double b) {
    double q = b;
    size_t k = 0;
    for (size_t i = 0; i < e.size(); ++i) {
        double qn, h;
        two_sum(q, e[i], qn, h);
        q = qn;
        if (h != 0.0) e[k++] = h;
    }
    e.resize(k);
    if (q != 0.0) e.push_back(q);
}

Ex ex_add(const Ex &a, const Ex &b) {
    Ex r = a.t.size() >= b.t.size() ? a : b;
    const Ex &s = a.t.size() >= b.t.size() ? b : a;
    for (double x : s.t) grow(r.t, x);
    return r;
}

Ex ex_neg(Ex a) {
    for (double &x : a.t) x = -x;
    return a;
}

Ex ex_sub(const Ex &a, const Ex &b) { return ex_add(a, ex_neg(b)); }

Ex ex_scale(const Ex &e, double b) {
    Ex r;
    for (double x : e.t) {
        double p, q;
        two_prod(x, b, p, q);
        grow(r.t, q);
        grow(r.t, p);
    }
    return r;
}

Ex ex_mul(const Ex &a, const Ex &b) {
    Ex r;
    for (double y : b.t) r = ex_add(r, ex_scale(a, y));
    return r;
}

int ex_sign(const Ex &e) {
    if (e.t.empty()) return 0;
    double v = e.t.back();
    return (v > 0) - (v < 0);
}

Ex ex_det3(const Ex m[3][3]) {
    Ex a = ex_sub(ex_mul(m[1][1], m[2][2]), ex_mul(m[1][2], m[2][1]));
    Ex b = ex_sub(ex_mul(m[1][2], m[2][0]), ex_mul(m[1][0], m[2][2]));
    Ex c = ex_sub(ex_mul(m[1][0], m[2][1]), ex_mul(m[1][1], m[2][0]));
    return ex_add(ex_add(ex_mul(m[0][0], a), ex_mul(m[0][1], b)), ex_mul(m[0][2], c));
}

// ---------------------------------------------------------------------------
// predicates
// ---------------------------------------------------------------------------
// sign of det[b-a; c-a; d-a]
int orient_sign(const double *a, const double *b, const double *c, const double *d) {
    double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    double vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    double wx = d[0] - a[0], wy = d[1] - a[1], wz = d[2] - a[2];
    double det = ux * (vy * wz - vz * wy) + uy * (vz * wx - vx * wz) + uz * (vx * wy - vy * wx);
    double perm = std::fabs(ux) * (std::fabs(vy * wz) + std::fabs(vz * wy)) +
                  std::fabs(uy) * (std::fabs(vz * wx) + std::fabs(vx * wz)) +
                  std::fabs(uz) * (std::fabs(vx * wy) + std::fabs(vy * wx));
    double bound = 8.0e-16 * perm;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    Ex m[3][3];
    for (int i = 0; i < 3; ++i) {
        m[0][i] = ex_diff(b[i], a[i]);
        m[1][i] = ex_diff(c[i], a[i]);
        m[2][i] = ex_diff(d[i], a[i]);
    }
    return ex_sign(ex_det3(m));
}

// D = det [[q - p, |q - p|^2]] over rows q = a, b, c, d; for a positively
// oriented (a, b, c, d), p is inside the circumsphere when D < 0.
int lifted_sign(const double *a, const double *b, const double *c, const double *d, const double *p) {
    const double *q[4] = {a, b, c, d};
    double r[4][3], w[4];
    for (int k = 0; k < 4; ++k) {
        for (int i = 0; i < 3; ++i) r[k][i] = q[k][i] - p[i];
        w[k] = r[k][0] * r[k][0] + r[k][1] * r[k][1] + r[k][2] * r[k][2];
    }
    double D = 0.0, P = 0.0;
    for (int k = 0; k < 4; ++k) {
        int o[3], n = 0;
        for (int j = 0; j < 4; ++j)
            if (j != k) o[n++] = j;
        const double *x = r[o[0]], *y = r[o[1]], *z = r[o[2]];
        double m = x[0] * (y[1] * z[2] - y[2] * z[1]) + x[1] * (y[2] * z[0] - y[0] * z[2]) +
                   x[2] * (y[0] * z[1] - y[1] * z[0]);
        double pm = std::fabs(x[0]) * (std::fabs(y[1] * z[2]) + std::fabs(y[2] * z[1])) +
                    std::fabs(x[1]) * (std::fabs(y[2] * z[0]) + std::fabs(y[0] * z[2])) +
                    std::fabs(x[2]) * (std::fabs(y[0] * z[1]) + std::fabs(y[1] * z[0]));
        double sgn = ((k + 3) % 2 == 0) ? 1.0 : -1.0;
        D += sgn * w[k] * m;
        P += w[k] * pm;
    }
    double bound = 1.0e-14 * P;
    if (D > bound) return 1;
    if (-D > bound) return -1;
    // exact
    Ex R[4][3], W[4];
    for (int k = 0; k < 4; ++k) {
        for (int i = 0; i < 3; ++i) R[k][i] = ex_diff(q[k][i], p[i]);
        W[k] = ex_add(ex_add(ex_mul(R[k][0], R[k][0]), ex_mul(R[k][1], R[k][1])), ex_mul(R[k][2], R[k][2]));
    }
    Ex acc;
    for (int k = 0; k < 4; ++k) {
        Ex m[3][3];
        int n = 0;
        for (int j = 0; j < 4; ++j)
            if (j != k) {
                for (int i = 0; i < 3; ++i) m[n][i] = R[j][i];
                ++n;
            }
        Ex term = ex_mul(W[k], ex_det3(m));
        acc = ((k + 3) % 2 == 0) ? ex_add(acc, term) : ex_sub(acc, term);
    }
    return ex_sign(acc);
}

constexpr int INF = -1;

struct Tet {
    int v[4];
    int n[4];
};

struct Delaunay {
    const double *pts;
    std::vector<Tet> tets;
    std::vector<int> free_list;
    std::vector<uint32_t> stamp;
    uint32_t cur_stamp = 0;
    std::vector<uint8_t> in_conflict;
    int last = 0;
    std::vector<int32_t> dup_of;  // duplicate point -> kept point (or -1)

    const double *P(int v) const { return pts + 3 * (int64_t)v; }

    int new_tet() {
        if (!free_list.empty()) {
            int t = free_list.back();
            free_list.pop_back();
            return t;
        }
        tets.push_back(Tet{});
        stamp.push_back(0);
        in_conflict.push_back(0);
        return (int)tets.size() - 1;
    }

    bool is_inf(const Tet &t) const { return t.v[0] == INF || t.v[1] == INF || t.v[2] == INF || t.v[3] == INF; }

    // orient of tet t with vertex slot i replaced by point p (positive: p on
    // the same side of face i as vertex i)
    int orient_replace(const Tet &t, int i, const double *p) const {
        const double *q[4];
        for (int k = 0; k < 4; ++k) q[k] = (k == i) ? p : P(t.v[k]);
        return orient_sign(q[0], q[1], q[2], q[3]);
    }

    // perturbed in-sphere for a finite tet (point ids for the perturbation order)
    bool insphere_sos(const Tet &t, int pid) const {
        const double *p = P(pid);
        int s = lifted_sign(P(t.v[0]), P(t.v[1]), P(t.v[2]), P(t.v[3]), p);
        if (s != 0) return s < 0;
        // D' = D + sum_k eps_k C_k, C_k = (-1)^k orient(other rows in order),
        // rows (v0, v1, v2, v3, p); eps larger for higher point id
        int ids[5] = {t.v[0], t.v[1], t.v[2], t.v[3], pid};
        int order[5] = {0, 1, 2, 3, 4};
        std::sort(order, order + 5, [&](int x, int y) { return ids[x] > ids[y]; });
        for (int oi = 0; oi < 5; ++oi) {
            int k = order[oi];
            const double *q[4];
            int n = 0;
            for (int j = 0; j < 5; ++j)
                if (j != k) q[n++] = P(ids[j]);
            int o = orient_sign(q[0], q[1], q[2], q[3]);
            if (o != 0) {
                int c = (k % 2 == 0) ? o : -o;
                return c < 0;
            }
        }
        return false;  // unreachable for a non-flat tetrahedron
    }

    bool conflict(int ti, int pid) const {
        const Tet &t = tets[ti];
        int k = -1;
        for (int i = 0; i < 4; ++i)
            if (t.v[i] == INF) k = i;
        if (k < 0) return insphere_sos(t, pid);
        int o = orient_replace(t, k, P(pid));
        if (o != 0) return o > 0;
        return insphere_sos(tets[t.n[k]], pid);
    }

    // visibility walk to a tetrahedron in conflict with point pid; returns
    // -2 - v when pid duplicates vertex v
    int locate(int pid) {
        const double *p = P(pid);
        int t = last;
        if (is_inf(tets[t])) {
            // start from the finite neighbour of an infinite tet
            int k = 0;
            while (tets[t].v[k] != INF) ++k;
            t = tets[t].n[k];
        }
        int prev = -1;
        uint32_t rng = 2463534242u ^ (uint32_t)pid;
        for (int64_t step = 0; step < 4 * (int64_t)tets.size() + 64; ++step) {
            const Tet &T = tets[t];
            if (is_inf(T)) return t;
            rng ^= rng << 13;
            rng ^= rng >> 17;
            rng ^= rng << 5;
            int off = rng & 3;
            int next = -1;
            for (int j = 0; j < 4; ++j) {
                int i = (j + off) & 3;
                if (T.n[i] == prev) continue;
                if (orient_replace(T, i, p) < 0) {
                    next = T.n[i];
                    break;
                }
            }
            if (next < 0) {
                for (int i = 0; i < 4; ++i) {
                    const double *q = P(T.v[i]);
                    if (q[0] == p[0] && q[1] == p[1] && q[2] == p[2]) return -2 - T.v[i];
                }
                return t;
            }
            prev = t;
            t = next;
        }
        return -1;
    }

    struct FaceRef {
        int a, b, c;  // sorted vertex ids of a face containing the new point
        int tet, slot;
    };

    // links the faces of the new tetrahedra that contain the new vertex
    static void link_faces(std::vector<Tet> &tets, std::vector<FaceRef> &faces) {
        std::sort(faces.begin(), faces.end(), [](const FaceRef &x, const FaceRef &y) {
            if (x.a != y.a) return x.a < y.a;
            if (x.b != y.b) return x.b < y.b;
            return x.c < y.c;
        });
        for (size_t i = 0; i + 1 < faces.size(); i += 2) {
            const FaceRef &x = faces[i], &y = faces[i + 1];
            tets[x.tet].n[x.slot] = y.tet;
            tets[y.tet].n[y.slot] = x.tet;
        }
    }

    static FaceRef face_of(const Tet &t, int ti, int slot) {
        int f[3], n = 0;
        for (int i = 0; i < 4; ++i)
            if (i != slot) f[n++] = t.v[i];
        std::sort(f, f + 3);
        return FaceRef{f[0], f[1], f[2], ti, slot};
    }

    bool init(int i0, int i1, int i2, int i3) {
        if (orient_sign(P(i0), P(i1), P(i2), P(i3)) < 0) std::swap(i2, i3);
        int t0 = new_tet();
        tets[t0] = Tet{{i0, i1, i2, i3}, {-1, -1, -1, -1}};
        std::vector<FaceRef> faces;
        for (int i = 0; i < 4; ++i) {
            int ti = new_tet();
            Tet t = tets[t0];
            t.v[i] = INF;
            // replacing vertex i by a point beyond face i flips the orientation:
            // swap two finite vertices to keep the positive convention
            int a = (i + 1) & 3, b = (i + 2) & 3;
            std::swap(t.v[a], t.v[b]);
            for (int k = 0; k < 4; ++k) t.n[k] = -1;
            t.n[i] = t0;
            tets[ti] = t;
            tets[t0].n[i] = ti;
            for (int k = 0; k < 4; ++k)
                if (k != i) faces.push_back(face_of(tets[ti], ti, k));
        }
        link_faces(tets, faces);
        last = t0;
        return true;
    }

    std::vector<int> stack, cavity, made;
    std::vector<FaceRef> faces;
    std::vector<Tet> newtets;

    // inserts point pid; returns false for a duplicate (dup_of set)
    bool insert(int pid) {
        int t0 = locate(pid);
        if (t0 <= -2) {
            dup_of[pid] = -2 - t0;
            return false;
        }
        if (t0 < 0) return false;
        ++cur_stamp;
        stack.clear();
        cavity.clear();
        faces.clear();
        stack.push_back(t0);
        stamp[t0] = cur_stamp;
        in_conflict[t0] = 1;
        struct Bnd {
            int tet, slot, outer;
        };
        static thread_local std::vector<Bnd> bnd;
        bnd.clear();
        while (!stack.empty()) {
            int t = stack.back();
            stack.pop_back();
            cavity.push_back(t);
            for (int i = 0; i < 4; ++i) {
                int nb = tets[t].n[i];
                if (stamp[nb] != cur_stamp) {
                    stamp[nb] = cur_stamp;
                    in_conflict[nb] = conflict(nb, pid) ? 1 : 0;
                    if (in_conflict[nb]) stack.push_back(nb);
                }
                if (!in_conflict[nb]) bnd.push_back(Bnd{t, i, nb});
            }
        }
        // new tetrahedra on the cavity boundary: face of the conflict tet
        // with the new point in place of the opposite vertex
        newtets.clear();
        for (const Bnd &b : bnd) {
            Tet nt = tets[b.tet];
            nt.v[b.slot] = pid;
            for (int k = 0; k < 4; ++k) nt.n[k] = -1;
            nt.n[b.slot] = b.outer;
            newtets.push_back(nt);
        }
        made.resize(bnd.size());
        for (int t : cavity) {
            in_conflict[t] = 0;
            free_list.push_back(t);
        }
        for (size_t q = 0; q < bnd.size(); ++q) {
            int ti = new_tet();
            made[q] = ti;
            tets[ti] = newtets[q];
            stamp[ti] = 0;
            in_conflict[ti] = 0;
            Tet &outer = tets[bnd[q].outer];
            for (int k = 0; k < 4; ++k)
                if (outer.n[k] == bnd[q].tet) {
                    // the outer tet may border the cavity on several faces and a
                    // freed index may already be reused: match the face by its vertices
                    bool same = true;
                    for (int j = 0; j < 4 && same; ++j) {
                        if (j == k) continue;
                        int v = outer.v[j];
                        bool found = false;
                        for (int m = 0; m < 4; ++m)
                            if (m != bnd[q].slot && newtets[q].v[m] == v) found = true;
                        same = found;
                    }
                    if (same) outer.n[k] = ti;
                }
            for (int k = 0; k < 4; ++k)
                if (k != bnd[q].slot) faces.push_back(face_of(tets[ti], ti, k));
        }
        link_faces(tets, faces);
        last = made.empty() ? last : made[0];
        return true;
    }
};

// Morton code of a point in [0,1)^3 with 21 bits per axis
inline uint64_t spread3(uint64_t x) {
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

std::vector<int> brio_order(int64_t n, const double *pts) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], pts[3 * i + k]);
            hi[k] = std::max(hi[k], pts[3 * i + k]);
        }
    double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    std::vector<uint64_t> code(n);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t c[3];
        for (int k = 0; k < 3; ++k) {
            double f = (pts[3 * i + k] - lo[k]) / ext;
            c[k] = (uint64_t)std::min(2097151.0, std::max(0.0, f * 2097152.0));
        }
        code[i] = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
    }
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(12345);
    std::shuffle(perm.begin(), perm.end(), rng);
    // rounds: [0, n/2^k) ... the last round holds half of the points
    std::vector<int64_t> bounds;
    int64_t m = n;
    while (m > 64) {
        bounds.push_back(m);
        m /= 2;
    }
    bounds.push_back(m);
    std::reverse(bounds.begin(), bounds.end());
    int64_t start = 0;
    for (int64_t b : bounds) {
        std::sort(perm.begin() + start, perm.begin() + b, [&](int x, int y) { return code[x] < code[y]; });
        start = b;
    }
    return perm;
}

}  // namespace

extern "C" {

struct dtfe_tess {
    std::vector<int32_t> cells, nbrs;
    std::vector<int32_t> dup_of;
    int64_t n_dup = 0;
};

// Builds the tessellation of n points ([n][3] FP64). Returns 0 on success,
// 1 for bad arguments, 2 when all points are coplanar (no tetrahedron).
int dtfe_tess_build(int64_t n, const double *pts, dtfe_tess **out) {
    if (!out || n < 0 || (n > 0 && !pts) || n >= INT32_MAX) return 1;
    *out = nullptr;
    dtfe_tess *res = new dtfe_tess();
    res->dup_of.assign(n, -1);
    if (n < 4) {
        *out = res;
        return 2;
    }
    std::vector<int> order = brio_order(n, pts);
    Delaunay D;
    D.pts = pts;
    D.dup_of.assign(n, -1);
    D.tets.reserve((size_t)(n * 7.5) + 64);
    D.stamp.reserve(D.tets.capacity());
    D.in_conflict.reserve(D.tets.capacity());
    // four affinely independent starting points
    int i0 = order[0], i1 = -1, i2 = -1, i3 = -1;
    size_t k = 1;
    auto same = [&](int a, int b) {
        return pts[3 * a] == pts[3 * b] && pts[3 * a + 1] == pts[3 * b + 1] && pts[3 * a + 2] == pts[3 * b + 2];
    };
    for (; k < order.size(); ++k)
        if (!same(order[k], i0)) {
            i1 = order[k++];
            break;
        }
    auto collinear = [&](int a, int b, int c) {
        // all three 2D projections degenerate
        const double *A = pts + 3 * a, *B = pts + 3 * b, *Cc = pts + 3 * c;
        double o[3] = {A[0] + 1.0, A[1] + 0.5, A[2] + 0.25};
        double e[3] = {A[0] - 0.25, A[1] + 1.0, A[2] + 0.5};
        return orient_sign(A, B, Cc, o) == 0 && orient_sign(A, B, Cc, e) == 0;
    };
    for (; i1 >= 0 && k < order.size(); ++k)
        if (!collinear(i0, i1, order[k])) {
            i2 = order[k++];
            break;
        }
    for (; i2 >= 0 && k < order.size(); ++k)
        if (orient_sign(pts + 3 * i0, pts + 3 * i1, pts + 3 * i2, pts + 3 * order[k]) != 0) {
            i3 = order[k];
            break;
        }
    if (i3 < 0) {
        *out = res;
        return 2;
    }
    D.init(i0, i1, i2, i3);
    // DTFE_TESS_PROGRESS=k: a line on stderr after every k inserted points
    int64_t every = 0;
    if (const char *e = std::getenv("DTFE_TESS_PROGRESS")) every = std::atoll(e);
    for (size_t j = 0; j < order.size(); ++j) {
        int p = order[j];
        if (p == i0 || p == i1 || p == i2 || p == i3) continue;
        D.insert(p);
        if (every > 0 && (int64_t)(j + 1) % every == 0) {
            std::fprintf(stderr, "[tessellation] %zu of %zu points inserted\n", j + 1, order.size());
            std::fflush(stderr);
        }
    }
    // finite tetrahedra
    std::vector<int> live(D.tets.size(), 0);
    for (int t : D.free_list) live[t] = -1;
    std::vector<int> fid(D.tets.size(), -1);
    int64_t nf = 0;
    for (size_t t = 0; t < D.tets.size(); ++t)
        if (live[t] == 0 && !D.is_inf(D.tets[t])) fid[t] = (int)nf++;
    res->cells.resize(4 * nf);
    res->nbrs.resize(4 * nf);
    for (size_t t = 0; t < D.tets.size(); ++t) {
        if (fid[t] < 0) continue;
        int64_t f = fid[t];
        for (int i = 0; i < 4; ++i) {
            res->cells[4 * f + i] = D.tets[t].v[i];
            res->nbrs[4 * f + i] = fid[D.tets[t].n[i]];
        }
    }
    res->dup_of = D.dup_of;
    for (int32_t d : res->dup_of) res->n_dup += d >= 0;
    *out = res;
    return 0;
}

int64_t dtfe_tess_n_cells(const dtfe_tess *t) { return t ? (int64_t)t->cells.size() / 4 : 0; }

int64_t dtfe_tess_n_duplicates(const dtfe_tess *t) { return t ? t->n_dup : 0; }

int dtfe_tess_copy(const dtfe_tess *t, int32_t *cells, int32_t *nbrs, int32_t *dup_of) {
    if (!t) return 1;
    if (cells) std::memcpy(cells, t->cells.data(), t->cells.size() * sizeof(int32_t));
    if (nbrs) std::memcpy(nbrs, t->nbrs.data(), t->nbrs.size() * sizeof(int32_t));
    if (dup_of) std::memcpy(dup_of, t->dup_of.data(), t->dup_of.size() * sizeof(int32_t));
    return 0;
}

void dtfe_tess_free(dtfe_tess *t) { delete t; }

// Renumbers the vertices of a mesh along a Morton curve and sorts its
// tetrahedra by the Morton code of their centroids, in place; neighbour
// indices follow. vertex_perm[new] = old. Spatially close vertices and
// tetrahedra then share cache lines, which is what the gathers, atomics and
// walks of the GPU kernels need.
int dtfe_mesh_spatial_order(int64_t nv, double *pos, int64_t nc, int32_t *cells, int32_t *nbrs,
                            int64_t *vertex_perm) {
    if (nv < 0 || nc < 0 || (nv && !pos) || (nc && (!cells || !nbrs))) return 1;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], pos[3 * i + k]);
            hi[k] = std::max(hi[k], pos[3 * i + k]);
        }
    double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    auto code_of = [&](const double *p) {
        uint64_t c[3];
        for (int k = 0; k < 3; ++k) {
            double f = (p[k] - lo[k]) / ext;
            c[k] = (uint64_t)std::min(2097151.0, std::max(0.0, f * 2097152.0));
        }
        return spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
    };
    // vertices
    std::vector<std::pair<uint64_t, int64_t>> vk(nv);
    for (int64_t i = 0; i < nv; ++i) vk[i] = {code_of(pos + 3 * i), i};
    std::sort(vk.begin(), vk.end());
    std::vector<int32_t> new_of(nv);
    std::vector<double> npos(3 * nv);
    for (int64_t j = 0; j < nv; ++j) {
        int64_t old = vk[j].second;
        new_of[old] = (int32_t)j;
        if (vertex_perm) vertex_perm[j] = old;
        for (int k = 0; k < 3; ++k) npos[3 * j + k] = pos[3 * old + k];
    }
    std::memcpy(pos, npos.data(), npos.size() * sizeof(double));
    for (int64_t i = 0; i < 4 * nc; ++i) cells[i] = new_of[cells[i]];
    // tetrahedra by centroid
    std::vector<std::pair<uint64_t, int64_t>> tk(nc);
    for (int64_t t = 0; t < nc; ++t) {
        double c[3] = {0, 0, 0};
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 3; ++k) c[k] += 0.25 * pos[3 * (int64_t)cells[4 * t + i] + k];
        tk[t] = {code_of(c), t};
    }
    std::sort(tk.begin(), tk.end());
    std::vector<int32_t> tnew(nc), nc_cells(4 * nc), nc_nbrs(4 * nc);
    for (int64_t j = 0; j < nc; ++j) tnew[tk[j].second] = (int32_t)j;
    for (int64_t j = 0; j < nc; ++j) {
        int64_t old = tk[j].second;
        for (int i = 0; i < 4; ++i) {
            nc_cells[4 * j + i] = cells[4 * old + i];
            int32_t n = nbrs[4 * old + i];
            nc_nbrs[4 * j + i] = n < 0 ? -1 : tnew[n];
        }
    }
    std::memcpy(cells, nc_cells.data(), nc_cells.size() * sizeof(int32_t));
    std::memcpy(nbrs, nc_nbrs.data(), nc_nbrs.size() * sizeof(int32_t));
    return 0;
}

// Exact predicate signs, exported for tests.
int dtfe_orient_sign(const double *a, const double *b, const double *c, const double *d) {
    return orient_sign(a, b, c, d);
}

int dtfe_lifted_sign(const double *a, const double *b, const double *c, const double *d, const double *p) {
    return lifted_sign(a, b, c, d, p);
}

}  // extern "C"
