"""CPU tests of the native libraries: the exact tessellator and the exports
of the C ABI (no GPU needed: nothing here launches a kernel)."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from dtfe_amd import mesh as M, native_delaunay as ND

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _validate(p, cells, nbrs, check_delaunay=3000):
    from scipy.spatial import ConvexHull
    X = p[cells]
    det = np.einsum("ij,ij->i", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))
    assert all(ND.orient_sign(*p[c]) > 0 for c in cells), "non-positive tetrahedron"
    # neighbour symmetry and shared faces
    for t in range(len(cells)):
        for i in range(4):
            n = nbrs[t, i]
            if n < 0:
                continue
            face = set(cells[t]) - {cells[t, i]}
            assert t in nbrs[n]
            j = list(nbrs[n]).index(t)
            assert set(cells[n]) - {cells[n, j]} == face
    assert det.sum() / 6 == pytest.approx(ConvexHull(p).volume, rel=1e-9)
    # local Delaunay: no neighbour's opposite vertex strictly inside a circumsphere
    for t in range(min(len(cells), check_delaunay)):
        for i in range(4):
            n = nbrs[t, i]
            if n < 0:
                continue
            opp = [v for v in cells[n] if v not in cells[t]][0]
            assert ND.lifted_sign(*p[cells[t]], p[opp]) >= 0


def test_random_points_match_qhull():
    rng = np.random.default_rng(0)
    p = rng.random((3000, 3)) * 10
    c, nb = ND.delaunay(p)
    cq, _ = M._qhull(p)
    assert set(map(tuple, np.sort(c, 1))) == set(map(tuple, np.sort(cq, 1)))
    _validate(p, c, nb)


@pytest.mark.parametrize("k", [3, 5])
def test_cospherical_lattice(k):
    g = np.arange(k) + 0.5
    p = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    c, nb = ND.delaunay(p)
    _validate(p, c, nb)


def test_points_on_a_sphere_plus_centre():
    rng = np.random.default_rng(1)
    th, ph = rng.random(300) * 2 * np.pi, np.arccos(rng.random(300) * 2 - 1)
    s = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], 1)
    p = np.concatenate([s, [[0.0, 0.0, 0.0]]])
    c, nb = ND.delaunay(p)
    _validate(p, c, nb)


def test_tessellate_dedupes_keeping_last_copy():
    rng = np.random.default_rng(2)
    p = rng.random((200, 3))
    q = np.concatenate([p, p[:10]])          # the last 10 repeat the first 10
    m = M.tessellate(q, backend="native")
    assert m.n_vertices == 200
    # the kept copies of the repeated points are the later ones
    for i in range(10):
        v = m.vertex_of_point[i]
        assert m.vertex_of_point[200 + i] == v and m.point_of_vertex[v] == 200 + i


def test_lifted_sign_convention():
    a, b, c, d = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    assert ND.orient_sign(a, b, c, d) == 1
    assert ND.lifted_sign(a, b, c, d, np.array([0.25, 0.25, 0.25])) == -1   # inside
    assert ND.lifted_sign(a, b, c, d, np.array([2.0, 2.0, 2.0])) == 1       # outside
    assert ND.lifted_sign(a, b, c, d, np.array([1.0, 1.0, 0.0])) == 0       # on the sphere


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "dtfe_hip.h")).read()
    return sorted(set(re.findall(r"\b(dtfe_hip_\w+)\s*\(", src)))


def test_c_abi_exports_every_declared_symbol():
    from dtfe_amd import _native
    lib = C.CDLL(_native.LIB_PATH)
    names = _declared_symbols()
    assert "dtfe_hip_sample" in names and "dtfe_hip_vertex_density" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_native.EXPORTED)


def test_library_version_string():
    from dtfe_amd import _native
    lib = _native.load()
    assert b"gfx950" in lib.dtfe_hip_version()
