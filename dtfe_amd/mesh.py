"""Host front end: periodic padding, average density and the flattened
Delaunay mesh handed to the GPU kernels.

Reference host steps mirrored here:

* ``findParticlesInBox`` / ``insertParticlesInBox`` (src/subpartition.h:79-154):
  the 27 periodic translates of the padded box, closed bounds, positions
  shifted by the translate.  The overlap test at src/subpartition.h:129
  accepts every translate, so all 27 are scanned.
* ``averageDensity`` (src/subpartition.h:437-449) with the reference's box
  volume of 35000^3 (src/box.h:141-148).
* ``updatePadding`` (src/user_options.cc:1062-1089): padding length
  ``paddingParticles / N^(1/3) * L`` for DTFE.
* ``delaunayTriangulation`` (src/CGAL_triangulation/triangulation.cpp:162-195):
  the tessellation itself.  CGAL is not available in this image, so the
  mesh comes from the native tessellator when it is built, or from Qhull
  (scipy) otherwise.  Duplicate points keep the data of the last copy, as
  CGAL's insert-then-setData does (triangulation.cpp:183-185).

The flattened mesh: ``pos`` [nv, 3] float64, ``cells`` [nc, 4] int32 with
every tetrahedron positively oriented (det[x1-x0; x2-x0; x3-x0] > 0),
``nbrs`` [nc, 4] int32 with the neighbour across the face opposite vertex i
or -1 on the convex hull.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

REFERENCE_BOX_VOLUME = 35000.0 ** 3


@dataclass
class Mesh:
    pos: np.ndarray                 # [nv, 3] float64
    cells: np.ndarray               # [nc, 4] int32
    nbrs: np.ndarray                # [nc, 4] int32
    point_of_vertex: np.ndarray     # [nv] int64: input point kept for this vertex
    vertex_of_point: np.ndarray     # [npoints] int64: vertex of each input point
    backend: str = "qhull"
    seconds: float = 0.0
    info: dict = field(default_factory=dict)

    @property
    def n_vertices(self) -> int:
        return int(self.pos.shape[0])

    @property
    def n_cells(self) -> int:
        return int(self.cells.shape[0])


def average_density(weight: np.ndarray) -> float:
    """rho_bar = sum(w) / 35000^3 (the reference's Box::volume constant)."""
    return float(np.sum(np.asarray(weight, dtype=np.float64))) / REFERENCE_BOX_VOLUME


def padding_length(n_particles: int, box_len: float, padding_particles: float = 5.0) -> float:
    if padding_particles <= 0:
        return 0.0
    return float(padding_particles) / float(n_particles) ** (1.0 / 3.0) * float(box_len)


def periodic_images(pos: np.ndarray, full_box, padded_box):
    """Images of the particles inside the padded box, in the reference's order.

    Returns (image_pos [k, 3], image_of [k]) where image_of maps an image to
    its particle.
    """
    pos = np.asarray(pos, dtype=np.float64)
    fb = np.asarray(full_box, dtype=np.float64)
    pb = np.asarray(padded_box, dtype=np.float64)
    L = fb[1::2] - fb[0::2]
    outs, idxs = [], []
    for i1 in range(3):
        for i2 in range(3):
            for i3 in range(3):
                off = np.array([(i1 - 1) * L[0], (i2 - 1) * L[1], (i3 - 1) * L[2]])
                lo = pb[0::2] + off
                hi = pb[1::2] + off
                m = np.all((pos >= lo) & (pos <= hi), axis=1)
                sel = np.nonzero(m)[0]
                if sel.size:
                    outs.append(pos[sel] - off)
                    idxs.append(sel)
    if not outs:
        return np.zeros((0, 3)), np.zeros(0, dtype=np.int64)
    return np.concatenate(outs), np.concatenate(idxs).astype(np.int64)


def _dedupe_keep_last(points: np.ndarray):
    """Unique points; each keeps the index of its LAST occurrence."""
    n = points.shape[0]
    rev = points[::-1]
    view = np.ascontiguousarray(rev).view(np.dtype((np.void, rev.dtype.itemsize * 3)))
    _, first_in_rev, inverse = np.unique(view.ravel(), return_index=True, return_inverse=True)
    last = n - 1 - first_in_rev              # original index of the kept copy
    order = np.argsort(last, kind="stable")  # vertices in input order
    kept = last[order]
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    vertex_of_point = rank[inverse.ravel()][::-1].copy()
    return kept.astype(np.int64), vertex_of_point.astype(np.int64)


def _orient_fp64(pos, cells):
    a = pos[cells[:, 0]]
    u = pos[cells[:, 1]] - a
    v = pos[cells[:, 2]] - a
    w = pos[cells[:, 3]] - a
    return np.einsum("ij,ij->i", u, np.cross(v, w))


def _qhull(pts: np.ndarray):
    from scipy.spatial import Delaunay
    d = Delaunay(pts, qhull_options="Qbb Qc Qz Q12 Qt")
    cells = d.simplices.astype(np.int32)
    nbrs = d.neighbors.astype(np.int32)
    return cells, nbrs


def tessellate(points: np.ndarray, backend: str = "auto", spatial_order: bool = True) -> Mesh:
    """Delaunay tessellation of the points as a flattened mesh.

    With spatial_order (the default) vertices and tetrahedra are renumbered
    along a Morton curve, so that neighbours in space are neighbours in
    memory; point_of_vertex / vertex_of_point keep the link to the input."""
    t0 = time.perf_counter()
    points = np.ascontiguousarray(points, dtype=np.float64)
    kept, vop = _dedupe_keep_last(points)
    pos = np.ascontiguousarray(points[kept])
    if backend == "auto":
        backend = "native"
        try:
            from . import native_delaunay  # noqa: F401
            if not native_delaunay.available():
                backend = "qhull"
        except Exception:
            backend = "qhull"
    if pos.shape[0] < 4:
        cells = np.zeros((0, 4), np.int32)
        nbrs = np.zeros((0, 4), np.int32)
    elif backend == "native":
        from . import native_delaunay
        cells, nbrs = native_delaunay.delaunay(pos)
    elif backend == "qhull":
        cells, nbrs = _qhull(pos)
    else:
        raise ValueError(f"unknown tessellation backend {backend!r}")
    if cells.shape[0]:
        det = _orient_fp64(pos, cells)
        neg = det < 0
        if np.any(neg):
            cells[neg] = cells[neg][:, [0, 1, 3, 2]]
            nbrs[neg] = nbrs[neg][:, [0, 1, 3, 2]]
    if spatial_order and cells.shape[0]:
        from . import native_delaunay
        pos, cells, nbrs, perm = native_delaunay.spatial_order(pos, cells, nbrs)
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size)
        kept = kept[perm]
        vop = inv[vop]
    m = Mesh(pos=pos, cells=np.ascontiguousarray(cells, dtype=np.int32),
             nbrs=np.ascontiguousarray(nbrs, dtype=np.int32), point_of_vertex=kept,
             vertex_of_point=vop, backend=backend, seconds=time.perf_counter() - t0)
    return m


@dataclass
class PreparedInput:
    """Particles after the reference's host steps, ready to tessellate."""
    pos: np.ndarray
    weight: np.ndarray
    vel: Optional[np.ndarray]
    scalar: Optional[np.ndarray]
    image_of: np.ndarray       # particle index of every image
    rho_bar: float
    region: tuple
    padding: float


def prepare_particles(ps, padding_particles: float = 5.0, rho_bar: Optional[float] = None) -> PreparedInput:
    """Periodic padding and average density as DTFE() does before its
    tessellation (src/DTFE.cpp:101-103, :222-237)."""
    box = np.asarray(ps.box, dtype=np.float64)
    L = box[1::2] - box[0::2]
    if rho_bar is None:
        rho_bar = average_density(ps.weight)
    pad = padding_length(ps.n, float(L[0]), padding_particles)
    if ps.periodic:
        padded = box.copy()
        padded[0::2] -= pad
        padded[1::2] += pad
        ipos, iof = periodic_images(ps.pos, box, padded)
    else:
        ipos, iof = np.asarray(ps.pos, dtype=np.float64), np.arange(ps.n, dtype=np.int64)
    return PreparedInput(pos=ipos, weight=np.asarray(ps.weight, np.float64)[iof],
                         vel=None if ps.vel is None else np.asarray(ps.vel, np.float64)[iof],
                         scalar=None if ps.scalar is None else np.asarray(ps.scalar, np.float64)[iof],
                         image_of=iof, rho_bar=float(rho_bar), region=tuple(box.tolist()), padding=pad)


def vertex_data(mesh: Mesh, values: Optional[np.ndarray]):
    """Per-vertex copy of per-point data (duplicates keep the last copy)."""
    if values is None:
        return None
    return np.ascontiguousarray(np.asarray(values)[mesh.point_of_vertex])
