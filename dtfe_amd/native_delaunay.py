"""ctypes binding of the native exact Delaunay tessellator (libdtfe_host.so)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdtfe_host.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dtfe_amd.build`")
        lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        lib.dtfe_tess_build.argtypes = [C.c_int64, vp, C.POINTER(vp)]
        lib.dtfe_tess_build.restype = C.c_int
        lib.dtfe_tess_n_cells.argtypes = [vp]
        lib.dtfe_tess_n_cells.restype = C.c_int64
        lib.dtfe_tess_n_duplicates.argtypes = [vp]
        lib.dtfe_tess_n_duplicates.restype = C.c_int64
        lib.dtfe_tess_copy.argtypes = [vp, vp, vp, vp]
        lib.dtfe_tess_free.argtypes = [vp]
        lib.dtfe_orient_sign.argtypes = [vp, vp, vp, vp]
        lib.dtfe_mesh_spatial_order.argtypes = [C.c_int64, vp, C.c_int64, vp, vp, vp]
        lib.dtfe_mesh_spatial_order.restype = C.c_int
        lib.dtfe_lifted_sign.argtypes = [vp, vp, vp, vp, vp]
        _lib = lib
    return _lib


def available() -> bool:
    return os.path.exists(LIB_PATH)


def delaunay(points: np.ndarray):
    """(cells [nc, 4] int32, nbrs [nc, 4] int32) of the distinct points."""
    lib = _load()
    pts = np.ascontiguousarray(points, dtype=np.float64)
    h = C.c_void_p()
    rc = lib.dtfe_tess_build(pts.shape[0], pts.ctypes.data_as(C.c_void_p), C.byref(h))
    try:
        if rc == 1:
            raise ValueError("bad arguments to the tessellator")
        nc = lib.dtfe_tess_n_cells(h)
        cells = np.empty((nc, 4), dtype=np.int32)
        nbrs = np.empty((nc, 4), dtype=np.int32)
        if nc:
            lib.dtfe_tess_copy(h, cells.ctypes.data_as(C.c_void_p), nbrs.ctypes.data_as(C.c_void_p), None)
        if lib.dtfe_tess_n_duplicates(h):
            raise ValueError("duplicate points passed to the tessellator")
        return cells, nbrs
    finally:
        lib.dtfe_tess_free(h)


def spatial_order(pos: np.ndarray, cells: np.ndarray, nbrs: np.ndarray):
    """Morton renumbering of vertices and tetrahedra (in place on copies).
    Returns (pos, cells, nbrs, vertex_perm) with vertex_perm[new] = old."""
    lib = _load()
    pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
    cells = np.ascontiguousarray(cells, dtype=np.int32).copy()
    nbrs = np.ascontiguousarray(nbrs, dtype=np.int32).copy()
    perm = np.empty(pos.shape[0], dtype=np.int64)
    rc = lib.dtfe_mesh_spatial_order(pos.shape[0], pos.ctypes.data_as(C.c_void_p), cells.shape[0],
                                     cells.ctypes.data_as(C.c_void_p), nbrs.ctypes.data_as(C.c_void_p),
                                     perm.ctypes.data_as(C.c_void_p))
    if rc:
        raise ValueError("bad mesh arrays")
    return pos, cells, nbrs, perm


def orient_sign(a, b, c, d) -> int:
    lib = _load()
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, d)]
    return int(lib.dtfe_orient_sign(*[x.ctypes.data_as(C.c_void_p) for x in arrs]))


def lifted_sign(a, b, c, d, p) -> int:
    lib = _load()
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c, d, p)]
    return int(lib.dtfe_lifted_sign(*[x.ctypes.data_as(C.c_void_p) for x in arrs]))
