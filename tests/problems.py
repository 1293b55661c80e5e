"""Small seeded DTFE problems shared by the CPU and GPU tests."""
from __future__ import annotations

import functools
from dataclasses import dataclass

import numpy as np

from dtfe_amd import mesh as M, synthetic as S


@dataclass
class Problem:
    ps: S.ParticleSet
    prep: M.PreparedInput
    mesh: M.Mesh
    w: np.ndarray
    vel: np.ndarray
    scalar: np.ndarray
    rho_bar: float
    region: tuple


def _make(ps, backend="native"):
    prep = M.prepare_particles(ps)
    m = M.tessellate(prep.pos, backend=backend)
    return Problem(ps=ps, prep=prep, mesh=m, w=M.vertex_data(m, prep.weight), vel=M.vertex_data(m, prep.vel),
                   scalar=M.vertex_data(m, prep.scalar), rho_bar=prep.rho_bar, region=prep.region)


@functools.lru_cache(maxsize=None)
def uniform_problem(n=3000, seed=3):
    return _make(S.small_random(n, seed, box_len=10.0, n_scalars=3))


@functools.lru_cache(maxsize=None)
def clustered_problem(n_side=16, seed=4):
    ps = S.clustered(n_side ** 3, 20.0, 20, 0.2, 1.0, seed, with_velocity=True, lognormal_weights=True,
                     n_scalars=2)
    return _make(ps)


@functools.lru_cache(maxsize=None)
def lattice_problem(k=6):
    """Particles on a regular lattice: cospherical points, and cell-centre
    samples on faces, edges and vertices of the tessellation."""
    g = np.arange(k) + 0.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    rng = np.random.default_rng(9)
    ps = S.ParticleSet(pos=pos, weight=np.ones(pos.shape[0]), vel=rng.standard_normal(pos.shape),
                       box=(0.0, float(k), 0.0, float(k), 0.0, float(k)), periodic=True)
    return _make(ps)


@functools.lru_cache(maxsize=None)
def snapped_problem(seed=9):
    """Random particles, a fifth of them moved onto sample-grid points
    (x = (i + 0.5) * 0.5): cell-centre samples then sit exactly on mesh
    vertices, so every face through the vertex is an exact tie."""
    rng = np.random.default_rng(seed)
    n_rand, n_snap = 1500, 300
    sites = (rng.integers(0, 12, size=(n_snap, 3)) + 0.5) * 0.5
    pos = np.concatenate([rng.random((n_rand, 3)) * 6.0, sites])
    ps = S.ParticleSet(pos=pos, weight=rng.lognormal(0, 0.3, pos.shape[0]), vel=rng.standard_normal(pos.shape),
                       box=(0.0, 6.0, 0.0, 6.0, 0.0, 6.0), periodic=True)
    return _make(ps)
