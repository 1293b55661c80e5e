"""The reference's 5-particle demo input (tests/golden/demo5.json, written by
tests/golden/make_demo_fixture.py from demo/example_input_data_from_text_file.txt:1-7)
through the default DTFE() steps: periodic images, rho_bar = sum w / 35000^3,
vertex densities, cell-centre and Monte-Carlo densities on a 5^3 grid.
The CPU test pins the host steps and the oracle against the fixture; the GPU
test runs the device path against it. The 27 translates of each particle
form cubic lattices, whose cospherical vertices leave the Delaunay
tessellation non-unique: CGAL may split them differently, so these values
pin this build, not the reference, there."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from dtfe_amd import mesh as M, synthetic as S
from oracle import oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "demo5.json")


def _load():
    with open(FIX) as f:
        return json.load(f)


def _prepared(fx):
    rows = np.asarray(fx["particles"])
    ps = S.ParticleSet(pos=rows[:, :3].copy(), weight=rows[:, 3].copy(), vel=np.zeros((len(rows), 3)),
                       box=tuple(fx["box"]), periodic=True)
    return ps, M.prepare_particles(ps, fx["padding_particles"])


def test_demo_host_steps_and_oracle():
    fx = _load()
    ps, prep = _prepared(fx)
    # padding 5 / 5^(1/3) * 10 = 29.2 exceeds the box: all 27 translates of all 5 particles
    assert prep.pos.shape[0] == 27 * 5
    assert np.array_equal(prep.pos, np.asarray(fx["image_positions"]))
    assert np.array_equal(prep.image_of, np.asarray(fx["image_of"]))
    assert prep.rho_bar == fx["rho_bar"] == float(np.sum(ps.weight)) / 35000.0 ** 3
    m = M.tessellate(prep.pos)
    assert m.n_cells == fx["n_cells"]
    rho, _, _ = O.vertex_density(m, M.vertex_data(m, prep.weight), prep.rho_bar)
    rp = np.zeros(prep.pos.shape[0])
    rp[m.point_of_vertex] = rho
    exp = np.asarray(fx["vertex_density_of_image"])
    assert np.allclose(rp, exp, rtol=1e-14, atol=0)
    # vertices on the convex hull (on a face without a neighbour) get 0, the
    # others a positive density (triangulation.cpp:200-279)
    faces = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    hull = np.zeros(m.n_vertices, dtype=bool)
    for f in range(4):
        open_cells = m.nbrs[:, f] < 0
        for k in faces[f]:
            hull[m.cells[open_cells, k]] = True
    assert hull.any() and (~hull).any()
    assert np.all(rho[hull] == 0.0) and np.all(rho[~hull] > 0.0)
    out = O.sample(m, rho, None, None, fields=("density",), kind="grid", box=fx["box"], n=fx["grid"], method=0)
    assert np.allclose(out["density"], fx["density_centre"], rtol=1e-14, atol=0)


@pytest.mark.gpu
def test_demo_device_path(gpu):
    import torch
    from dtfe_amd import api
    fx = _load()
    ps, _ = _prepared(fx)
    dev = api.prepare_particles(torch.from_numpy(ps.pos).cuda(), torch.from_numpy(ps.weight).cuda(), ps.box)
    assert np.array_equal(dev.pos.cpu().numpy(), np.asarray(fx["image_positions"]))
    # the device's fixed-order tree sum against the reference's serial one: rounding only
    assert abs(dev.rho_bar - fx["rho_bar"]) <= 4 * 5 * np.finfo(np.float64).eps * fx["rho_bar"]
    m = M.tessellate(dev.pos.cpu().numpy())
    dm = api.DeviceMesh.from_mesh(m, gpu)
    w = api.gather_rows(torch.from_numpy(ps.weight).cuda(), api.vertex_particles(dev, m))
    rho = api.vertex_density(dm, w, dev.rho_bar)
    rp = np.zeros(dev.pos.shape[0])
    rp[m.point_of_vertex] = rho.cpu().numpy()
    assert np.allclose(rp, fx["vertex_density_of_image"], rtol=1e-12, atol=0)
    spec = api.GridSpec(fx["box"], fx["grid"])
    c = api.sample(dm, rho, None, None, ["density"], spec, method=0)
    assert np.allclose(c["density"].cpu().numpy(), fx["density_centre"], rtol=1e-12, atol=0)
    mc = api.sample(dm, rho, None, None, ["density"], spec, method=2, samples=8, seed=1)
    assert np.allclose(mc["density"].cpu().numpy(), fx["density_mc8_seed1"], rtol=1e-12, atol=0)
