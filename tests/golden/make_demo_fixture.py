"""Writes tests/golden/demo5.json: the 5 particles of the reference's demo
input (demo/example_input_data_from_text_file.txt:1-7: box [0, 10]^3,
positions and weights) through the path as DTFE() runs it by default
(periodic, padding of 5 particles), with the oracle's outputs as the
expected values: periodic images, average density, vertex densities at the
vertices' particles, cell-centre density on a 5^3 grid and a Monte-Carlo
average (8 samples per cell, seed 1). Run from the repository root with
the reference checked out at /root/reference; the fixture holds the data,
the script is not needed to run the tests."""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dtfe_amd import mesh as M, synthetic as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

DEMO = "/root/reference/demo/example_input_data_from_text_file.txt"


def read_demo(path):
    lines = open(path).read().splitlines()
    n = int(lines[0])
    box = [float(x) for x in lines[1].split()]
    rows = np.array([[float(x) for x in lines[2 + i].split()] for i in range(n)])
    return box, rows


def main():
    box, rows = read_demo(DEMO)
    ps = S.ParticleSet(pos=rows[:, :3].copy(), weight=rows[:, 3].copy(), vel=np.zeros((len(rows), 3)),
                       box=tuple(box), periodic=True)
    prep = M.prepare_particles(ps)
    m = M.tessellate(prep.pos)
    w = M.vertex_data(m, prep.weight)
    rho, W, _ = O.vertex_density(m, w, prep.rho_bar)
    grid = (5, 5, 5)
    centre = O.sample(m, rho, None, None, fields=("density",), kind="grid", box=box, n=grid, method=0)
    mc = O.sample(m, rho, None, None, fields=("density",), kind="grid", box=box, n=grid, method=2, samples=8,
                  seed=1)
    # vertex density per padded particle (the vertex of each image)
    rho_particle = np.zeros(prep.pos.shape[0])
    rho_particle[m.point_of_vertex] = rho
    out = {
        "source": "demo/example_input_data_from_text_file.txt:1-7 of the reference",
        "box": box,
        "particles": rows.tolist(),
        "periodic": True,
        "padding_particles": 5.0,
        "image_positions": prep.pos.tolist(),
        "image_of": prep.image_of.tolist(),
        "rho_bar": prep.rho_bar,
        "n_cells": int(m.n_cells),
        "vertex_density_of_image": rho_particle.tolist(),
        "grid": list(grid),
        "density_centre": centre["density"].tolist(),
        "density_mc8_seed1": mc["density"].tolist(),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "demo5.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"{path}: {len(prep.pos)} images, {m.n_cells} tetrahedra, rho_bar {prep.rho_bar:.6e}")


if __name__ == "__main__":
    main()
