"""The C++ DTFE() drop-in (libDTFE.so, through its command-line driver)
against the oracle on the same particles: the driver's own tessellation,
periodic padding and average density feed the GPU kernels."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from dtfe_amd import build as B, mesh as M
from oracle import oracle as O
from tests import problems as P

pytestmark = pytest.mark.gpu


def _run(tmp_path, opts, parts, samples=None, real="double"):
    exe = B.DROPIN_RUN[real]
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built")
    (tmp_path / "o.txt").write_text("\n".join(opts) + "\n")
    parts.astype(np.float64).tofile(tmp_path / "p.bin")
    args = [exe, str(tmp_path / "o.txt"), str(tmp_path / "p.bin")]
    if samples is not None:
        samples.astype(np.float64).tofile(tmp_path / "s.bin")
        args.append(str(tmp_path / "s.bin"))
    args.append(str(tmp_path / "out"))
    r = subprocess.run(args, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def _load(tmp_path, which, field, dtype, comps):
    a = np.fromfile(tmp_path / f"out.{which}.{field}", dtype=dtype)
    return a.reshape(-1, comps) if comps > 1 else a


def _reference(prob, fields, **kw):
    rho, _, _ = O.vertex_density(prob.mesh, prob.w, prob.rho_bar)
    return O.sample(prob.mesh, rho, prob.vel, None, fields=fields, **kw)


OPTS_GRID = ["grid 16 16 16", "periodic 1", "fields_u density,velocity,divergence",
             "fields_a density,gradient,vorticity", "method 2", "points 8", "seed 4242"]


def test_dropin_grid_centre_and_mc(gpu, tmp_path):
    prob = P.clustered_problem()
    ps = prob.ps
    parts = np.concatenate([ps.pos, ps.weight[:, None], ps.vel], 1)
    box = " ".join(str(x) for x in ps.box)
    out = _run(tmp_path, [f"box {box}"] + OPTS_GRID, parts, real="double")
    assert "particles_left=0" in out
    refu = _reference(prob, ("density", "velocity", "velocity_gradient", "velocity_divergence"), kind="grid",
                      box=ps.box, n=(16, 16, 16), method=0)
    refa = _reference(prob, ("density", "velocity_gradient", "velocity_vorticity"), kind="grid", box=ps.box,
                      n=(16, 16, 16), method=2, samples=8, seed=4242)
    for which, ref, fields in (("u", refu, (("density", 1), ("velocity", 3), ("velocity_divergence", 1))),
                               ("a", refa, (("density", 1), ("velocity_gradient", 9), ("velocity_vorticity", 3)))):
        for f, k in fields:
            got = _load(tmp_path, which, f, np.float64, k)
            exp = ref[f]
            err = np.max(np.abs(got - exp)) / np.max(np.abs(exp))
            print(f"drop-in double {which}.{f}: {err:.3e}")
            assert err < 1e-10, (which, f)
    # fields that were not requested stay empty (the gradient is dropped when only
    # divergence was asked for, as computeDivergenceShearVorticity does)
    assert not os.path.exists(tmp_path / "out.u.velocity_gradient")


def test_dropin_float_build_against_double_build(gpu, tmp_path):
    """libDTFE.so (Real = float, mixed path) against libDTFE_double.so on
    float-representable particles; periodic images are shifted in Real, so the
    two builds tessellate slightly different points near the box faces."""
    prob = P.clustered_problem()
    ps = prob.ps
    parts = np.concatenate([ps.pos, ps.weight[:, None], ps.vel], 1).astype(np.float32).astype(np.float64)
    box = " ".join(str(x) for x in ps.box)
    d64, d32 = tmp_path / "d", tmp_path / "f"
    d64.mkdir()
    d32.mkdir()
    _run(d64, [f"box {box}"] + OPTS_GRID, parts, real="double")
    _run(d32, [f"box {box}"] + OPTS_GRID, parts, real="float")
    for which, f, k in (("u", "density", 1), ("a", "density", 1), ("a", "velocity_gradient", 9)):
        a = _load(d32, which, f, np.float32, k).astype(np.float64)
        b = _load(d64, which, f, np.float64, k)
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"drop-in float vs double {which}.{f}: {err:.3e}")
        assert err < 1e-4, (which, f)


def test_dropin_user_points_and_cic(gpu, tmp_path):
    prob = P.clustered_problem()
    ps = prob.ps
    parts = np.concatenate([ps.pos, ps.weight[:, None], ps.vel], 1)
    box = " ".join(str(x) for x in ps.box)
    rng = np.random.default_rng(8)
    pts = rng.random((500, 3)) * 22.0 - 1.0
    samples = np.concatenate([pts, np.full_like(pts, 0.5)], 1)
    opts = [f"box {box}", "periodic 1", "fields_a density,velocity", "points 6", "seed 17"]
    _run(tmp_path, opts, parts, samples=samples, real="double")
    ref = _reference(prob, ("density", "velocity"), kind="points", box=ps.box, points=pts,
                     deltas=np.full_like(pts, 0.5), method=2, samples=6, seed=17)
    got = _load(tmp_path, "a", "density", np.float64, 1)
    assert got.shape[0] == ref["density"].shape[0]
    assert np.max(np.abs(got - ref["density"])) / np.max(np.abs(ref["density"])) < 1e-10
    # CIC on the same particles: non-periodic, no padding (the reference pads CIC by 0)
    tmp2 = tmp_path / "cic"
    tmp2.mkdir()
    _run(tmp2, [f"box {box}", "grid 8 8 8", "periodic 1", "fields_a density,velocity", "CIC"], parts,
         real="double")
    inside = np.all((ps.pos >= 0) & (ps.pos <= 20.0), axis=1)
    d, v = O.deposit("cic", ps.pos[inside], ps.weight[inside], ps.vel[inside], ps.box, (8, 8, 8),
                     M.average_density(ps.weight))
    got = np.fromfile(tmp2 / "out.a.density", dtype=np.float64)
    assert np.max(np.abs(got - d)) / np.max(np.abs(d)) < 1e-10


def test_dropin_method1(gpu, tmp_path):
    """Averaging method 1 through DTFE(): the driver's padding and
    tessellation, then the tetrahedron points kernel, against the oracle."""
    prob = P.clustered_problem()
    ps = prob.ps
    parts = np.concatenate([ps.pos, ps.weight[:, None], ps.vel], 1)
    box = " ".join(str(x) for x in ps.box)
    opts = [f"box {box}", "grid 12 12 12", "periodic 1", "fields_a density,velocity,gradient,divergence",
            "method 1", "points 8"]
    out = _run(tmp_path, opts, parts, real="double")
    assert "particles_left=0" in out
    rho, _, _ = O.vertex_density(prob.mesh, prob.w, prob.rho_bar)
    ref = O.grid_method1(prob.mesh, rho, prob.vel, fields=("density", "velocity", "velocity_gradient",
                                                            "velocity_divergence"),
                         box=ps.box, n=(12, 12, 12), samples=8)
    for f, k in (("density", 1), ("velocity", 3), ("velocity_gradient", 9), ("velocity_divergence", 1)):
        got = _load(tmp_path, "a", f, np.float64, k)
        err = np.max(np.abs(got - ref[f])) / np.max(np.abs(ref[f]))
        print(f"drop-in method 1 {f}: {err:.3e}")
        assert err < 1e-10, f


def test_dropin_sph(gpu, tmp_path):
    """SPH through DTFE(): the driver's periodic padding (the reference pads
    for SPH as for DTFE), then the GPU neighbour searches, against the oracle
    on the same padded particles."""
    prob = P.clustered_problem()
    ps = prob.ps
    parts = np.concatenate([ps.pos, ps.weight[:, None], ps.vel], 1)
    box = " ".join(str(x) for x in ps.box)
    _run(tmp_path, [f"box {box}", "grid 10 10 10", "periodic 1", "fields_a density,velocity", "SPH 24",
                    f"output 101 {tmp_path / 'file'}"], parts, real="double")
    # writeOutputData's raw format holds the same values as the fields
    assert (tmp_path / "file.a_den").read_bytes() == (tmp_path / "out.a.density").read_bytes()
    assert (tmp_path / "file.a_vel").read_bytes() == (tmp_path / "out.a.velocity").read_bytes()
    prep = M.prepare_particles(ps)
    iof = prep.image_of
    ref = O.sph(prep.pos, ps.weight[iof], ps.vel[iof], None, 24, prep.rho_bar, kind="grid", box=ps.box,
                n=(10, 10, 10))
    for f, k in (("density", 1), ("velocity", 3)):
        got = _load(tmp_path, "a", f, np.float64, k)
        err = np.max(np.abs(got - ref[f])) / np.max(np.abs(ref[f]))
        print(f"drop-in SPH {f}: {err:.3e}")
        assert err < 1e-10, f
