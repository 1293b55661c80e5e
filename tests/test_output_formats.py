"""The drop-in's output writers (writeOutputData, src/input_output.cc:213-298)
in both builds, on known field values written without a DTFE() call: file
names and field order; text formats 111-114 byte for byte against C's %g
(the reference writes with a default ostream, 6 significant digits); raw
format 101; format 100's 1024-byte header, framing and data."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from dtfe_amd import build as B

GRID = (3, 2, 4)
BOX = (0.0, 10.0, -2.0, 2.0, 0.0, 1.0)
FIELDS = {  # file extension: (field, components, header file type)
    ".den": ("u", 1, 1), ".vel": ("u", 3, 11), ".a_velDiv": ("a", 1, 13), ".a_velGrad": ("a", 9, 12),
    ".a_velVort": ("a", 3, 15),
}
REALS = {"float": np.float32, "double": np.float64}


def _exe(real):
    exe = B.DROPIN_RUN[real]
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built")
    return exe


def _run(tmp, real, fmt, extra=(), expect_ok=True):
    rows = int(np.prod(GRID))
    opts = [f"box {' '.join(map(str, BOX))}", f"grid {' '.join(map(str, GRID))}",
            "fields_u density,velocity", "fields_a divergence,gradient,vorticity",
            f"output {fmt} {tmp / 'out'}", "program DTFE in.bin out --grid 3 2 4", f"synthetic_rows {rows}", *extra]
    (tmp / "o.txt").write_text("\n".join(opts) + "\n")
    (tmp / "p.bin").write_bytes(b"")
    r = subprocess.run([_exe(real), str(tmp / "o.txt"), str(tmp / "p.bin"), str(tmp / "dump")],
                       capture_output=True, text=True, timeout=60)
    if expect_ok:
        assert r.returncode == 0, r.stderr + r.stdout
    return r


def _values(c, R):
    n = int(np.prod(GRID))
    i, j = np.meshgrid(np.arange(n), np.arange(c), indexing="ij")
    return ((i * c + j) * 0.125 - 3.0 + 1.0 / 3.0).astype(R)


def _g(x):
    return "%g" % float(x)


def _row(v):
    return "".join(_g(x) + "\t" for x in v) + "\n" if len(v) > 1 else _g(v[0]) + "\n"


def _lines(path):
    return open(path, newline="").read()


@pytest.mark.parametrize("real", ["float", "double"])
def test_raw_binary_101(tmp_path, real):
    _run(tmp_path, real, 101)
    R = REALS[real]
    for ext, (_, c, _) in FIELDS.items():
        got = np.fromfile(tmp_path / f"out{ext}", dtype=R)
        assert np.array_equal(got, _values(c, R).ravel()), ext


@pytest.mark.parametrize("real", ["float", "double"])
def test_text_111_112(tmp_path, real):
    R = REALS[real]
    _run(tmp_path, real, 111)
    for ext, (_, c, _) in FIELDS.items():
        assert _lines(tmp_path / f"out{ext}") == "".join(_row(v) for v in _values(c, R)), ext
    t2 = tmp_path / "t112"
    t2.mkdir()
    _run(t2, real, 112)
    for ext, (_, c, _) in FIELDS.items():
        vals = _values(c, R)
        exp = []
        for i in range(GRID[0]):
            for j in range(GRID[1]):
                for k in range(GRID[2]):
                    exp.append(f"{i}\t{j}\t{k}\t" + _row(vals[(i * GRID[1] + j) * GRID[2] + k]))
        assert _lines(t2 / f"out{ext}") == "".join(exp), ext


@pytest.mark.parametrize("real", ["float", "double"])
def test_text_113_cell_centres(tmp_path, real):
    """x = xmin + dx (i + 1/2), dx = (xmax - xmin) / n in Real (text_io.cc:395-410)."""
    R = REALS[real]
    _run(tmp_path, real, 113)
    b = np.asarray(BOX, dtype=R)
    dx = [(b[2 * a + 1] - b[2 * a]) / R(GRID[a]) for a in range(3)]
    vals = _values(1, R)
    exp = []
    for i in range(GRID[0]):
        for j in range(GRID[1]):
            for k in range(GRID[2]):
                pos = [R(float(b[2 * a]) + float(dx[a]) * (ii + 0.5)) for a, ii in enumerate((i, j, k))]
                exp.append("".join(_g(p) + "\t" for p in pos) + _row(vals[(i * GRID[1] + j) * GRID[2] + k]))
    assert _lines(tmp_path / "out.den") == "".join(exp)
    # grid indices and cell centres do not exist for user points or cones
    t2 = tmp_path / "cone"
    t2.mkdir()
    r = _run(t2, real, 113, extra=("cone 10 100 20 60 30 80", "origin 0 0 0"), expect_ok=False)
    assert r.returncode != 0 and "writeTextFile_samplingPosition" in r.stderr


def test_text_114_cone_positions(tmp_path):
    """Cone rows as Cartesian positions with the writer's own 3.14/180
    degree factor (text_io.cc:511-548), double build."""
    _run(tmp_path, "double", 114, extra=("cone 10 100 20 60 30 80", "origin 1 2 3"))
    cone = np.array([10, 100, 20, 60, 30, 80.0])
    dx = [(cone[2 * a + 1] - cone[2 * a]) / GRID[a] for a in range(3)]
    f = 3.14 / 180.0
    vals = _values(1, np.float64)
    exp = []
    for i in range(GRID[0]):
        for j in range(GRID[1]):
            for k in range(GRID[2]):
                r, th, ph = (cone[2 * a] + dx[a] * (ii + 0.5) for a, ii in enumerate((i, j, k)))
                pos = (1 + r * np.sin(th * f) * np.cos(ph * f), 2 + r * np.sin(th * f) * np.sin(ph * f),
                       3 + r * np.cos(th * f))
                exp.append("".join(_g(p) + "\t" for p in pos) + _row(vals[(i * GRID[1] + j) * GRID[2] + k]))
    assert _lines(tmp_path / "out.den") == "".join(exp)
    t2 = tmp_path / "grid"
    t2.mkdir()
    r = _run(t2, "double", 114, expect_ok=False)
    assert r.returncode != 0 and "redshift cone" in r.stderr


HEADER = np.dtype([("gridSize", "<u8", 3), ("totalGrid", "<u8"), ("fileType", "<i4"), ("noDensityFiles", "<i4"),
                   ("densityFileGrid", "<i4", 3), ("indexDensityFile", "<i4"), ("box", "<f8", 6),
                   ("npartTotal", "<u8", 6), ("mass", "<f8", 6), ("time", "<f8"), ("redshift", "<f8"),
                   ("BoxSize", "<f8"), ("Omega0", "<f8"), ("OmegaLambda", "<f8"), ("HubbleParam", "<f8"),
                   ("method", "<u8"), ("fill", "S760"), ("FILE_ID", "<u8")])


@pytest.mark.parametrize("real", ["float", "double"])
def test_density_header_100(tmp_path, real):
    """Density_header (density_file_io.cc:44-160) is 1024 bytes on x86-64;
    the file is [1024][header][1024][n bytes][data][n bytes]."""
    assert HEADER.itemsize == 1024
    R = REALS[real]
    _run(tmp_path, real, 100, extra=("SPH 16",))
    for ext, (_, c, ftype) in FIELDS.items():
        raw = (tmp_path / f"out{ext}").read_bytes()
        n = int(np.prod(GRID)) * c * np.dtype(R).itemsize
        assert len(raw) == 8 + 1024 + 8 + 8 + n + 8
        assert np.frombuffer(raw[:8], "<u8")[0] == 1024 and np.frombuffer(raw[1032:1040], "<u8")[0] == 1024
        assert np.frombuffer(raw[1040:1048], "<u8")[0] == n and np.frombuffer(raw[-8:], "<u8")[0] == n
        h = np.frombuffer(raw[8:1032], HEADER)[0]
        assert list(h["gridSize"]) == list(GRID) and h["totalGrid"] == np.prod(GRID)
        assert h["fileType"] == ftype, ext
        assert h["noDensityFiles"] == 1 and list(h["densityFileGrid"]) == [1, 1, 1] and h["indexDensityFile"] == -1
        assert np.array_equal(h["box"], np.asarray(BOX))
        assert h["method"] == 3                      # SPH
        assert h["time"] == -1 and h["HubbleParam"] == -1 and h["FILE_ID"] == 1
        assert h["fill"].startswith(b"DTFE in.bin out --grid 3 2 4 ;  ")
        assert np.array_equal(np.frombuffer(raw[1048:-8], R), _values(c, R).ravel())
