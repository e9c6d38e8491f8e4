"""The host VTK writer (mcm_vtk_write) against the reference's writer layout
(mcm_VTK.f90) and against flang's formatted output for the same values
(tests/golden/data/fortran_formats.json, made by make_fortran_formats.py
from fortran_formats.f90).  Host only: no GPU is needed."""
import json
import math
import os
import struct

import numpy as np
import pytest

from mcmtop_amd import gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "data", "fortran_formats.json")
W = gpu.OUT_WORDS


def _fixture():
    d = json.load(open(GOLD))
    vals = [struct.unpack("<d", bytes.fromhex(h))[0] for h in d["values_hex"]]
    return vals, d


def _read(path):
    with open(path) as f:
        return f.read().split("\n")


def _block(lines, header, n):
    k = lines.index(header)
    return lines[k + 1:k + 1 + n]


@pytest.fixture(scope="module")
def lib():
    return gpu.load_library()


def test_number_formats_equal_fortran(lib, tmp_path):
    """F15.8 (coordinates, velocities), E24.5 (scalars) and E24.5 of the
    single-precision von Mises stress, value by value, equal flang's output:
    rounding, signed zeros, three-digit exponents, overflow asterisks,
    subnormals, Inf and NaN."""
    vals, d = _fixture()
    n = len(vals)
    f = np.zeros((n, W))
    for a in range(3):
        f[:, a] = vals
        f[:, 3 + a] = vals
    f[:, 6] = vals
    vm_ok = [math.isfinite(v) and 1e-150 < v < 1e150 for v in vals]
    f[:, 7] = [v if ok else 0.0 for v, ok in zip(vals, vm_ok)]
    path = str(tmp_path / "f.vtk")
    gpu.vtk_write(path, f)
    lines = _read(path)
    pts = _block(lines, "POINTS %d double" % n, n)
    vel = _block(lines, "VECTORS Velocities double", n)
    for k in range(n):
        assert pts[k] == d["f15_8"][k] * 3, (vals[k], pts[k])
        assert vel[k] == d["f15_8"][k] * 3, (vals[k], vel[k])
    pres = _block(lines, "LOOKUP_TABLE default", n)   # the first scalar block is the pressure
    assert lines[lines.index("LOOKUP_TABLE default") - 1] == "SCALARS Pressure double 1"
    for k in range(n):
        assert pres[k] == d["e24_5"][k], (vals[k], pres[k])
    vm = _block(lines, "SCALARS VM_Stress double 1", n + 1)[1:]
    for k in range(n):
        if vm_ok[k]:
            assert vm[k] == d["e24_5_real4"][k], (vals[k], vm[k])


def test_integer_fields_follow_fortran_and_counts_do_not_overflow(lib, tmp_path):
    """Cell lines are "1 " + I6 and counts I5 as in the reference where they
    fit (flang's I6 / I5 right-justify with blanks); 123,456 particles, where
    the reference's I5 count fields would print asterisks, get full counts."""
    vals, d = _fixture()
    for k, v in enumerate(vals):
        if not math.isfinite(v):
            continue
        i6 = int(min(max(v * 1.0e6, -2.0e9), 2.0e9))
        i5 = int(min(max(v * 1.0e3, -2.0e9), 2.0e9))
        if len(str(i6)) <= 6:
            assert d["i6"][k] == "1 " + str(i6).rjust(6)
        else:
            assert d["i6"][k] == "1 ******"
        if len(str(i5)) <= 5:
            assert d["i5"][k] == str(i5).rjust(5)
    n = 123456
    path = str(tmp_path / "big.vtk")
    gpu.vtk_write(path, np.zeros((n, W)))
    lines = _read(path)
    assert lines[:5] == ["# vtk DataFile Version 3.0", "TEST", "ASCII", "DATASET UNSTRUCTURED_GRID",
                         "POINTS %d double" % n]
    k = lines.index("CELLS %d %d" % (n, 2 * n))
    assert lines[k + 1] == "1      0" and lines[k + n] == "1 123455"
    assert lines[k + n + 1] == "CELL_TYPES 123456"
    assert lines[k + 2 * n + 2] == "POINT_DATA 123456"
    assert "*" not in "".join(lines[k:k + 2 * n + 3])


def test_material_file_layout(lib, tmp_path):
    """WriteVTKmat: only the particles of the material, renumbered cells, the
    same blocks, then a PTYPE block (I5)."""
    rng = np.random.default_rng(5)
    n = 50
    f = rng.uniform(-1, 1, (n, W))
    mat = np.where(np.arange(n) % 3 == 0, 2, 1).astype(np.int32)
    p_all, p2 = str(tmp_path / "all.vtk"), str(tmp_path / "m2.vtk")
    gpu.vtk_write(p_all, f, mat, 0)
    gpu.vtk_write(p2, f, mat, 2)
    la, l2 = _read(p_all), _read(p2)
    m = int((mat == 2).sum())
    sel = np.nonzero(mat == 2)[0]
    heads = ["VECTORS Velocities double"] + ["SCALARS %s double 1" % s for s in (
        "Pressure", "Stress_X", "Stress_Y", "Stress_Z", "Shear_Stress_XY", "Shear_Stress_XZ",
        "Shear_Stress_YZ", "VM_Stress")]
    assert [h for h in la if h in heads] == heads and [h for h in l2 if h in heads] == heads
    assert _block(l2, "POINTS %d double" % m, m) == [_block(la, "POINTS %d double" % n, n)[i] for i in sel]
    assert _block(l2, "CELLS %d %d" % (m, 2 * m), m) == ["1 " + str(k).rjust(6) for k in range(m)]
    for h in heads[1:]:
        a = _block(la, h, n + 1)[1:]
        b = _block(l2, h, m + 1)[1:]
        assert b == [a[i] for i in sel], h
    assert "SCALARS PTYPE double 1" not in la
    assert _block(l2, "SCALARS PTYPE double 1", m + 1)[1:] == ["    0"] * m
    with pytest.raises(gpu.McmGpuError):
        gpu.vtk_write(str(tmp_path / "bad.vtk"), f, None, 1)
