"""Output sync and the VTK writer on the GPU path: the packed output fields
of a run equal the oracle's state, and the VTK file written from the device
equals, byte for byte, the file written from the oracle's state (the plain
f3dm3 deck, whose state is bit-identical to the reference's)."""
import os

import numpy as np
import pytest

from mcmtop_amd import gpu
from oracle import Oracle

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")
# column-major 3x3 in the dumps: sigma 11 22 33 12 13 23 at 0 4 8 3 6 7
SIG = (0, 4, 8, 3, 6, 7)


def _records(dump):
    n = dump.np
    f = np.zeros((n, gpu.OUT_WORDS))
    f[:, 0:3] = dump.f("x")
    f[:, 3:6] = dump.f("v")
    f[:, 6] = dump.f("p")
    f[:, 7:13] = dump.f("sigma")[:, SIG]
    f[:, 13] = dump.f("e")
    f[:, 14] = dump.f("efps")
    f[:, 15] = dump.f("rho")
    f[:, 16] = dump.f("h")
    return f


def test_output_fields_and_vtk_equal_the_oracle(tmp_path):
    path = os.path.join(DATA, "plain_f3dm3.bin")
    o = Oracle(path)
    o.initial()
    o.run(40)
    ref = _records(o.state(40, 10))
    with gpu.Solver(path) as s:
        s.initial()
        s.step(40)
        got = s.output_fields(8)
        got4 = s.output_fields(4)
        s.write_vtk(str(tmp_path / "gpu.vtk"))
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got4, ref.astype(np.float32))
    gpu.vtk_write(str(tmp_path / "oracle.vtk"), ref)
    assert open(tmp_path / "gpu.vtk", "rb").read() == open(tmp_path / "oracle.vtk", "rb").read()


def test_material_files_of_a_two_material_run(tmp_path):
    path = os.path.join(DATA, "two_materials_contact.bin")
    with gpu.Solver(path) as s:
        s.initial()
        s.step(10)
        f = s.output_fields(8)
        mat = s.scalars().nummat
        assert mat == 2
        for m in (0, 1, 2):
            s.write_vtk(str(tmp_path / ("m%d.vtk" % m)), m)
        # a material without particles gives a file without points, as the
        # reference's WriteVTKmat(2) does for a one-material run
        s.write_vtk(str(tmp_path / "m3.vtk"), 3)
        with pytest.raises(gpu.McmGpuError):
            s.write_vtk(str(tmp_path / "mneg.vtk"), -1)
    assert "POINTS 0 double" in open(tmp_path / "m3.vtk").read()
    from mcmtop_amd import deck
    ids = deck.Deck.read_bin(path).mat
    for m in (0, 1, 2):
        gpu.vtk_write(str(tmp_path / ("h%d.vtk" % m)), f, ids, m)
        assert open(tmp_path / ("m%d.vtk" % m), "rb").read() == open(tmp_path / ("h%d.vtk" % m), "rb").read()


def test_material_file_of_a_cap_run_writes_the_modes(tmp_path):
    """WriteVTKmat's PTYPE block (mcm_VTK.f90) holds par%ptype, the cap
    model's mode of each particle's last update: the oracle's modes after
    the same steps."""
    path = os.path.join(DATA, "cap_soil_impact.bin")
    o = Oracle(path)
    o.initial()
    o.run(20)
    want = o.state(20, 10).ptype
    with gpu.Solver(path) as s:
        s.initial()
        s.step(20)
        s.write_vtk(str(tmp_path / "m1.vtk"), 1)
    lines = open(tmp_path / "m1.vtk").read().splitlines()
    k = lines.index("SCALARS PTYPE double 1")
    got = np.array([int(v) for v in lines[k + 2:k + 2 + len(want)]])
    assert len(got) == len(want) and len(set(want.tolist())) > 1
    np.testing.assert_array_equal(got, want)
