"""The entry point that allocates nothing: mcm_gpu_calctotale_noalloc (the
energy totals of mcm_calctotale over particle arrays the caller holds, with
its scratch in a fixed device array of the library) against
mcm_gpu_calctotale of a run holding the same state, through the C ABI.  Two
calls in a row, the second with a smaller input (and another material
count), each bit for bit."""
import ctypes
import os

import numpy as np
import pytest

from mcmtop_amd import deck, gpu

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")


class _Hip:
    """Device buffers for the test's inputs and outputs, from the HIP runtime
    the library itself runs on (the test allocates; the entry point does not)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, n):
        out = np.zeros(n)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc_totals(lib, s, matid):
    sc = s.scalars()
    n, nm = sc.np, sc.nummat
    v = s.field(gpu.F_V).reshape(-1, 3)
    hip = _Hip()
    try:
        cols = [hip.put(c) for c in (s.field(gpu.F_MASS), s.field(gpu.F_VABS), s.field(gpu.F_E), v[:, 0], v[:, 1],
                                     v[:, 2])]
        mat = hip.put((np.asarray(matid) - 1).astype(np.int32))
        out = hip.put(np.full(5 * nm + 3, np.nan))
        rc = lib.mcm_gpu_calctotale_noalloc(*cols, mat, n, nm, out, None)
        assert rc == 0
        return hip.get(out, 5 * nm + 3)
    finally:
        hip.free()


def test_calctotale_without_allocation_equals_calctotale(tmp_path):
    lib = gpu.load_library()
    big = deck.taylor_bar(dx=2.54e-4, model="f3dm10", fit_text=False)       # 49,608 particles, one material
    small = os.path.join(DATA, "two_materials_contact.bin")                # 145 particles, two materials
    path = str(tmp_path / "bar.bin")
    big.write_bin(path)
    sizes = []
    for p, matid in ((path, big.mat), (small, deck.Deck.read_bin(small).mat)):
        with gpu.Solver(p) as s:
            s.initial()
            s.step(7)
            s.calctotale()
            want = s.energies()
            sc = s.scalars()
            got = _noalloc_totals(lib, s, matid)
        nm = want.shape[0]
        np.testing.assert_array_equal(got[:5 * nm].reshape(nm, 5), want)
        assert got[5 * nm] == sc.kinetice and got[5 * nm + 1] == sc.internale and got[5 * nm + 2] == sc.totale
        sizes.append(sc.np)
    assert sizes[1] < sizes[0]
    # inputs the scratch cannot hold are refused before anything is launched
    x = ctypes.c_void_p(1)
    assert lib.mcm_gpu_calctotale_noalloc(x, x, x, x, x, x, x, (1 << 22) + 1, 1, x, None) == -1   # MCM_GPU_ERR_ARG
