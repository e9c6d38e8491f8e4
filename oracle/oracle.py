"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never imports
this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

from dumpio import Dump

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

STAGES = ("neighbours", "strain", "update_h", "constitutive", "dt_update", "momentum",
          "boundcond", "calctotale", "update_velocity", "move")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def _lib():
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    lib.orc_load_deck.restype = ctypes.c_void_p
    lib.orc_load_deck.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.orc_free.argtypes = [ctypes.c_void_p]
    for name in ("orc_initial", "orc_step", "orc_rep_cont") + tuple("orc_" + s for s in STAGES):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
        getattr(lib, name).restype = ctypes.c_int
    lib.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_run.restype = ctypes.c_int
    lib.orc_dump.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.orc_dump.restype = ctypes.c_int
    dbl3 = ctypes.POINTER(ctypes.c_double)
    lib.orc_kernel_w.argtypes = [dbl3, dbl3, ctypes.c_double, ctypes.c_double]
    lib.orc_kernel_w.restype = ctypes.c_double
    lib.orc_kernel_gradw.argtypes = [dbl3, dbl3, ctypes.c_double, ctypes.c_double, dbl3]
    lib.orc_inverse3.argtypes = [dbl3, dbl3]
    lib.orc_get_x.argtypes = [ctypes.c_void_p, dbl3]
    lib.orc_get_x.restype = None
    lib.orc_get_rows.argtypes = [ctypes.c_void_p, dbl3]
    lib.orc_get_rows.restype = None
    i32p = ctypes.POINTER(ctypes.c_int32)
    lib.orc_get_lists.argtypes = [ctypes.c_void_p, i32p, i32p, i32p, ctypes.c_int64]
    lib.orc_get_lists.restype = ctypes.c_int
    lib.orc_get_scalars.argtypes = [ctypes.c_void_p, dbl3]
    lib.orc_get_scalars.restype = None
    lib.orc_get_capa.argtypes = [ctypes.c_void_p, dbl3, i32p]
    lib.orc_get_capa.restype = None
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


class Oracle:
    """One CPU-oracle run over a binary deck."""

    def __init__(self, deck_path: str, ncap: int = 160, gcap: int = 160):
        self._l = lib()
        self.h = self._l.orc_load_deck(deck_path.encode(), ncap, gcap)
        if not self.h:
            raise IOError("oracle could not load deck %s" % deck_path)
        self.step_no = 0
        # penalty contact (deck int 6, contacttype > 0): mcm_solution calls
        # mcm_rep_cont between the dt update and mcm_momentum
        with open(deck_path, "rb") as f:
            self.contact = int.from_bytes(f.read(28)[24:28], "little", signed=True) > 0

    def close(self):
        if self.h:
            self._l.orc_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def initial(self) -> int:
        return self._l.orc_initial(self.h)

    def stage(self, name: str) -> int:
        return getattr(self._l, "orc_" + name)(self.h)

    def step(self) -> int:
        self.step_no += 1
        return self._l.orc_step(self.h)

    def run(self, n: int) -> int:
        self.step_no += n
        return self._l.orc_run(self.h, n)

    def positions(self, np_: int):
        """(np, 3) positions of the current state."""
        import numpy
        out = numpy.zeros((np_, 3))
        self._l.orc_get_x(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def rows(self, np_: int):
        """(np, 86) particle rows of the current state (dump row layout), in memory."""
        import numpy
        out = numpy.zeros((np_, 86))
        self._l.orc_get_rows(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def lists(self, np_: int, cap: int):
        """(counts, gcounts, ids[np, cap]): the lists in order, 0-based, ghosts as np + ghost id."""
        import numpy
        c = numpy.zeros(np_, numpy.int32)
        g = numpy.zeros(np_, numpy.int32)
        ids = numpy.full((np_, cap), -1, numpy.int32)
        i32p = ctypes.POINTER(ctypes.c_int32)
        rc = self._l.orc_get_lists(self.h, c.ctypes.data_as(i32p), g.ctypes.data_as(i32p), ids.ctypes.data_as(i32p),
                                   cap)
        if rc != 0:
            raise ValueError("oracle lists longer than %d" % cap)
        return c, g, ids

    def capa(self, np_: int):
        """(capa, ptype) of the cap model (material model 4), per particle."""
        import numpy
        c = numpy.zeros(np_)
        t = numpy.zeros(np_, numpy.int32)
        self._l.orc_get_capa(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return c, t

    def scalars(self) -> dict:
        import numpy
        v = numpy.zeros(9)
        self._l.orc_get_scalars(self.h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        keys = ("dt", "dtold", "ptime", "critts", "hmax", "kinetice", "internale", "totale", "ngp")
        return dict(zip(keys, v.tolist()))

    def dump(self, path: str, step: int = 0, stage: int = 0) -> None:
        rc = self._l.orc_dump(self.h, path.encode(), step, stage)
        if rc != 0:
            raise IOError("oracle dump failed")

    def state(self, step: int = 0, stage: int = 0) -> Dump:
        fd, path = tempfile.mkstemp(suffix=".bin")
        os.close(fd)
        try:
            self.dump(path, step, stage)
            return Dump(path)
        finally:
            os.unlink(path)
