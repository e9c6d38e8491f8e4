"""Reader for the state dumps written by oracle/ref_probe.f90 and by the CPU
oracle (same layout).  TEST INFRASTRUCTURE ONLY.

Layout (little-endian): see the comment above ``probe_dump`` in ref_deck.f90;
``probe_dump_rows`` writes the same file up to the particle rows.
Particle rows hold NF=86 doubles; the column names are ``FIELDS`` below.
Neighbour ids are 1-based as in the reference (real ids into the particle
array, ghost ids into the ghost array).  Decks with a cap-model material
(model 4) end with 'CAP4', capa(np) and ptype(np).
"""
from __future__ import annotations

import numpy as np

DUMP_MAGIC = int.from_bytes(b"MCMD", "big")
ROWS_MAGIC = int.from_bytes(b"MCMR", "big")   # state rows only (probe_dump_rows in ref_deck.f90)
CAP_TAG = int.from_bytes(b"CAP4", "big")      # cap-model section: capa(np), ptype(np)
NF = 86

FIELDS = {
    "x": slice(0, 3), "v": slice(3, 6), "a": slice(6, 9),
    "rho": 9, "rhoold": 10, "h": 11, "hold": 12, "p": 13, "e": 14, "etry": 15,
    "einc": 16, "c": 17, "vabs": 18, "tracerod": 19, "pcut": 20, "p_cut": 21,
    "efps": 22, "fail": 23, "mindist": 24, "critts": 25, "mass": 26, "rho0": 27,
    "sigma": slice(28, 37), "s": slice(37, 46), "alfa": slice(46, 55),
    "q": slice(55, 64), "rod": slice(64, 73), "spin": slice(73, 82),
    "h0": 82, "xzero": slice(83, 86),
}
# 3x3 tensors are stored column-major (Fortran reshape): index 3*col + row.


class Dump:
    def __init__(self, path: str):
        buf = open(path, "rb").read()
        off = 0
        hdr = np.frombuffer(buf, "<i4", 6, off); off += 24
        if hdr[0] not in (DUMP_MAGIC, ROWS_MAGIC):
            raise ValueError("bad dump magic in %s" % path)
        self.step, self.stage, self.np, self.ngp, self.nummat = (int(v) for v in hdr[1:])
        sc = np.frombuffer(buf, "<f8", 8, off); off += 64
        (self.dt, self.dtold, self.ptime, self.critts, self.hmax,
         self.kinetice, self.internale, self.totale) = (float(v) for v in sc)
        g = np.frombuffer(buf, "<f8", 12, off); off += 96
        self.coord_maxmin = g[:6].reshape(3, 2)      # [dim][min,max]
        self.gridmin = g[6:9].copy()
        self.gridsize = g[9:12].copy()
        self.gridlim = np.frombuffer(buf, "<i4", 3, off).copy(); off += 12
        m = np.frombuffer(buf, "<f8", 5 * self.nummat, off).reshape(self.nummat, 5); off += 40 * self.nummat
        self.mat_ke = m[:, 0].copy(); self.mat_ie = m[:, 1].copy(); self.mat_mom = m[:, 2:5].copy()
        n = self.np
        self.rows = np.frombuffer(buf, "<f8", n * NF, off).reshape(n, NF).copy(); off += 8 * n * NF
        if hdr[0] == ROWS_MAGIC:
            self.nnbr = self.g_nnbr = self.llpointer = None
            return
        self.nnbr = np.frombuffer(buf, "<i4", n, off).copy(); off += 4 * n
        self.g_nnbr = np.frombuffer(buf, "<i4", n, off).copy(); off += 4 * n
        self.llpointer = np.frombuffer(buf, "<i4", n, off).copy(); off += 4 * n
        tot = int(self.nnbr.sum())
        flat = np.frombuffer(buf, "<i4", tot, off).copy(); off += 4 * tot
        self.nbr_off = np.concatenate([[0], np.cumsum(self.nnbr)])
        self.nbr = flat
        self.gnbr_off = np.concatenate([[0], np.cumsum(self.g_nnbr)])
        self.capa = self.ptype = None
        if off < len(buf) and int(np.frombuffer(buf, "<i4", 1, off)[0]) != CAP_TAG:
            gt = int(self.g_nnbr.sum())
            self.gnbr = np.frombuffer(buf, "<i4", gt, off).copy(); off += 4 * gt
            k = self.ngp
            self.g_par = np.frombuffer(buf, "<i4", k, off).copy(); off += 4 * k
            self.g_mat = np.frombuffer(buf, "<i4", k, off).copy(); off += 4 * k
            gr = np.frombuffer(buf, "<f8", 8 * k, off).reshape(k, 8).copy(); off += 64 * k
            self.g_x = gr[:, 0:3]; self.g_v = gr[:, 3:6]; self.g_mass = gr[:, 6]; self.g_h = gr[:, 7]
            self.g_rho = np.frombuffer(buf, "<f8", k, off).copy(); off += 8 * k
            self.g_hold = np.frombuffer(buf, "<f8", k, off).copy(); off += 8 * k
        else:
            self.gnbr = np.zeros(0, np.int32)
            self.g_par = np.zeros(0, np.int32)
        if off < len(buf) and int(np.frombuffer(buf, "<i4", 1, off)[0]) == CAP_TAG:
            off += 4
            self.capa = np.frombuffer(buf, "<f8", n, off).copy(); off += 8 * n
            self.ptype = np.frombuffer(buf, "<i4", n, off).copy(); off += 4 * n

    def f(self, name: str) -> np.ndarray:
        return self.rows[:, FIELDS[name]]

    def neighbours(self, i: int) -> np.ndarray:
        """Real neighbour ids (1-based) of particle i (0-based), in list order."""
        return self.nbr[self.nbr_off[i]:self.nbr_off[i + 1]]

    def ghost_neighbours(self, i: int) -> np.ndarray:
        return self.gnbr[self.gnbr_off[i]:self.gnbr_off[i + 1]]
