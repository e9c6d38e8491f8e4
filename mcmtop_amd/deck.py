"""Taylor-bar deck generator and binary deck I/O.

A deck is the complete input of one MCM run: control options, material cards,
particle positions / velocities / materials / boundary codes and the symmetry
planes.  It is written two ways from the same numbers:

* ``write_mcm``: the reference's version-2 fixed-column text deck
  (title ``(a78,i2)`` mcm_getinput.f90:42,145; control cards
  mcm_control_v2.f90:88-123,279-284; material cards mcm_matin.f90:41-203;
  node cards ``(i8,i5,3e20.0,i7)`` mcm_nodein.f90:42; velocity cards
  ``(i8,3e10.0)`` mcm_inputvel.f90:41-131; symmetry cards ``(6i10)/(6e10.0)``
  mcm_inp_symmetry_v2.f90:30-53).
* ``write_bin``: a little-endian binary image of the same values, read by the
  C-ABI host (``mcm_gpu_load_deck``), by the CPU oracle and by the reference
  probe driver.

Every real number is first rounded to what its fixed-width text field can
hold (``_fit``), so the text deck and the binary deck describe bit-identical
doubles: the text field is the definition of the input.
"""
from __future__ import annotations

import dataclasses
import io
import math
import struct
from typing import List, Optional

import numpy as np

DECK_MAGIC = int.from_bytes(b"MCM1", "big")
CONTACT_TAG = int.from_bytes(b"CONT", "big")   # optional penalty-contact section after the particles
OPTS_TAG = int.from_bytes(b"OPTS", "big")      # optional section: dynamic relaxation, base acceleration


def _fit(value: float, width: int) -> float:
    """Round ``value`` to the closest number printable in a ``width``-column
    Fortran ``E`` field and return the double that field reads back as."""
    return float(_fmt(value, width))


def _fmt(value: float, width: int) -> str:
    v = float(value)
    if v == 0.0:
        return ("0.0").rjust(width)
    # 'E' format with the largest mantissa that fits: sign + d + '.' + m + 'E+xx'
    for prec in range(width - 6, 0, -1):
        s = f"{v:.{prec}E}"
        if len(s) <= width:
            return s.rjust(width)
    raise ValueError(f"{value} does not fit in {width} columns")


@dataclasses.dataclass
class Material:
    """One material card set (mcm_matin.f90:41-203, mcm_database.f90:235-257).

    ``card`` holds the 48 raw property-card values exactly as the reference
    reads them into ``strinput`` (models 1/3/4/9: field k of card j goes to
    5(j-1)+k; model 10: 8 per card); ``eoscard`` holds the raw ``eosinput``.
    """
    model: int
    eos: int
    rho: float
    av_q: float
    av_l: float
    mass: float
    h: float
    rho_min: float = 0.0
    rho_max: float = 0.0
    visc_type: int = 1
    card: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(48))
    eoscard: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(24))


@dataclasses.dataclass
class Deck:
    np_: int
    disctype: int
    h_opt: int
    tcrit_opt: int
    maxnbr: int
    tssf: float
    itss: float
    endtime: float
    materials: List[Material]
    x: np.ndarray          # (N, 3) float64
    v: np.ndarray          # (N, 3) float64
    mat: np.ndarray        # (N,) int32, 1-based material id
    dispbc: np.ndarray     # (N,) int32
    sym_flag: np.ndarray   # (6,) int32: xmin xmax ymin ymax zmin zmax (1 sym, 2 periodic)
    sym_x: np.ndarray      # (6,) float64 plane coordinates
    contacttype: int = 0
    krtype: int = 1
    massopt: int = 0
    init_h_opt: int = 0
    init_rhoe: int = 0
    title: str = "taylor bar"
    pmass: Optional[np.ndarray] = None
    ph: Optional[np.ndarray] = None
    prho: Optional[np.ndarray] = None
    pe: Optional[np.ndarray] = None
    veloc_opt: int = 0          # 1: XSPH velocity smoothing (mcm_update_velocity.f90:61-87)
    ncontmats: int = -1         # penalty contact (contacttype > 0): -1 = every material pair
    cont_opt: int = 0
    contact_k: Optional[List[float]] = None   # per material K and n of mcm_rep_cont
    contact_n: Optional[List[float]] = None
    drelax_scale: float = 0.0   # dynamic relaxation when 0.1 < scale < 1 (mcm_control_v2.f90:223-225)
    nthp: tuple = (0, 0, 0)     # base acceleration on in x, y, z where 1 (mcm_control_v2.f90:241-243)
    base_a: tuple = (0.0, 0.0, 0.0)

    @property
    def mirr_opt(self) -> int:
        return 1 if np.any(self.sym_flag != 0) else 0

    # ------------------------------------------------------------------ binary
    def write_bin(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(self.to_bytes())

    def to_bytes(self) -> bytes:
        """The binary deck (little endian): magic, 19 ints, 9 reals, per material
        3 ints + 7 reals + 48 property + 24 EOS reals, then x, v, material,
        dispbc, mass, h, rho, e per particle."""
        n = self.np_
        f = io.BytesIO()
        f.write(struct.pack("<i", DECK_MAGIC))
        ints = [n, len(self.materials), self.disctype, self.h_opt, self.tcrit_opt,
                self.contacttype, self.krtype, self.maxnbr, self.massopt,
                self.init_h_opt, self.init_rhoe, self.mirr_opt]
        ints += [int(s) for s in self.sym_flag] + [self.veloc_opt]
        f.write(struct.pack("<%di" % len(ints), *ints))
        f.write(struct.pack("<9d", self.tssf, self.itss, self.endtime, *[float(s) for s in self.sym_x]))
        for m in self.materials:
            f.write(struct.pack("<3i", m.model, m.eos, m.visc_type))
            f.write(struct.pack("<7d", m.rho, m.av_q, m.av_l, m.mass, m.h, m.rho_min, m.rho_max))
            f.write(np.asarray(m.card, dtype="<f8").tobytes())
            f.write(np.asarray(m.eoscard, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(self.x, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(self.v, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(self.mat, dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(self.dispbc, dtype="<i4").tobytes())
        for arr in (self.pmass, self.ph, self.prho, self.pe):
            a = np.zeros(n) if arr is None else np.asarray(arr, dtype="<f8")
            f.write(a.tobytes())
        if self.contacttype > 0:
            # penalty contact cards (mcm_contact_in): tag, ncontmats, cont_opt, K and n per material
            nm = len(self.materials)
            f.write(struct.pack("<3i", CONTACT_TAG, self.ncontmats, self.cont_opt))
            k = list(self.contact_k or [0.0] * nm)
            c = list(self.contact_n or [0.0] * nm)
            f.write(struct.pack("<%dd" % (2 * nm), *[v for pair in zip(k, c) for v in pair]))
        if self.drelax_scale != 0.0 or any(self.nthp):
            f.write(struct.pack("<4i", OPTS_TAG, *[int(t) for t in self.nthp]))
            f.write(struct.pack("<4d", self.drelax_scale, *[float(a) for a in self.base_a]))
        return f.getvalue()

    @staticmethod
    def read_bin(path: str) -> "Deck":
        with open(path, "rb") as f:
            buf = f.read()
        off = 0

        def take(fmt):
            nonlocal off
            vals = struct.unpack_from(fmt, buf, off)
            off += struct.calcsize(fmt)
            return vals

        (magic,) = take("<i")
        if magic != DECK_MAGIC:
            raise ValueError("not an MCM binary deck")
        ints = take("<19i")
        n, nummat, disctype, h_opt, tcrit, contact, krtype, maxnbr, massopt, ihopt, irhoe, mirr = ints[:12]
        sym_flag = np.array(ints[12:18], dtype=np.int32)
        reals = take("<9d")
        mats = []
        for _ in range(nummat):
            model, eos, visc = take("<3i")
            rho, avq, avl, mass, h, rmin, rmax = take("<7d")
            card = np.array(take("<48d"))
            eoscard = np.array(take("<24d"))
            mats.append(Material(model, eos, rho, avq, avl, mass, h, rmin, rmax, visc, card, eoscard))

        def arr(dtype, count):
            nonlocal off
            a = np.frombuffer(buf, dtype=dtype, count=count, offset=off).copy()
            off += a.nbytes
            return a

        x = arr("<f8", 3 * n).reshape(n, 3)
        v = arr("<f8", 3 * n).reshape(n, 3)
        mat = arr("<i4", n)
        bc = arr("<i4", n)
        pm, ph, pr, pe = (arr("<f8", n) for _ in range(4))
        dk = Deck(n, disctype, h_opt, tcrit, maxnbr, reals[0], reals[1], reals[2], mats, x, v,
                  mat, bc, sym_flag, np.array(reals[3:9]), contact, krtype, massopt, ihopt,
                  irhoe, pmass=pm, ph=ph, prho=pr, pe=pe, veloc_opt=ints[18])
        if off + 12 <= len(buf) and struct.unpack_from("<i", buf, off)[0] == CONTACT_TAG:
            _, dk.ncontmats, dk.cont_opt = take("<3i")
            kn = take("<%dd" % (2 * nummat))
            dk.contact_k, dk.contact_n = list(kn[0::2]), list(kn[1::2])
        if off + 16 <= len(buf) and struct.unpack_from("<i", buf, off)[0] == OPTS_TAG:
            dk.nthp = tuple(take("<4i")[1:])
            vals = take("<4d")
            dk.drelax_scale, dk.base_a = vals[0], tuple(vals[1:])
        return dk

    # -------------------------------------------------------------------- text
    def write_mcm(self, path: str) -> None:
        """Reference version-2 text deck with the same numbers as ``write_bin``."""
        n = self.np_
        lines = []
        lines.append(self.title[:78].ljust(78) + "%2d" % 2)
        lines.append("%5d%5d%10d%10d%10d%10d%10d" % (3, self.disctype, len(self.materials), n, n, 0, 0))
        lines.append(_fmt(self.endtime, 10) + _fmt(self.tssf, 10) + _fmt(self.itss, 10) + _fmt(self.drelax_scale, 10))
        lines.append(_fmt(1.0e9, 10) + "%5d" % 0 + _fmt(1.0e9, 10) + "%10d%5d%10d%5d%5d" % (0, 0, 100, 0, 0))
        lines.append("%5d%5d%5d%5d%5d" % (1, 1, self.massopt, self.init_h_opt, self.init_rhoe))
        lines.append("%5d%5d%5d%5d%5d%5d%5d%5d" % (self.contacttype, self.tcrit_opt, self.veloc_opt, self.mirr_opt,
                                                    0, *[int(t) for t in self.nthp]))
        lines.append("%5d%5d%5d" % (self.h_opt, self.krtype, self.maxnbr))
        lines.append("")
        for k, m in enumerate(self.materials, start=1):
            lines.append("%5d%5d" % (k, m.model) + _fmt(m.rho, 10) + "%5d%5d" % (m.eos, 0) + _fmt(0.0, 10)
                         + "%5d" % m.visc_type + _fmt(m.av_q, 10) + _fmt(m.av_l, 10))
            lines.append(_fmt(m.mass, 10) + _fmt(m.h, 10) + _fmt(m.rho_min, 10) + _fmt(m.rho_max, 10))
            lines.append(("material %d" % k).ljust(72))
            if m.model == 10:
                for j in range(6):
                    lines.append("".join(_fmt(m.card[8 * j + i], 10) for i in range(8)))
            else:
                for j in range(6):
                    lines.append("".join(_fmt(m.card[5 * j + i], 10) for i in range(5)))
            if m.eos != 0:
                lines.append(("eos %d" % m.eos).ljust(72))
                if m.eos == 1:
                    lines.append("".join(_fmt(m.eoscard[i], 10) for i in range(8)))
                    lines.append(_fmt(m.eoscard[8], 10))
                elif m.eos in (4, 41):
                    lines.append("".join(_fmt(m.eoscard[i], 10) for i in range(8)))
                elif m.eos == 13:
                    lines.append("".join(_fmt(m.eoscard[i], 10) for i in range(5)))
                elif m.eos == 28:
                    lines.append("".join(_fmt(m.eoscard[i], 10) for i in range(2)))
        for i in range(n):
            lines.append("%8d%5d" % (i + 1, self.dispbc[i]) + "".join(_fmt(c, 20) for c in self.x[i])
                         + "%7d" % self.mat[i])
        if self.massopt or self.init_h_opt or self.init_rhoe:
            for i in range(n):
                vals = [self.pmass[i] if self.pmass is not None else 0.0,
                        self.ph[i] if self.ph is not None else 0.0,
                        self.prho[i] if self.prho is not None else 0.0,
                        self.pe[i] if self.pe is not None else 0.0]
                lines.append("%8d" % (i + 1) + "".join(_fmt(c, 20) for c in vals))
        for i in range(n):
            lines.append("%8d" % (i + 1) + "".join(_fmt(c, 10) for c in self.v[i]))
        if self.mirr_opt:
            lines.append("".join("%10d" % s for s in self.sym_flag))
            lines.append("".join(_fmt(s, 10) for s in self.sym_x))
        if self.contacttype > 0:
            lines.append("%5d%5d" % (self.ncontmats, self.cont_opt))
            if self.ncontmats == -1:
                nm = len(self.materials)
                for k, c in zip(self.contact_k or [0.0] * nm, self.contact_n or [0.0] * nm):
                    lines.append(_fmt(k, 10) + _fmt(c, 10))
        if any(int(t) == 1 for t in self.nthp):
            lines.append("".join(_fmt(a, 20) for a in self.base_a))
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")


# ---------------------------------------------------------------- generators
def bar_lattice(dx: float, radius: float, length: float) -> np.ndarray:
    """Simple-cubic lattice nodes at ((i+1/2)dx, (j+1/2)dx, (k+1/2)dx) inside the
    cylinder x^2+y^2 <= r^2, 0 < z < L, in C order (k slowest, then j, then i)."""
    nr = int(math.ceil(radius / dx)) + 1
    ij = (np.arange(-nr, nr) + 0.5) * dx
    xx, yy = np.meshgrid(ij, ij, indexing="xy")       # rows = j (y), cols = i (x)
    keep = xx ** 2 + yy ** 2 <= radius ** 2
    xy = np.stack([xx[keep], yy[keep]], axis=1)
    nz = int(math.floor(length / dx + 1e-9))
    if (nz + 0.5) * dx > length + 1e-12:
        nz -= 0
    zs = (np.arange(nz) + 0.5) * dx
    zs = zs[zs < length]
    pts = np.empty((len(zs) * len(xy), 3))
    for k, z in enumerate(zs):
        sl = slice(k * len(xy), (k + 1) * len(xy))
        pts[sl, 0:2] = xy
        pts[sl, 2] = z
    return pts


def _material_f3dm3(rho, E, nu, sy, Et, beta, mass, h, av_q, av_l) -> Material:
    card = np.zeros(48)
    card[0] = _fit(E, 10)      # strinput(1)  E        (card 1)
    card[5] = _fit(nu, 10)     # strinput(6)  nu       (card 2)
    card[10] = _fit(sy, 10)    # strinput(11) sigma_y  (card 3)
    card[15] = _fit(Et, 10)    # strinput(16) E_tan    (card 4)
    card[20] = _fit(beta, 10)  # strinput(21) beta     (card 5)
    return Material(3, 0, _fit(rho, 10), _fit(av_q, 10), _fit(av_l, 10), _fit(mass, 10),
                    _fit(h, 10), card=card)


def _material_f3dm10_eos4(rho, E, nu, sy, Et, mass, h, av_q, av_l, C0, S1, gamma0, a) -> Material:
    G = E / (2.0 * (1.0 + nu))
    Eh = E * Et / (E - Et)          # plastic (hardening) modulus, as mcm_sets3.f90:42 derives it
    card = np.zeros(48)
    card[0] = _fit(G, 10)           # strinput(1) shear modulus
    card[1] = _fit(sy, 10)          # strinput(2) yield stress
    card[2] = _fit(Eh, 10)          # strinput(3) plastic modulus
    eos = np.zeros(24)
    eos[0] = _fit(C0, 10)
    eos[1] = _fit(S1, 10)
    eos[4] = _fit(gamma0, 10)
    eos[5] = _fit(a, 10)
    return Material(10, 4, _fit(rho, 10), _fit(av_q, 10), _fit(av_l, 10), _fit(mass, 10),
                    _fit(h, 10), card=card, eoscard=eos)


def taylor_bar(dx=4.0e-4, radius=3.2e-3, length=25.4e-3, v0=190.0, seed=1,
               jitter=0.01, model="f3dm3", corrected=False, h_opt=0,
               E=117e9, nu=0.35, sy=400e6, Et=100e6, rho=8930.0,
               av_l=1.0, av_q=2.0, tssf=0.3, maxnbr=80, sym=True,
               C0=3940.0, S1=1.49, gamma0=2.0, a_eos=0.47, tcrit_opt=1, fit_text=True) -> Deck:
    """Taylor impact bar on a symmetry plane at z=0, impacting at -v0 along z.

    Defaults are the 12.8k-particle plain-SPH case: dx = 0.4 mm, r = 3.2 mm,
    L = 25.4 mm, h = 1.2 dx, f3dm3 (rho 8930, E 117 GPa, nu 0.35, sigma_y 400 MPa,
    E_tan 100 MPa, kinematic beta=0), nodal AV av_l=1 av_q=2, CFL 0.3, seeded
    U(+-0.01 dx) jitter from numpy default_rng(seed) in C order.

    fit_text: round every coordinate to what its 20-column text field holds,
    so that the text deck (write_mcm) carries exactly the same values; off
    for decks that only travel in binary form (large batches).
    """
    pts = bar_lattice(dx, radius, length)
    rng = np.random.default_rng(seed)
    pts = pts + rng.uniform(-jitter * dx, jitter * dx, size=pts.shape)
    if fit_text:
        pts = np.vectorize(lambda c: _fit(c, 20))(pts) if len(pts) < 200000 else _fit_array(pts, 20)
    n = len(pts)
    total_mass = rho * dx ** 3 * n
    h = 1.2 * dx
    if model == "f3dm3":
        mat = _material_f3dm3(rho, E, nu, sy, Et, 0.0, total_mass, h, av_q, av_l)
    elif model == "f3dm10":
        mat = _material_f3dm10_eos4(rho, E, nu, sy, Et, total_mass, h, av_q, av_l, C0, S1, gamma0, a_eos)
    else:
        raise ValueError(model)
    v = np.zeros((n, 3))
    v[:, 2] = -_fit(v0, 10)
    sym_flag = np.array([0, 0, 0, 0, 1 if sym else 0, 0], dtype=np.int32)
    return Deck(n, 1 if corrected else 0, h_opt, tcrit_opt, maxnbr, _fit(tssf, 10), 0.0, _fit(1.0, 10),
                [mat], pts, v, np.ones(n, dtype=np.int32), np.zeros(n, dtype=np.int32),
                sym_flag, np.zeros(6))


def batch_bars(count=256, seed=3, nmin=5000, nmax=200000, corrected=True, h_opt=1, dx_range=(0.05e-3, 0.2e-3)):
    """Independent Taylor bars of uneven size, every parameter from one
    numpy default_rng(seed): radius U[2, 5] mm, L/D U[3, 8], dx U[dx_range]
    (default [0.05, 0.2] mm), v0 U[100, 250] m/s, sigma_y U[200, 600] MPa,
    E U[70, 210] GPa, drawn in that order per candidate, kept when the
    particle count lies in [nmin, nmax] (rejection sampling); then a jitter
    seed per bar.  f3dm10 + eqos4 as the scaled bar (nu 0.35, E_h 100 MPa,
    copper Gruneisen).  Returns (decks, parameters)."""
    rng = np.random.default_rng(seed)
    decks, params = [], []
    while len(decks) < count:
        r = rng.uniform(2.0e-3, 5.0e-3)
        ld = rng.uniform(3.0, 8.0)
        dx = rng.uniform(*dx_range)
        v0 = rng.uniform(100.0, 250.0)
        sy = rng.uniform(200e6, 600e6)
        E = rng.uniform(70e9, 210e9)
        length = 2.0 * r * ld
        est = math.pi * r * r * length / dx ** 3
        if est < 0.7 * nmin or est > 1.4 * nmax:
            continue
        n = len(bar_lattice(dx, r, length))
        if n < nmin or n > nmax:
            continue
        jseed = int(rng.integers(0, 2 ** 31 - 1))
        decks.append(taylor_bar(dx=dx, radius=r, length=length, v0=v0, seed=jseed, model="f3dm10",
                                corrected=corrected, h_opt=h_opt, E=E, sy=sy, fit_text=False))
        params.append({"radius": r, "length": length, "dx": dx, "v0": v0, "sigma_y": sy, "E": E, "np": n,
                       "seed": jseed})
    return decks, params


def _fit_array(a: np.ndarray, width: int) -> np.ndarray:
    """Vectorised ``_fit`` for large arrays: 14 significant digits, which every
    value of a Taylor-bar coordinate field (|c| < 1, 20 columns) can hold."""
    flat = a.ravel()
    out = np.array([float(s) for s in np.char.mod("%.13E", flat)])
    return out.reshape(a.shape)


def config1(seed=1) -> Deck:
    """Plain SPH, 13,104 particles, f3dm3, 1000 steps in the benchmark plan."""
    return taylor_bar(seed=seed)


def config2(dx=7.3e-5, seed=1) -> Deck:
    """Normalised (corrected) SPH at scale: f3dm10 + eqos4 (the nearest form the
    reference accepts: it rejects f3dm3 with an EOS, mcm_matin.f90:72-79),
    variable h (h_opt=1)."""
    return taylor_bar(dx=dx, seed=seed, model="f3dm10", corrected=True, h_opt=1)


# ---------------------------------------------------- generic deck builders
def material(model: int, rho: float, mass: float, h: float, card=None, eos: int = 0, eoscard=None,
             av_q: float = 2.0, av_l: float = 1.0, rho_max: float = 0.0) -> Material:
    """A material card set with every real fitted to its text field.  ``card``
    and ``eoscard`` map strinput/eosinput index (0-based) -> value."""
    c = np.zeros(48)
    for k, val in (card or {}).items():
        c[k] = _fit(val, 10)
    e = np.zeros(24)
    for k, val in (eoscard or {}).items():
        e[k] = _fit(val, 10)
    return Material(model, eos, _fit(rho, 10), _fit(av_q, 10), _fit(av_l, 10), _fit(mass, 10), _fit(h, 10),
                    0.0, _fit(rho_max, 10) if rho_max else 0.0, card=c, eoscard=e)


def f3dm1(E, nu):
    """model 1 (elastic) property cards: E (card 1), nu (card 2)"""
    return {0: E, 5: nu}


def f3dm3(E, nu, sy, Et, beta=0.0, fail_flag=0, efps_fail=0.0):
    return {0: E, 5: nu, 10: sy, 15: Et, 20: beta, 25: fail_flag, 26: efps_fail}


def f3dm9(pcut, mu, mat_id=1):
    """model 9 (fluid): the reference reads cm(1:2, mat) (mcm_f3dm9.f90:27-28)"""
    b = 5 * (mat_id - 1)
    return {0: pcut, b + 1: mu} if b == 0 else {0: pcut, b: pcut, b + 1: mu}


def f3dm10(G, sy, Eh, pcut=0.0, a1=0.0, a2=0.0, spall=0.0, epf=0.0):
    return {0: G, 1: sy, 2: Eh, 3: pcut, 4: a1, 5: a2, 6: spall, 7: epf}


def f3dm25(K, G, alpha, theta, gamma, beta, R, D, W, X0, Xint, ltype=1, tcut=0.0, vec=0.0, falfac=0.0):
    """model 4 (the cap model, mcm_f3dm25.f90) property cards as
    cm(5, 6): bulk and shear modulus, the failure surface
    Fe = alpha - gamma exp(-beta J1) + theta J1, cap aspect ratio R, the
    hardening law W (1 - exp(D (X0 - X))), the initial cap position Xint
    (mcm_sets4.f90 solves the initial cap parameter from it), soil (ltype 1)
    or rock (2), the tension cut-off tcut; vec != 0 selects the iterated
    (rather than the four-step) return of modes 10 and 11; falfac lowers
    alpha.  Stresses compression positive, in the deck's stress unit."""
    def cm(a, b):
        return (a - 1) + 5 * (b - 1)
    return {cm(1, 1): K, cm(1, 2): G, cm(1, 3): alpha, cm(1, 4): theta, cm(1, 5): gamma, cm(1, 6): beta,
            cm(2, 1): R, cm(2, 2): D, cm(2, 3): W, cm(2, 4): X0, cm(2, 5): Xint, cm(2, 6): float(ltype),
            cm(3, 1): tcut, cm(3, 5): vec, cm(4, 1): falfac}


def block(dx, nx, ny, nz, origin=(0.0, 0.0, 0.0)):
    """simple-cubic lattice block, nodes at origin + (i+1/2) dx"""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pts = np.stack([i.ravel(order="F"), j.ravel(order="F"), k.ravel(order="F")], axis=1).astype(float)
    return (pts + 0.5) * dx + np.asarray(origin)


def make_deck(pts, vel, mats, matid=None, dispbc=None, disctype=0, h_opt=0, tcrit_opt=1, tssf=0.3,
              sym_flag=(0, 0, 0, 0, 0, 0), sym_x=(0, 0, 0, 0, 0, 0), seed=1, jitter_dx=0.0,
              massopt=0, pmass=None, veloc_opt=0, contacttype=0, contact_k=None, contact_n=None,
              drelax_scale=0.0, nthp=(0, 0, 0), base_a=(0.0, 0.0, 0.0)) -> Deck:
    """Generic deck from points, velocities and materials (1-based matid);
    veloc_opt 1 turns on XSPH, contacttype > 0 penalty contact with K and n
    per material."""
    pts = np.asarray(pts, dtype=float)
    n = len(pts)
    if jitter_dx:
        pts = pts + np.random.default_rng(seed).uniform(-jitter_dx, jitter_dx, size=pts.shape)
    pts = _fit_array(pts, 20)
    v = np.asarray(vel, dtype=float) * np.ones((n, 3))
    v = np.vectorize(lambda c: _fit(c, 10))(v)
    matid = np.ones(n, np.int32) if matid is None else np.asarray(matid, np.int32)
    bc = np.zeros(n, np.int32) if dispbc is None else np.asarray(dispbc, np.int32)
    d = Deck(n, disctype, h_opt, tcrit_opt, 80, _fit(tssf, 10), 0.0, _fit(1.0, 10), list(mats), pts, v, matid,
             bc, np.asarray(sym_flag, np.int32), np.array([_fit(s, 10) if s else 0.0 for s in sym_x]),
             massopt=massopt, veloc_opt=veloc_opt)
    d.contacttype = contacttype
    d.drelax_scale = _fit(drelax_scale, 10) if drelax_scale else 0.0
    d.nthp = tuple(int(t) for t in nthp)
    d.base_a = tuple(_fit(a, 20) if a else 0.0 for a in base_a)
    if contacttype > 0:
        d.contact_k = [_fit(k, 10) for k in contact_k]
        d.contact_n = [_fit(c, 10) for c in contact_n]
    if massopt:
        d.pmass = np.array([_fit(m, 20) for m in pmass])
    return d
