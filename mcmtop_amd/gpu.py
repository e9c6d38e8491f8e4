"""Python front end of the C ABI in include/mcm_gpu.h (libmcm_gpu.so).

The reference's time step is a set of argument-less subroutines called from
mcm_solution (MCM/mcm_solution.f90:41-116).  ``Solver`` exposes the same
stages under the same names (``neighbours``, ``strain``, ``update_h``,
``constitutive``, ``momentum``, ``boundcond``, ``calctotale``,
``update_velocity``, ``move``) plus the fused ``step(n)`` fast path.  Every
call goes through the C ABI; there is no CPU fallback: if the library is
missing or no GPU is visible the constructor raises.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcm_gpu.so")

ERRORS = {0: "ok", -1: "bad argument", -2: "io error", -3: "HIP error", -4: "capacity exceeded",
          -5: "unsupported model", -6: "no HIP device"}

F_X, F_V, F_A, F_RHO, F_H, F_P, F_E, F_C, F_EFPS, F_SIGMA, F_VABS, F_MASS = range(12)

# C-ABI symbols declared in include/mcm_gpu.h
SYMBOLS = (
    "mcm_gpu_create_from_deck", "mcm_gpu_destroy", "mcm_gpu_last_error", "mcm_gpu_initial",
    "mcm_gpu_neighbours", "mcm_gpu_strain", "mcm_gpu_update_h", "mcm_gpu_constitutive",
    "mcm_gpu_dt_update", "mcm_gpu_momentum", "mcm_gpu_boundcond", "mcm_gpu_calctotale",
    "mcm_gpu_update_velocity", "mcm_gpu_move", "mcm_gpu_calccritts", "mcm_gpu_step",
    "mcm_gpu_get_scalars", "mcm_gpu_download", "mcm_gpu_download_neighbours", "mcm_gpu_dump",
    "mcm_gpu_time_steps", "mcm_gpu_time_kernels", "mcm_gpu_convert_deck", "mcm_gpu_create_from_memory",
    "mcm_gpu_batch_create", "mcm_gpu_batch_add_deck", "mcm_gpu_batch_add_memory", "mcm_gpu_batch_size",
    "mcm_gpu_batch_run", "mcm_gpu_batch_last_error", "mcm_gpu_batch_initial", "mcm_gpu_batch_step",
    "mcm_gpu_batch_time_steps", "mcm_gpu_batch_destroy", "mcm_gpu_output_fields", "mcm_vtk_write",
    "mcm_gpu_write_vtk", "mcm_gpu_rep_cont", "mcm_gpu_batch_tune",
    "mcm_gpu_set_material_params", "mcm_gpu_get_material_params", "mcm_gpu_set_initial_velocity",
    "mcm_gpu_adj_forward", "mcm_gpu_adj_backward", "mcm_gpu_adj_last_error", "mcm_gpu_adj_destroy",
    "mcm_gpu_set_time", "mcm_gpu_download_state", "mcm_gpu_upload_state", "mcm_gpu_get_energies",
    "mcm_gpu_calctotale_noalloc", "mcm_gpu_cap_state",
)

STATE_WORDS = 86   # row of the per-particle state (include/mcm_gpu.h, the dump row layout)

N_PARAMS = 4   # per material: model 10 strinput(1:3) (G, sigma_y, Eh) and eosinput(2) (eos4 S1); model 1 E, nu

OUT_WORDS = 17   # packed output record: x(3), v(3), p, sigma 11 22 33 12 13 23, e, efps, rho, h


class McmGpuError(RuntimeError):
    pass


class Options(ctypes.Structure):
    _fields_ = [("nbr_capacity", ctypes.c_int32), ("ghost_capacity", ctypes.c_int32),
                ("device", ctypes.c_int32), ("corrected_one_pass", ctypes.c_int32),
                ("pair_mapping", ctypes.c_int32), ("nbr_skin_milli", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 2)]


class Scalars(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("dtold", ctypes.c_double), ("ptime", ctypes.c_double),
                ("critts", ctypes.c_double), ("hmax", ctypes.c_double), ("kinetice", ctypes.c_double),
                ("internale", ctypes.c_double), ("totale", ctypes.c_double),
                ("coord_min", ctypes.c_double * 3), ("coord_max", ctypes.c_double * 3),
                ("np", ctypes.c_int32), ("ngp", ctypes.c_int32), ("max_count", ctypes.c_int32),
                ("timestep", ctypes.c_int32), ("error_flags", ctypes.c_int32), ("nummat", ctypes.c_int32),
                ("skin_builds", ctypes.c_int32), ("reserved", ctypes.c_int32 * 1)]


def _options(nbr_capacity, ghost_capacity, device, corrected_one_pass, pair_mapping, nbr_skin) -> Options:
    return Options(nbr_capacity, ghost_capacity, device, int(bool(corrected_one_pass)), int(pair_mapping),
                   int(round(nbr_skin * 1000)))


_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libmcm_gpu.so and declare its signatures; raises if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise McmGpuError("libmcm_gpu.so not built (%s); run __graft_entry__.build()" % path)
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.mcm_gpu_create_from_deck.argtypes = [ctypes.c_char_p, ctypes.POINTER(Options), ctypes.POINTER(vp)]
    lib.mcm_gpu_destroy.argtypes = [vp]
    lib.mcm_gpu_destroy.restype = None
    lib.mcm_gpu_last_error.argtypes = [vp]
    lib.mcm_gpu_last_error.restype = ctypes.c_char_p
    for name in ("initial", "neighbours", "strain", "update_h", "constitutive", "dt_update", "rep_cont", "momentum",
                 "boundcond", "calctotale", "update_velocity", "move"):
        getattr(lib, "mcm_gpu_" + name).argtypes = [vp]
    lib.mcm_gpu_calccritts.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    lib.mcm_gpu_step.argtypes = [vp, ctypes.c_int32]
    lib.mcm_gpu_get_scalars.argtypes = [vp, ctypes.POINTER(Scalars)]
    lib.mcm_gpu_download.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_int64]
    i32p = ctypes.POINTER(ctypes.c_int32)
    lib.mcm_gpu_download_neighbours.argtypes = [vp, i32p, i32p, i32p, ctypes.c_int64, i32p]
    lib.mcm_gpu_dump.argtypes = [vp, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32]
    lib.mcm_gpu_time_steps.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    lib.mcm_gpu_time_kernels.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]
    lib.mcm_gpu_convert_deck.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32]
    lib.mcm_gpu_create_from_memory.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(Options),
                                               ctypes.POINTER(vp)]
    lib.mcm_gpu_batch_create.argtypes = [ctypes.POINTER(Options), ctypes.POINTER(vp)]
    lib.mcm_gpu_batch_add_deck.argtypes = [vp, ctypes.c_char_p]
    lib.mcm_gpu_batch_add_memory.argtypes = [vp, ctypes.c_char_p, ctypes.c_int64]
    lib.mcm_gpu_batch_size.argtypes = [vp]
    lib.mcm_gpu_batch_size.restype = ctypes.c_int32
    lib.mcm_gpu_batch_run.argtypes = [vp, ctypes.c_int32]
    lib.mcm_gpu_batch_run.restype = vp
    lib.mcm_gpu_batch_last_error.argtypes = [vp]
    lib.mcm_gpu_batch_last_error.restype = ctypes.c_char_p
    lib.mcm_gpu_batch_initial.argtypes = [vp]
    lib.mcm_gpu_batch_step.argtypes = [vp, ctypes.c_int32]
    lib.mcm_gpu_batch_time_steps.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    lib.mcm_gpu_output_fields.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int64]
    lib.mcm_vtk_write.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    lib.mcm_gpu_write_vtk.argtypes = [vp, ctypes.c_char_p, ctypes.c_int32]
    lib.mcm_gpu_batch_tune.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    lib.mcm_gpu_batch_destroy.argtypes = [vp]
    lib.mcm_gpu_batch_destroy.restype = None
    dp = ctypes.POINTER(ctypes.c_double)
    lib.mcm_gpu_set_material_params.argtypes = [vp, ctypes.c_int32, dp]
    lib.mcm_gpu_get_material_params.argtypes = [vp, ctypes.c_int32, dp]
    lib.mcm_gpu_set_initial_velocity.argtypes = [vp, dp, ctypes.c_int64]
    lib.mcm_gpu_adj_forward.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, i32p, ctypes.c_int32, dp,
                                        ctypes.POINTER(vp)]
    lib.mcm_gpu_adj_backward.argtypes = [vp, dp, dp, dp, dp, dp]
    lib.mcm_gpu_adj_last_error.argtypes = [vp]
    lib.mcm_gpu_adj_last_error.restype = ctypes.c_char_p
    lib.mcm_gpu_adj_destroy.argtypes = [vp]
    lib.mcm_gpu_adj_destroy.restype = None
    lib.mcm_gpu_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.mcm_gpu_cap_state.argtypes = [vp, dp, i32p, ctypes.c_int64, ctypes.c_int32]
    lib.mcm_gpu_download_state.argtypes = [vp, dp, ctypes.c_int64, ctypes.POINTER(Scalars)]
    lib.mcm_gpu_upload_state.argtypes = [vp, dp, ctypes.c_int64, ctypes.POINTER(Scalars)]
    lib.mcm_gpu_get_energies.argtypes = [vp, dp, ctypes.c_int32]
    lib.mcm_gpu_calctotale_noalloc.argtypes = [vp] * 7 + [ctypes.c_int64, ctypes.c_int32, vp, vp]
    special = ("mcm_gpu_destroy", "mcm_gpu_last_error", "mcm_gpu_batch_size", "mcm_gpu_batch_run",
               "mcm_gpu_batch_last_error", "mcm_gpu_batch_destroy", "mcm_gpu_adj_last_error",
               "mcm_gpu_adj_destroy")
    for name in SYMBOLS:
        if name not in special:
            getattr(lib, name).restype = ctypes.c_int
    _LIB = lib
    return lib


def convert_deck(in_path: str, out_path: str) -> None:
    """Text (reference version 2) or binary deck -> binary deck, values as
    written.  Runs on the host; no GPU needed."""
    lib = load_library()
    msg = ctypes.create_string_buffer(512)
    rc = lib.mcm_gpu_convert_deck(in_path.encode(), out_path.encode(), msg, 512)
    if rc != 0:
        raise McmGpuError("convert_deck failed: %s (%s)" % (ERRORS.get(rc, rc), msg.value.decode()))


def vtk_write(path: str, fields: np.ndarray, material_ids=None, material: int = 0) -> None:
    """The reference's VTK particle file from packed output records
    fields[np, OUT_WORDS] (host only, no GPU needed)."""
    lib = load_library()
    f = np.ascontiguousarray(fields, dtype=np.float64).reshape(-1, OUT_WORDS)
    ids = None
    if material_ids is not None:
        ids = np.ascontiguousarray(material_ids, dtype=np.int32)
    rc = lib.mcm_vtk_write(path.encode(), len(f), f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if ids is not None else None,
                           int(material))
    if rc != 0:
        raise McmGpuError("vtk_write failed: %s" % ERRORS.get(rc, rc))


def _deck_bytes(deck) -> bytes:
    """Binary deck bytes of a Deck object or of bytes as given."""
    if isinstance(deck, (bytes, bytearray)):
        return bytes(deck)
    return deck.to_bytes()


class Solver:
    """One MCM run on one GPU, created from an input deck: a path to the
    reference's version-2 text deck or to the binary deck of
    mcmtop_amd/deck.py, or binary deck bytes / a Deck object."""

    def __init__(self, deck, nbr_capacity: int = 0, ghost_capacity: int = 0, device: int = 0,
                 corrected_one_pass: bool = False, pair_mapping: int = 0, nbr_skin: float = 0.0):
        """pair_mapping: 0 one thread per particle, 1 one wavefront per particle (same results).
        nbr_skin: neighbour skin width as a fraction of the largest h (0: lists searched every step
        in the reference's order; > 0: the same neighbour sets from skin lists, in another order)."""
        self._lib = load_library()
        self._owned = True
        opt = _options(nbr_capacity, ghost_capacity, device, corrected_one_pass, pair_mapping, nbr_skin)
        h = ctypes.c_void_p()
        if isinstance(deck, str):
            rc = self._lib.mcm_gpu_create_from_deck(deck.encode(), ctypes.byref(opt), ctypes.byref(h))
        else:
            data = _deck_bytes(deck)
            rc = self._lib.mcm_gpu_create_from_memory(data, len(data), ctypes.byref(opt), ctypes.byref(h))
        self._h = h
        self._check(rc, "create")

    @classmethod
    def _run_of(cls, lib, handle: int) -> "Solver":
        """A run owned by a batch (closing it does nothing)."""
        s = cls.__new__(cls)
        s._lib = lib
        s._owned = False
        s._h = ctypes.c_void_p(handle)
        return s

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self._lib.mcm_gpu_last_error(self._h).decode() if self._h else ""
            raise McmGpuError("%s failed: %s (%s)" % (what, ERRORS.get(rc, rc), msg))

    def close(self):
        spare = getattr(self, "_adj_spare", None)
        if spare:
            self._lib.mcm_gpu_adj_destroy(spare)
            self._adj_spare = None
        # live recordings are released with the run (mcm_gpu_destroy)
        for rec in list(getattr(self, "_recs", ())):
            rec._h = None
        if getattr(self, "_h", None):
            if self._owned:
                self._lib.mcm_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # stages, named as the reference subroutines they replace
    def initial(self):
        self._check(self._lib.mcm_gpu_initial(self._h), "initial")

    def _stage(self, name):
        self._check(getattr(self._lib, "mcm_gpu_" + name)(self._h), name)

    def neighbours(self): self._stage("neighbours")
    def strain(self): self._stage("strain")
    def update_h(self): self._stage("update_h")
    def constitutive(self): self._stage("constitutive")
    def dt_update(self): self._stage("dt_update")
    def rep_cont(self): self._stage("rep_cont")
    def momentum(self): self._stage("momentum")
    def boundcond(self): self._stage("boundcond")
    def calctotale(self): self._stage("calctotale")
    def update_velocity(self): self._stage("update_velocity")
    def move(self): self._stage("move")

    def stage(self, name: str):
        self._stage(name)

    def calccritts(self) -> float:
        out = ctypes.c_double()
        self._check(self._lib.mcm_gpu_calccritts(self._h, ctypes.byref(out)), "calccritts")
        return out.value

    def step_granular(self):
        """One step through the stage entry points, in mcm_solution's order."""
        for name in ("neighbours", "strain", "update_h", "constitutive", "dt_update", "rep_cont", "momentum",
                     "boundcond", "calctotale", "update_velocity", "move"):
            self._stage(name)

    def energies(self) -> np.ndarray:
        """(nummat, 5): ke, ie, mom(3) per material from the last calctotale or fused step."""
        nm = self.scalars().nummat
        out = np.zeros(5 * nm)
        self._check(self._lib.mcm_gpu_get_energies(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 5 * nm),
                    "get_energies")
        return out.reshape(nm, 5)

    def set_time(self, dt: float, dtold: float, ptime: float):
        """The host's dt, dtold and ptime (mcm_solution.f90:58-62 on the host)."""
        self._check(self._lib.mcm_gpu_set_time(self._h, float(dt), float(dtold), float(ptime)), "set_time")

    def download_state(self):
        """(rows[np, STATE_WORDS], Scalars): the state as the host's particle records hold it."""
        n = self.scalars().np
        rows = np.zeros((n, STATE_WORDS))
        s = Scalars()
        self._check(self._lib.mcm_gpu_download_state(self._h, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                     n, ctypes.byref(s)), "download_state")
        return rows, s

    def upload_state(self, rows, sc: Scalars):
        r = np.ascontiguousarray(rows, dtype=np.float64)
        self._check(self._lib.mcm_gpu_upload_state(self._h, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                   r.shape[0], ctypes.byref(sc)), "upload_state")

    def cap_state(self):
        """(capa[np], ptype[np]): the cap model's state (zeros without a cap-model material)."""
        n = self.scalars().np
        capa, ptype = np.zeros(n), np.zeros(n, np.int32)
        self._check(self._lib.mcm_gpu_cap_state(self._h, capa.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                ptype.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n, 0),
                    "cap_state")
        return capa, ptype

    def upload_cap_state(self, capa, ptype):
        c = np.ascontiguousarray(capa, dtype=np.float64)
        t = np.ascontiguousarray(ptype, dtype=np.int32)
        self._check(self._lib.mcm_gpu_cap_state(self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(c), 1),
                    "cap_state")

    def step(self, n: int = 1):
        """n fused steps (device-side dt, captured hipGraph)."""
        self._check(self._lib.mcm_gpu_step(self._h, int(n)), "step")

    def scalars(self) -> Scalars:
        s = Scalars()
        self._check(self._lib.mcm_gpu_get_scalars(self._h, ctypes.byref(s)), "get_scalars")
        return s

    def field(self, code: int) -> np.ndarray:
        n = self.scalars().np
        width = 3 if code in (F_X, F_V, F_A) else (9 if code == F_SIGMA else 1)
        out = np.zeros(n * width)
        self._check(self._lib.mcm_gpu_download(self._h, code, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               out.size), "download")
        return out.reshape(n, width) if width > 1 else out

    def neighbour_lists(self):
        """(counts, ghost_counts, ids[np, cap]) with 0-based ids, ghosts as np + ghost index."""
        n = self.scalars().np
        cap = ctypes.c_int32()
        counts = np.zeros(n, np.int32)
        gcounts = np.zeros(n, np.int32)
        # query capacity with an empty call first
        i32p = ctypes.POINTER(ctypes.c_int32)
        rc = self._lib.mcm_gpu_download_neighbours(self._h, counts.ctypes.data_as(i32p), gcounts.ctypes.data_as(i32p),
                                                   counts.ctypes.data_as(i32p), 0, ctypes.byref(cap))
        ids = np.full((n, cap.value), -1, np.int32)
        rc = self._lib.mcm_gpu_download_neighbours(self._h, counts.ctypes.data_as(i32p), gcounts.ctypes.data_as(i32p),
                                                   ids.ctypes.data_as(i32p), ids.size, ctypes.byref(cap))
        self._check(rc, "download_neighbours")
        return counts, gcounts, ids

    def output_fields(self, precision: int = 8) -> np.ndarray:
        """Packed output records [np, OUT_WORDS] in float64 (precision 8) or
        float32 (4, converted on the device), one transfer."""
        n = self.scalars().np
        out = np.zeros((n, OUT_WORDS), dtype=np.float64 if precision == 8 else np.float32)
        self._check(self._lib.mcm_gpu_output_fields(self._h, int(precision), out.ctypes.data, out.nbytes),
                    "output_fields")
        return out

    def write_vtk(self, path: str, material: int = 0):
        self._check(self._lib.mcm_gpu_write_vtk(self._h, path.encode(), int(material)), "write_vtk")

    def dump(self, path: str, step: int = 0, stage: int = 0):
        self._check(self._lib.mcm_gpu_dump(self._h, path.encode(), step, stage), "dump")

    def time_steps(self, n: int) -> np.ndarray:
        """Average ms per step of: neighbour build, strain-pass kernel,
        constitutive update, momentum kernel, whole step, and the list build
        kernel alone (part of the neighbour build), with HIP events on the
        run's stream."""
        ms = (ctypes.c_double * 6)()
        self._check(self._lib.mcm_gpu_time_kernels(self._h, int(n), ms, 6), "time_kernels")
        return np.array(ms[:])

    # ---- differentiation (plain SPH, material model 10)
    def set_material_params(self, mat: int, params) -> None:
        """The 4 parameters of material mat (0-based): model 10 G, sigma_y,
        Eh (strinput(1:3)) and S1 (eosinput(2)); model 1 E, nu (then 2
        unused).  Derived constants follow; takes effect at the next
        initial()."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        assert p.size == N_PARAMS
        self._check(self._lib.mcm_gpu_set_material_params(self._h, int(mat), p.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))), "set_material_params")

    def material_params(self, mat: int = 0) -> np.ndarray:
        p = np.zeros(N_PARAMS)
        self._check(self._lib.mcm_gpu_get_material_params(self._h, int(mat), p.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))), "get_material_params")
        return p

    def set_initial_velocity(self, v0) -> None:
        """The deck's initial velocities, (np, 3); takes effect at the next initial()."""
        v = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1)
        self._check(self._lib.mcm_gpu_set_initial_velocity(self._h, v.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)), v.size), "set_initial_velocity")

    def release_recordings(self) -> None:
        """Free the device buffers kept from the last recording."""
        spare = getattr(self, "_adj_spare", None)
        if spare:
            self._lib.mcm_gpu_adj_destroy(spare)
            self._adj_spare = None

    def record(self, nsteps: int, ckpt_every: int = 50, sample=()) -> "Recording":
        """nsteps fused steps from the current state (normally right after
        initial()), with a state checkpoint every ckpt_every steps and the
        positions of the sampled particles after every step, for a later
        reverse sweep (Recording.backward)."""
        smp = np.ascontiguousarray(sample, dtype=np.int32)
        hist = np.zeros((int(nsteps), len(smp), 3))
        # a closed recording of this run is reused (its buffers, when the shape matches)
        h = getattr(self, "_adj_spare", None) or ctypes.c_void_p()
        self._adj_spare = None
        i32p = ctypes.POINTER(ctypes.c_int32)
        rc = self._lib.mcm_gpu_adj_forward(self._h, int(nsteps), int(ckpt_every), smp.ctypes.data_as(i32p),
                                           len(smp), hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           ctypes.byref(h))
        if rc != 0:
            msg = self._lib.mcm_gpu_adj_last_error(h).decode() if h else \
                self._lib.mcm_gpu_last_error(self._h).decode()
            if h:
                self._lib.mcm_gpu_adj_destroy(h)
            raise McmGpuError("record failed: %s (%s)" % (ERRORS.get(rc, rc), msg))
        rec = Recording(self, h, int(nsteps), len(smp), hist)
        if not hasattr(self, "_recs"):
            self._recs = weakref.WeakSet()
        self._recs.add(rec)
        return rec


class Recording:
    """A recorded forward run (mcm_gpu_adj_forward): the sampled positions
    after every step in ``hist`` (nsteps, nsample, 3) and a reverse sweep."""

    def __init__(self, solver: Solver, handle, nsteps: int, nsample: int, hist: np.ndarray):
        self._s = solver
        self._lib = solver._lib
        self._h = handle
        self.nsteps = nsteps
        self.nsample = nsample
        self.hist = hist

    def backward(self, xbar=None, vbar=None, histbar=None):
        """Adjoint of the recorded run: dL/dx and dL/dv of the final state
        (np, 3) and dL/dhist (nsteps, nsample, 3), any of them None for zero.
        Returns (dL/dv0 (np, 3) of the deck's initial velocities,
        dL/dparams (nummat, 4))."""
        n = self._s.scalars().np
        nm = self._s.scalars().nummat
        dp = ctypes.POINTER(ctypes.c_double)

        def arr(a, shape):
            if a is None:
                return None, None
            b = np.ascontiguousarray(a, dtype=np.float64)
            assert b.shape == shape, (b.shape, shape)
            return b, b.ctypes.data_as(dp)

        xb, xp = arr(xbar, (n, 3))
        vb, vp = arr(vbar, (n, 3))
        hb, hp = arr(histbar, (self.nsteps, self.nsample, 3))
        v0 = np.zeros((n, 3))
        pb = np.zeros((nm, N_PARAMS))
        rc = self._lib.mcm_gpu_adj_backward(self._h, xp, vp, hp, v0.ctypes.data_as(dp), pb.ctypes.data_as(dp))
        if rc != 0:
            raise McmGpuError("backward failed: %s (%s)" % (ERRORS.get(rc, rc),
                                                            self._lib.mcm_gpu_adj_last_error(self._h).decode()))
        return v0, pb

    def close(self):
        """Release the recording: its buffers go back to the run for the next
        record() (or are freed when the run is closed)."""
        if getattr(self, "_h", None):
            s = self._s
            if getattr(s, "_h", None) and not getattr(s, "_adj_spare", None):
                s._adj_spare = self._h
            else:
                self._lib.mcm_gpu_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Batch:
    """Independent runs stepped together in one launch set (mcm_gpu_batch_*):
    each run has its own deck, particle count and scalars, and gives the
    same result as when stepped alone."""

    def __init__(self, decks=(), nbr_capacity: int = 0, ghost_capacity: int = 0, device: int = 0,
                 corrected_one_pass: bool = False, pair_mapping: int = 0, nbr_skin: float = 0.0):
        self._lib = load_library()
        opt = _options(nbr_capacity, ghost_capacity, device, corrected_one_pass, pair_mapping, nbr_skin)
        h = ctypes.c_void_p()
        rc = self._lib.mcm_gpu_batch_create(ctypes.byref(opt), ctypes.byref(h))
        self._h = h
        self._check(rc, "batch_create")
        for d in decks:
            self.add(d)

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self._lib.mcm_gpu_batch_last_error(self._h).decode() if self._h else ""
            raise McmGpuError("%s failed: %s (%s)" % (what, ERRORS.get(rc, rc), msg))

    def add(self, deck):
        """Add a run: a deck path, binary deck bytes or a Deck object."""
        if isinstance(deck, str):
            rc = self._lib.mcm_gpu_batch_add_deck(self._h, deck.encode())
        else:
            data = _deck_bytes(deck)
            rc = self._lib.mcm_gpu_batch_add_memory(self._h, data, len(data))
        self._check(rc, "batch_add")

    def __len__(self):
        return int(self._lib.mcm_gpu_batch_size(self._h))

    def run(self, k: int) -> Solver:
        """Run k (owned by the batch): scalars, fields, lists, dumps."""
        h = self._lib.mcm_gpu_batch_run(self._h, int(k))
        if not h:
            raise McmGpuError("batch has no run %d" % k)
        return Solver._run_of(self._lib, h)

    def initial(self):
        self._check(self._lib.mcm_gpu_batch_initial(self._h), "batch_initial")

    def step(self, n: int = 1):
        self._check(self._lib.mcm_gpu_batch_step(self._h, int(n)), "batch_step")

    def time_steps(self, n: int) -> np.ndarray:
        ms = (ctypes.c_double * 5)()
        self._check(self._lib.mcm_gpu_batch_time_steps(self._h, int(n), ms), "batch_time_steps")
        return np.array(ms[:])

    def tune(self, n: int = 2) -> int:
        """Time n steps with each pair mapping and keep the faster (0 thread,
        1 wavefront per particle); advances every run by 2 (n + 2) steps."""
        m = ctypes.c_int32()
        self._check(self._lib.mcm_gpu_batch_tune(self._h, int(n), ctypes.byref(m)), "batch_tune")
        return m.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mcm_gpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
