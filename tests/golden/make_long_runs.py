#!/usr/bin/env python3
"""Write tests/golden/data/long_runs.json: the reference's state after 1000
steps of two Taylor bars, for the long-run GPU tests.

Runs only where the reference build exists (oracle/build_ref.sh): the flang
probe of the reference (oracle/_ref/ref_probe) runs each bar for 1000 steps
and dumps the end state; the CPU oracle runs it too and must give the same
dump bit for bit (checked here, once: about four minutes of CPU).  Recorded
per bar: the generator call, the dump's SHA-256, SHA-256 digests of the
state fields as little-endian float64 arrays (np x k, deck order) and their
L2 norms, the energies and ptime.  The GPU tests compare their own fields
with these digests (the fused path is bit-exact on both bars) and with the
norms for the 1e-7 L2 bar.

Usage: python tests/golden/make_long_runs.py
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from dumpio import Dump  # noqa: E402
from oracle import Oracle  # noqa: E402

DATA = os.path.join(HERE, "data")
PROBE = os.path.join(ROOT, "oracle", "_ref", "ref_probe")
FIELDS = ("x", "v", "sigma", "rho", "h", "p", "e", "efps", "c", "vabs")
# name -> (mcmtop_amd.deck generator and keyword arguments, steps)
RUNS = {
    # 13,104 particles: the plain, CPU-runnable Taylor bar (f3dm3, fixed h)
    "config1_plain_f3dm3": ("config1", {}, 1000),
    # 6,240 particles: the corrected, variable-h bar (f3dm10 + eqos4, h_opt 1),
    # at 150 m/s: at 190 m/s a particle leaves this bar's free surface near
    # step 800 and the reference's own cell grid (mcm_setuplist.f90:85-102)
    # grows until its allocation fails, in the reference and the oracle alike
    "corrected_f3dm10_eos4_hopt1_6k": ("taylor_bar", {"dx": 2.0e-4, "radius": 1.6e-3, "length": 6.0e-3,
                                                      "model": "f3dm10", "corrected": True, "h_opt": 1,
                                                      "v0": 150.0}, 1000),
}


def make(gen: str, kw: dict):
    from mcmtop_amd import deck
    return getattr(deck, gen)(**kw)


def field_record(d: Dump) -> dict:
    out = {}
    for k in FIELDS:
        a = np.ascontiguousarray(d.f(k), dtype="<f8")
        out[k] = {"sha256": hashlib.sha256(a.tobytes()).hexdigest(), "norm": float(np.linalg.norm(a))}
    dx = np.ascontiguousarray(d.f("x") - d.f("xzero"), dtype="<f8")
    out["x-x0"] = {"sha256": hashlib.sha256(dx.tobytes()).hexdigest(), "norm": float(np.linalg.norm(dx))}
    return out


def main():
    if not os.path.exists(PROBE):
        print("make_long_runs: %s not built" % PROBE)
        return 0
    out = {}
    for name, (gen, kw, nsteps) in RUNS.items():
        dk = make(gen, kw)
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "deck.bin")
            dk.write_bin(path)
            subprocess.run([PROBE, path, tmp, str(nsteps), "2", str(nsteps)], check=True, stdout=subprocess.DEVNULL)
            ref_path = os.path.join(tmp, "dump_%06d_10.bin" % nsteps)
            raw = open(ref_path, "rb").read()
            ref = Dump(ref_path)
            o = Oracle(path, ncap=200, gcap=200)
            o.initial()
            o.run(nsteps)
            orc = os.path.join(tmp, "oracle.bin")
            o.dump(orc, nsteps, 10)
            if open(orc, "rb").read() != raw:
                raise SystemExit("%s: the oracle's dump after %d steps differs from the reference's" % (name, nsteps))
        out[name] = {"generator": gen, "kwargs": kw, "nsteps": nsteps, "np": ref.np,
                     "dump_sha256": hashlib.sha256(raw).hexdigest(), "fields": field_record(ref),
                     "ptime": ref.ptime, "kinetice": ref.kinetice, "internale": ref.internale,
                     "oracle_equal": True}
        print("%-34s np=%d steps=%d ptime=%.6e ke=%.6e ie=%.6e (oracle equal)" % (
            name, ref.np, nsteps, ref.ptime, ref.kinetice, ref.internale))
    with open(os.path.join(DATA, "long_runs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
