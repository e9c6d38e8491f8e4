#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/data from the reference itself.

Runs only where the reference sources are mounted (/root/reference): builds
the flang probe of the reference (oracle/build_ref.sh), runs it on each deck
of cases.py and records, for the initial state, every stage of the listed
steps and the end of the last step, the SHA-256 digest of the reference's
state dump together with per-field L2 norms.  The dumps themselves are not
kept; the CPU oracle writes byte-identical dumps, so the digests pin it bit
for bit, and the norms locate a mismatch.

Usage: python tests/golden/make_golden.py
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
sys.path.insert(0, HERE)

from dumpio import FIELDS, Dump  # noqa: E402
import cases  # noqa: E402

DATA = os.path.join(HERE, "data")


def digest(path: str) -> dict:
    raw = open(path, "rb").read()
    d = Dump(path)
    norms = {k: float(np.linalg.norm(d.f(k))) for k in FIELDS}
    return {"sha256": hashlib.sha256(raw).hexdigest(), "norms": norms, "ptime": d.ptime, "dt": d.dt,
            "kinetice": d.kinetice, "internale": d.internale, "ngp": d.ngp}


def main():
    if not os.path.isdir("/root/reference/MCM"):
        print("make_golden: reference sources not present; nothing to do")
        return 0
    subprocess.run([os.path.join(ROOT, "oracle", "build_ref.sh")], check=True)
    probe = os.path.join(ROOT, "oracle", "_ref", "ref_probe")
    os.makedirs(DATA, exist_ok=True)
    out = {}
    for name, (nsteps, stage_steps) in cases.CASES.items():
        dk = cases.build(name)
        deck_path = os.path.join(DATA, name + ".bin")
        dk.write_bin(deck_path)
        with tempfile.TemporaryDirectory() as tmp:
            steps = sorted(set(stage_steps) | {nsteps})
            run = subprocess.run([probe, deck_path, tmp, str(nsteps), "1"] + [str(s) for s in steps], check=True,
                                 stdout=subprocess.PIPE, text=True)
            rec = {"np": dk.np_, "nsteps": nsteps, "stage_steps": list(stage_steps), "dumps": {},
                   # the reference's own warnings: branches it reports taking
                   "warnings": {"complex_sound_speed": run.stdout.count("Complex speed of sound")}}
            rec["dumps"]["0_0"] = digest(os.path.join(tmp, "dump_000000_00.bin"))
            for s in steps:
                stages = range(1, 11) if s in stage_steps else (10,)
                for k in stages:
                    rec["dumps"]["%d_%d" % (s, k)] = digest(os.path.join(tmp, "dump_%06d_%02d.bin" % (s, k)))
        out[name] = rec
        last = rec["dumps"]["%d_10" % nsteps]
        print("%-32s N=%4d steps=%3d ngp=%4d ptime=%.4e ke=%.5e ie=%.5e" % (
            name, dk.np_, nsteps, last["ngp"], last["ptime"], last["kinetice"], last["internale"]))
    with open(os.path.join(DATA, "digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
