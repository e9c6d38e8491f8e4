#!/usr/bin/env python3
"""Write tests/golden/data/dropin.json: what the reference's own time loop
writes on a few decks, for the drop-in test of the GPU wrappers.

Runs only where the reference build exists (oracle/build_ref.sh, which needs
/root/reference): oracle/_ref/ref_solution is the reference's unchanged
mcm_initial and mcm_solution with its own step routines and writers, driven
by oracle/dropin_main.f90.  For each deck the run controls are chosen from
the oracle's step times so that the loop ends after the deck's step count
(mcm_solution.f90:101 stops on time): endtime halfway between the last two
steps' start times, a state output (EnSight, state_opt 1) about every
third of the run and a time-history line about every ninth, three time-history
particles.  Recorded: those controls and the SHA-256 digest of every file the
reference's writers produced (its log aside, which also echoes host-only
messages), plus the time-history table itself for diagnosis.

Usage: python tests/golden/make_dropin.py
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from oracle import Oracle  # noqa: E402

DATA = os.path.join(HERE, "data")
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_solution")
# decks without symmetry or periodic planes: with ghosts the reference's own
# loop stops on its ghost-list capacity (mcm_ll_neighbours.f90:83, :301-309
# give the ghost lists no growth margin), which the probe works around and
# the unchanged loop cannot
DECKS = {"two_materials_contact": 30, "corr_collide_f3dm10_eos4_hopt1": 30, "f3dm3_efps_failure": 30,
         "f3dm10_spall2_failure": 30}
SKIP = ("run.log", "final.bin", "deck.bin")


def controls(deck_path: str, nsteps: int) -> list:
    o = Oracle(deck_path)
    o.initial()
    pt = []
    for k in range(nsteps):
        o.step()
        pt.append(o.state(k + 1, 10).ptime)
    endtime = 0.5 * (pt[nsteps - 3] + pt[nsteps - 2])
    stpl = pt[nsteps - 1] / 3.5
    np_ = o.state(0, 0).np
    return [repr(endtime), repr(stpl), repr(stpl / 3.0), "1", "1", str(np_ // 2), str(np_)]


def outputs(tmp: str) -> dict:
    return {f: hashlib.sha256(open(os.path.join(tmp, f), "rb").read()).hexdigest()
            for f in sorted(os.listdir(tmp)) if f not in SKIP and not f.startswith("fort.")}


def main():
    if not os.path.exists(EXE):
        print("make_dropin: %s not built (oracle/build_ref.sh needs the reference sources)" % EXE)
        return 0
    out = {}
    for name, nsteps in DECKS.items():
        src = os.path.join(DATA, name + ".bin")
        args = controls(src, nsteps)
        with tempfile.TemporaryDirectory() as tmp:
            shutil.copy(src, os.path.join(tmp, "deck.bin"))
            run = subprocess.run([EXE, "deck.bin"] + args, cwd=tmp, capture_output=True, text=True, timeout=600)
            if run.returncode != 0:
                raise SystemExit("%s: reference loop failed\n%s%s" % (name, run.stdout[-2000:], run.stderr[-2000:]))
            files = outputs(tmp)
            thist = open(os.path.join(tmp, "run_thist.csv")).read()
        out[name] = {"nsteps": nsteps, "args": args, "files": files, "thist": thist}
        print("%-32s %d steps, %d output files" % (name, nsteps, len(files)))
    with open(os.path.join(DATA, "dropin.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
