"""The CPU oracle against the reference's own results.

tests/golden/data/digests.json holds SHA-256 digests of state dumps written
by the flang build of the reference (tests/golden/make_golden.py) for eleven
decks covering both discretisations, models 1/3/9/10, EOS 1/4/13/28/41,
symmetry / periodic / corner ghosts, two materials in kernel contact and in
penalty contact, XSPH, boundary codes and the three timestep options.  The oracle writes the same dump
layout, so equal digests mean bit-identical state: positions, velocities,
stresses, energies, neighbour lists (in order) and ghosts.
"""
import hashlib
import json
import os
import tempfile

import numpy as np
import pytest

import cases
from dumpio import FIELDS, Dump
from oracle import STAGES, Oracle

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")
DIGESTS = json.load(open(os.path.join(DATA, "digests.json")))


def _check(o, rec, key, step, stage):
    fd, path = tempfile.mkstemp(suffix=".bin")
    os.close(fd)
    try:
        o.dump(path, step, stage)
        raw = open(path, "rb").read()
        want = rec["dumps"][key]
        if hashlib.sha256(raw).hexdigest() != want["sha256"]:
            d = Dump(path)
            bad = {k: (float(np.linalg.norm(d.f(k))), want["norms"][k]) for k in FIELDS
                   if float(np.linalg.norm(d.f(k))) != want["norms"][k]}
            pytest.fail("dump %s differs from the reference; fields (oracle, reference) norms: %s" % (key, bad))
    finally:
        os.unlink(path)


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_oracle_matches_reference_digests(name):
    rec = DIGESTS[name]
    o = Oracle(os.path.join(DATA, name + ".bin"))
    assert o.initial() == 0
    _check(o, rec, "0_0", 0, 0)
    nsteps = rec["nsteps"]
    for s in range(1, nsteps + 1):
        if s in rec["stage_steps"]:
            for k, st in enumerate(STAGES, start=1):
                assert o.stage(st) == 0
                if st == "dt_update" and o.contact:   # no dump falls between the two
                    assert o.stage("rep_cont") == 0
                _check(o, rec, "%d_%d" % (s, k), s, k)
        else:
            assert o.step() == 0
    _check(o, rec, "%d_10" % nsteps, nsteps, 10)


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_deck_generator_is_deterministic(name, tmp_path):
    p = tmp_path / "d.bin"
    cases.build(name).write_bin(str(p))
    assert p.read_bytes() == open(os.path.join(DATA, name + ".bin"), "rb").read()
