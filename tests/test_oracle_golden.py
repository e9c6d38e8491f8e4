"""The CPU oracle against the reference's own results.

tests/golden/data/digests.json holds SHA-256 digests of state dumps written
by the flang build of the reference (tests/golden/make_golden.py) for the
decks of tests/golden/cases.py, covering both discretisations, models
1/3/4/9/10 (4: the cap model, with its capa and mode), EOS 1/4/13/28/41,
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


def test_decks_reach_their_branches():
    """The decks added for branches no other deck runs do run them, on the
    oracle state that the digests pin to the reference's."""
    def final(name):
        o = Oracle(os.path.join(DATA, name + ".bin"))
        assert o.initial() == 0
        assert o.run(DIGESTS[name]["nsteps"]) == 0
        return o.state(DIGESTS[name]["nsteps"], 10)

    # f3dm3 failure (mcm_f3dm3.f90:137-150): particles past efps_fail keep
    # only the mean stress, the others keep a deviatoric part
    d = final("f3dm3_efps_failure")
    sig, efps = d.f("sigma").reshape(-1, 3, 3), d.f("efps")
    failed = efps > 0.05
    assert 0 < failed.sum() < len(efps)
    off = np.abs(sig[:, [0, 0, 1], [1, 2, 2]]).max(axis=1)
    spread = np.ptp(np.diagonal(sig, axis1=1, axis2=2), axis=1)
    assert np.all(off[failed] == 0.0) and np.all(spread[failed] == 0.0)
    assert np.all(spread[~failed] > 0.0)
    # f3dm9 with mu = 0 (mcm_f3dm9.f90:30-33): no deviatoric stress anywhere
    d = final("fluid_f3dm9_mu0_eos1")
    assert np.all(d.f("s") == 0.0)
    assert np.abs(d.f("rod")).max() > 0.0
    # eos4 sound speed with c^2 <= 0 (mcm_eos4c.f90:76-89): the reference
    # reports taking that branch
    assert DIGESTS["eos4_negative_c2"]["warnings"]["complex_sound_speed"] > 0
    assert all(DIGESTS[k]["warnings"]["complex_sound_speed"] == 0 for k in DIGESTS if k != "eos4_negative_c2")


def test_cap_decks_reach_the_cap_models_modes():
    """The cap model (mcm_f3dm25.f90:229-261 selects one of eleven modes per
    particle and step): over the cap decks, every step, the modes taken by
    the vector (four-step) and the iterated returns.  Modes 6 and 7 need
    0 < capa < 1e-7 Fe and mode 8 needs capa < 0, which no deck reaches: the
    update keeps capa >= 0 from its zero start."""
    seen = {}
    for name in [n for n in cases.CASES if n.startswith("cap_")]:
        o = Oracle(os.path.join(DATA, name + ".bin"))
        assert o.initial() == 0
        np_ = DIGESTS[name]["np"]
        modes = set()
        for _ in range(DIGESTS[name]["nsteps"]):
            assert o.step() == 0
            modes |= set(o.capa(np_)[1].tolist())
        seen[name] = modes
    vector = seen["cap_soil_impact"] | seen["cap_soil_stretch_corr"]
    iterated = seen["cap_rock_iterated_impact"] | seen["cap_soil_iterated_impact"]
    assert {1, 2, 4, 9, 10, 11} <= vector, vector
    assert {1, 2, 3, 5, 9, 10, 11} <= iterated, iterated
    assert not (vector | iterated) & {0, 6, 7, 8}


@pytest.mark.parametrize("name", sorted(json.load(open(os.path.join(DATA, "dropin.json")))))
def test_reference_loop_matches_the_oracle(name, tmp_path):
    """The reference's own unchanged time loop (oracle/_ref/ref_solution,
    where the reference build exists) writes the files recorded in
    dropin.json and ends in the oracle's state after the last step's
    mcm_calctotale, bit for bit."""
    import shutil
    import subprocess
    exe = os.path.join(os.path.dirname(DATA), "..", "..", "oracle", "_ref", "ref_solution")
    if not os.path.exists(exe):
        pytest.skip("reference build not present")
    rec = json.load(open(os.path.join(DATA, "dropin.json")))[name]
    shutil.copy(os.path.join(DATA, name + ".bin"), tmp_path / "deck.bin")
    out = subprocess.run([exe, "deck.bin"] + rec["args"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    for f, digest in rec["files"].items():
        assert hashlib.sha256((tmp_path / f).read_bytes()).hexdigest() == digest, f
    o = Oracle(os.path.join(DATA, name + ".bin"))
    assert o.initial() == 0
    assert o.run(rec["nsteps"] - 1) == 0
    for st in STAGES[:8]:
        assert o.stage(st) == 0
    want, got = o.state(rec["nsteps"], 8), Dump(str(tmp_path / "final.bin"))
    assert np.array_equal(got.rows, want.rows)
    assert got.ptime == want.ptime and got.kinetice == want.kinetice and got.internale == want.internale


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_reference_card_routines_read_the_text_deck_alike(name, tmp_path):
    """The reference's own card routines (mcm_matin, mcm_nodein,
    mcm_inputvel, mcm_inp_symmetry_v2, mcm_contact_in, mcm_baseaccel_in,
    compiled in place; only mcm_control_v2 restated) read the deck's text
    form (Deck.write_mcm) into exactly the state the binary deck gives: every
    dump of the case equals the digests of tests/golden/data/digests.json.
    With test_textdeck.py (the library's reader turns the same text into the
    binary deck's bytes) this pins the text reader to the reference's."""
    import subprocess
    probe = os.path.join(os.path.dirname(DATA), "..", "..", "oracle", "_ref", "ref_probe")
    if not os.path.exists(probe):
        pytest.skip("reference build not present")
    rec = DIGESTS[name]
    txt = str(tmp_path / "deck.mcm")
    cases.build(name).write_mcm(txt)
    steps = sorted(set(rec["stage_steps"]) | {rec["nsteps"]})
    out = subprocess.run([probe, txt, str(tmp_path), str(rec["nsteps"]), "1"] + [str(s) for s in steps],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    for key, want in rec["dumps"].items():
        step, stage = (int(v) for v in key.split("_"))
        raw = open(tmp_path / ("dump_%06d_%02d.bin" % (step, stage)), "rb").read()
        assert hashlib.sha256(raw).hexdigest() == want["sha256"], key
