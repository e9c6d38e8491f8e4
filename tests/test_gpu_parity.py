"""Parity of the HIP path with the CPU oracle, through the C ABI.

Per-step fields must agree to 1e-10 norm-relative (the per-step tolerance
of the target); neighbour lists must be equal as sets, and are checked to
come out in the reference's order as well.  Long runs check x - x0, v and
sigma to 1e-7 L2 and the energy totals to 1e-9 after 1000 steps.
"""
import json
import os

import numpy as np
import pytest

import cases
from dumpio import FIELDS, Dump
from mcmtop_amd import deck, gpu
from oracle import STAGES, Oracle

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")
PER_STEP_TOL = 1e-10
STATE = ("x", "v", "rho", "h", "p", "e", "sigma", "alfa", "efps", "q", "c", "vabs", "pcut", "fail")


def rel(a, b):
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def field_errors(g: Dump, o: Dump, fields=None):
    out = {}
    for k in fields or FIELDS:
        if k in ("hold",) and fields is None:
            continue
        out[k] = rel(g.f(k), o.f(k))
    return out


def lists_equal_as_sets(g: Dump, o: Dump):
    if not (np.array_equal(g.nnbr, o.nnbr) and np.array_equal(g.g_nnbr, o.g_nnbr)):
        return False
    for i in range(o.np):
        if set(g.neighbours(i)) != set(o.neighbours(i)):
            return False
        if o.g_nnbr[i]:
            gp = set(g.g_par[g.ghost_neighbours(i) - 1])
            op = set(o.g_par[o.ghost_neighbours(i) - 1])
            if gp != op:
                return False
    return True


def check_cap(g: Dump, o: Dump, tol, where):
    """The cap model's state (decks with a model-4 material): the mode of
    each particle's last update equal, capa within tol."""
    if o.capa is None:
        assert g.capa is None, where
        return
    assert g.capa is not None, where
    assert np.array_equal(g.ptype, o.ptype), (where, "ptype", int((g.ptype != o.ptype).sum()))
    assert rel(g.capa, o.capa) <= tol, (where, "capa", rel(g.capa, o.capa))


def gdump(s: gpu.Solver, tmp_path, step, stage):
    p = str(tmp_path / ("g_%d_%d.bin" % (step, stage)))
    s.dump(p, step, stage)
    return Dump(p)


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_stage_by_stage(name, tmp_path):
    path = os.path.join(DATA, name + ".bin")
    o = Oracle(path)
    o.initial()
    with gpu.Solver(path) as s:
        s.initial()
        g0, o0 = gdump(s, tmp_path, 0, 0), o.state(0, 0)
        assert lists_equal_as_sets(g0, o0)
        errs = field_errors(g0, o0)
        assert max(errs.values()) <= PER_STEP_TOL, errs
        for step in (1, 2):
            for k, st in enumerate(STAGES, start=1):
                s.stage(st)
                o.stage(st)
                if st == "dt_update" and o.contact:   # mcm_solution.f90:64, no dump in between
                    s.stage("rep_cont")
                    o.stage("rep_cont")
                gd, od = gdump(s, tmp_path, step, k), o.state(step, k)
                assert lists_equal_as_sets(gd, od), (step, st)
                # same order as the reference walks its linked lists
                assert np.array_equal(gd.nbr, od.nbr) and np.array_equal(gd.gnbr, od.gnbr), (step, st)
                assert gd.ngp == od.ngp
                errs = field_errors(gd, od)
                worst = max(errs, key=errs.get)
                assert errs[worst] <= PER_STEP_TOL, (step, st, worst, errs[worst])
                check_cap(gd, od, PER_STEP_TOL, (step, st))
                for sc in ("dt", "dtold", "ptime", "critts", "kinetice", "internale"):
                    a, b = getattr(gd, sc), getattr(od, sc)
                    assert abs(a - b) <= PER_STEP_TOL * max(abs(b), 1e-300), (step, st, sc, a, b)


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_fused_steps(name, tmp_path):
    """The fused kernels with device-side dt over the whole case."""
    path = os.path.join(DATA, name + ".bin")
    nsteps = cases.CASES[name][0]
    o = Oracle(path)
    o.initial()
    o.run(nsteps)
    od = o.state(nsteps, 10)
    with gpu.Solver(path) as s:
        s.initial()
        s.step(nsteps)
        gd = gdump(s, tmp_path, nsteps, 10)
    assert lists_equal_as_sets(gd, od)
    errs = field_errors(gd, od, STATE)
    assert max(errs.values()) <= 1e-9, errs
    check_cap(gd, od, 1e-9, nsteps)
    assert gd.ptime == pytest.approx(od.ptime, rel=1e-12)
    assert gd.kinetice == pytest.approx(od.kinetice, rel=1e-9)
    assert gd.internale == pytest.approx(od.internale, rel=1e-9)


def test_plain_path_is_bit_exact(tmp_path):
    """Plain SPH with f3dm3 takes no reassociated sums: the GPU state equals
    the oracle's (hence the reference's) bit for bit."""
    path = os.path.join(DATA, "plain_f3dm3.bin")
    o = Oracle(path)
    o.initial()
    o.run(40)
    od = o.state(40, 10)
    with gpu.Solver(path) as s:
        s.initial()
        s.step(40)
        gd = gdump(s, tmp_path, 40, 10)
    for k in STATE:
        np.testing.assert_array_equal(gd.f(k), od.f(k), err_msg=k)
    assert gd.ptime == od.ptime


def test_runs_are_deterministic(tmp_path):
    path = os.path.join(DATA, "corr_f3dm10_eos4_hopt1.bin")
    outs = []
    for r in range(2):
        with gpu.Solver(path) as s:
            s.initial()
            s.step(25)
            outs.append(gdump(s, tmp_path, 25, r).rows.copy())
    np.testing.assert_array_equal(outs[0], outs[1])


def test_capacity_regrow_gives_the_same_lists(tmp_path):
    """Tiny neighbour and ghost capacities overflow on the device; the host
    grows them and repeats the search (the reference's regrow,
    mcm_ll_neighbours.f90:44-311, aborts instead).  Lists and results equal
    a large-capacity run."""
    path = os.path.join(DATA, "plain_f3dm3.bin")
    with gpu.Solver(path) as big:
        big.initial()
        big.step(12)
        ref = gdump(big, tmp_path, 12, 0)
    with gpu.Solver(path, nbr_capacity=4, ghost_capacity=8) as small:
        small.initial()
        small.step(12)
        got = gdump(small, tmp_path, 12, 1)
    assert lists_equal_as_sets(got, ref)
    np.testing.assert_array_equal(got.rows, ref.rows)


def test_momentum_is_conserved_on_the_gpu(tmp_path):
    path = os.path.join(DATA, "two_materials_contact.bin")
    with gpu.Solver(path) as s:
        s.initial()
        s.step(3)
        for st in ("neighbours", "strain", "update_h", "constitutive", "dt_update", "momentum"):
            s.stage(st)
        d = gdump(s, tmp_path, 4, 6)
    m = d.f("mass")[:, None]
    a = d.f("a")
    assert np.abs((m * a).sum(axis=0)).max() <= 1e-12 * (m * np.abs(a)).sum(axis=0).max()


LONG = json.load(open(os.path.join(DATA, "long_runs.json")))


@pytest.mark.parametrize("name", sorted(LONG))
def test_taylor_bar_1000_steps(name, tmp_path):
    """The 13,104-particle plain bar (f3dm3, the CPU-runnable configuration)
    and a 6,240-particle corrected bar (f3dm10 + eqos4, h_opt 1), 1000 fused
    steps, against the reference's own state after 1000 steps
    (tests/golden/data/long_runs.json, from tests/golden/make_long_runs.py,
    where the CPU oracle gave the same dump).  The fused path is bit-exact on
    both, so every field's digest must match; the L2 bar of 1e-7 on x - x0,
    v and sigma is then met with zero error, and the energy totals (summed
    over blocks in another order) agree to 1e-9."""
    import hashlib
    rec = LONG[name]
    dk = getattr(deck, rec["generator"])(**rec["kwargs"])
    path = str(tmp_path / "deck.bin")
    dk.write_bin(path)
    with gpu.Solver(path) as s:
        s.initial()
        s.step(rec["nsteps"])
        gd = gdump(s, tmp_path, rec["nsteps"], 10)
    assert gd.np == rec["np"]
    bad = {}
    for k, want in rec["fields"].items():
        a = gd.f("x") - gd.f("xzero") if k == "x-x0" else gd.f(k)
        a = np.ascontiguousarray(a, dtype="<f8")
        if hashlib.sha256(a.tobytes()).hexdigest() != want["sha256"]:
            bad[k] = abs(float(np.linalg.norm(a)) - want["norm"]) / want["norm"]
    assert not bad, "fields differ from the reference (relative norm differences, a lower bound of the L2 error): %s" % bad
    assert gd.ptime == rec["ptime"]
    assert gd.kinetice == pytest.approx(rec["kinetice"], rel=1e-9)
    assert gd.internale == pytest.approx(rec["internale"], rel=1e-9)


def _assert_same_state(a: Dump, b: Dump):
    """State fields bit for bit (the fused paths stage different diagnostics:
    the wavefront strain pass keeps rod and spin, the thread path does not)."""
    for k in STATE:
        np.testing.assert_array_equal(a.f(k), b.f(k), err_msg=k)


def _fused_dump(path, tmp_path, nsteps, tag, **opts):
    with gpu.Solver(path, **opts) as s:
        s.initial()
        s.step(nsteps)
        return gdump(s, tmp_path, nsteps, tag)


@pytest.mark.parametrize("name", sorted(cases.CASES))
def test_wave_and_thread_mappings_agree(name, tmp_path):
    """The wavefront-per-particle list build and pair kernels fold every sum
    in list order, as the thread-per-particle kernels do: same state bit for
    bit."""
    path = os.path.join(DATA, name + ".bin")
    n = cases.CASES[name][0]
    a = _fused_dump(path, tmp_path, n, 1, pair_mapping=1)
    b = _fused_dump(path, tmp_path, n, 2, pair_mapping=0)
    _assert_same_state(a, b)
    np.testing.assert_array_equal(a.nbr, b.nbr)
    np.testing.assert_array_equal(a.gnbr, b.gnbr)
    assert (a.ptime, a.kinetice, a.internale) == (b.ptime, b.kinetice, b.internale)


def _dense_deck(corrected):
    """Smoothing length 1.45 dx: 80-120 neighbours per particle, so lists
    spill into the second wavefront slot (entries 64..127)."""
    pts = deck.block(4.0e-4, 7, 7, 7)
    rho = 8930.0
    h = 1.45 * 4.0e-4
    if corrected:
        m = deck.material(10, rho, rho * 4.0e-4 ** 3 * len(pts), h, deck.f3dm10(45e9, 400e6, 100e6), eos=4,
                          eoscard={0: 3940.0, 1: 1.49, 4: 2.0, 5: 0.47})
    else:
        m = deck.material(3, rho, rho * 4.0e-4 ** 3 * len(pts), h, deck.f3dm3(117e9, 0.35, 400e6, 100e6))
    return deck.make_deck(pts, (0.0, 0.0, -190.0), [m], disctype=1 if corrected else 0, h_opt=1 if corrected else 0,
                          jitter_dx=0.01 * 4.0e-4, sym_flag=(0, 0, 0, 0, 1, 0))


@pytest.mark.parametrize("corrected", [False, True])
def test_long_lists_use_both_wave_slots(corrected, tmp_path):
    path = str(tmp_path / "dense.bin")
    _dense_deck(corrected).write_bin(path)
    o = Oracle(path)
    o.initial()
    o.run(15)
    od = o.state(15, 10)
    a = _fused_dump(path, tmp_path, 15, 1, pair_mapping=1)
    assert (a.nnbr + a.g_nnbr).max() > 64
    assert lists_equal_as_sets(a, od)
    errs = field_errors(a, od, STATE)
    assert max(errs.values()) <= 1e-9, errs
    # lists longer than 2 x 64 slots go to the thread-per-particle kernels
    b = _fused_dump(path, tmp_path, 15, 2, pair_mapping=1, nbr_capacity=200)
    _assert_same_state(a, b)
    c = _fused_dump(path, tmp_path, 15, 3, pair_mapping=0)
    _assert_same_state(a, c)


# decks whose step calls pow (eos28: (rho/rho0)^gamma; h_opt 2: h0 (rho0/rho)^(1/3); tcrit_opt 0:
# (m/rho)^(1/3) and ^(2/3); penalty contact: f^n) or the double exp (the cap model's yield surface
# exp(-beta capa)): the device's and the host libm's may differ in the last bit, so these are held
# to 1e-9 by test_fused_steps only
POW_DECKS = {"f3dm10_eos41_hopt2_tcrit0", "fluid_f3dm9_eos28_dispbc", "penalty_contact_f3dm1"} | \
    {n for n in cases.CASES if n.startswith("cap_")}


@pytest.mark.parametrize("name", sorted(set(cases.CASES) - POW_DECKS))
def test_fused_state_is_bit_exact(name, tmp_path):
    """Every per-particle sum is folded in the reference's list order and
    every quotient is correctly rounded, so the state after the whole case
    equals the oracle's bit for bit (energy totals, summed over blocks in a
    different order, are checked to 1e-9 in test_fused_steps)."""
    path = os.path.join(DATA, name + ".bin")
    nsteps = cases.CASES[name][0]
    o = Oracle(path)
    o.initial()
    o.run(nsteps)
    od = o.state(nsteps, 10)
    gd = _fused_dump(path, tmp_path, nsteps, 1)
    bad = [k for k in STATE if not np.array_equal(gd.f(k), od.f(k))]
    assert not bad, {k: rel(gd.f(k), od.f(k)) for k in bad}
    assert gd.ptime == od.ptime


@pytest.mark.parametrize("name", ["corr_f3dm10_eos4_hopt1", "quarter_f3dm1_xyz_sym_tcrit2", "two_materials_contact",
                                  "penalty_contact_f3dm1", "xsph_corr_f3dm10_eos4_sym"])
def test_text_deck_runs_like_the_binary_deck(name, tmp_path):
    """A run created from the reference's text deck equals the run from the
    binary deck of the same content, bit for bit."""
    txt = str(tmp_path / (name + ".mcm"))
    cases.build(name).write_mcm(txt)
    n = cases.CASES[name][0]
    a = _fused_dump(os.path.join(DATA, name + ".bin"), tmp_path, n, 1)
    b = _fused_dump(txt, tmp_path, n, 2)
    np.testing.assert_array_equal(a.rows, b.rows)
    np.testing.assert_array_equal(a.nbr, b.nbr)


@pytest.mark.parametrize("name", ["plain_f3dm3", "corr_f3dm10_eos4_hopt1", "quarter_f3dm1_xyz_sym_tcrit2"])
def test_first_step_may_reuse_the_initial_search(name, tmp_path):
    """mcm_gpu_initial searches before it sets the first dt, and the strain
    pass's gather records carry positions half a step back with the step's
    dt.  Its lists hold until the particles move, so the first step may start
    with the strain pass: the records must then have been refreshed with the
    first dt, and the step equal one that searches again, bit for bit."""
    path = os.path.join(DATA, name + ".bin")

    def first_step(search: bool):
        with gpu.Solver(path) as s:
            s.initial()
            for st in STAGES:
                if st != "neighbours" or search:
                    s.stage(st)
            return gdump(s, tmp_path, 1, 0 if search else 1)

    a, b = first_step(True), first_step(False)
    _assert_same_state(a, b)
    assert np.array_equal(a.nbr, b.nbr) and np.array_equal(a.gnbr, b.gnbr)


@pytest.mark.parametrize("name", ["plain_f3dm3", "corr_f3dm10_eos4_hopt1"])
def test_dt_update_between_search_and_strain(name, tmp_path):
    """A dt update between the neighbour search and the strain pass (off the
    reference's order) moves the strain records' half step with it: the
    strain pass then equals the one after a search made with the new dt."""
    path = os.path.join(DATA, name + ".bin")

    def strain_after(order, tag):
        with gpu.Solver(path) as s:
            s.initial()
            for st in order:
                s.stage(st)
            return gdump(s, tmp_path, 1, tag)

    a = strain_after(("neighbours", "dt_update", "strain"), 1)
    b = strain_after(("dt_update", "neighbours", "strain"), 2)
    for k in ("x", "rod", "spin", "tracerod"):
        np.testing.assert_array_equal(a.f(k), b.f(k), err_msg=k)
    assert a.dt == b.dt and np.array_equal(a.nbr, b.nbr)
