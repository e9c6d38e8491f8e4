"""Parity of the HIP path at the workloads' own sizes.

* The 2,094,492-particle corrected Taylor bar that the bench times (dx =
  0.073 mm, f3dm10 + eqos4, h_opt 1), two steps stage by stage against the
  CPU oracle (its pair loops on OpenMP threads): neighbour lists equal to the
  oracle's in order (hence as sets), every state field within 1e-10
  norm-relative after every stage, the scalars likewise; then the fused step
  is deterministic at that size.
* A plain bar of the same size without its symmetry plane: the momentum
  stage conserves momentum (sum of m a = 0 up to rounding).
* The 256-bar batched workload: a sample of the runs equals the same decks
  stepped alone, bit for bit.
"""
import os

import numpy as np
import pytest

from mcmtop_amd import deck, gpu
from oracle import STAGES, Oracle

pytestmark = pytest.mark.gpu

# per-step tolerance of the accuracy gates (norm-relative)
TOL = 1e-10
# fields the dump rows hold that the stages compute (not the host-side x0)
FIELDS = {"x": slice(0, 3), "v": slice(3, 6), "a": slice(6, 9), "rho": 9, "rhoold": 10, "h": 11, "hold": 12,
          "p": 13, "e": 14, "etry": 15, "einc": 16, "c": 17, "vabs": 18, "tracerod": 19, "pcut": 20,
          "p_cut": 21, "efps": 22, "fail": 23, "mindist": 24, "critts": 25, "mass": 26, "rho0": 27,
          "sigma": slice(28, 37), "s": slice(37, 46), "alfa": slice(46, 55), "q": slice(55, 64),
          "rod": slice(64, 73), "spin": slice(73, 82), "h0": 82}


def _rel(a, b):
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def _row_errors(g, o):
    return {k: _rel(g[:, s], o[:, s]) for k, s in FIELDS.items()}


def _gpu_lists(s, n):
    counts = np.zeros(n, np.int32)
    gcounts = np.zeros(n, np.int32)
    cap = gpu.ctypes.c_int32()
    lib = s._lib
    i32p = gpu.ctypes.POINTER(gpu.ctypes.c_int32)
    # capacity first, then the ids
    rc = lib.mcm_gpu_download_neighbours(s._h, counts.ctypes.data_as(i32p), gcounts.ctypes.data_as(i32p), None, 0,
                                         gpu.ctypes.byref(cap))
    assert rc == 0
    ids = np.full((n, cap.value), -1, np.int32)
    rc = lib.mcm_gpu_download_neighbours(s._h, counts.ctypes.data_as(i32p), gcounts.ctypes.data_as(i32p),
                                         ids.ctypes.data_as(i32p), ids.size, gpu.ctypes.byref(cap))
    assert rc == 0
    return counts, gcounts, ids


def _assert_lists_equal(s, o, n):
    gc, gg, gids = _gpu_lists(s, n)
    oc, og, oids = o.lists(n, gids.shape[1])
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gg, og)
    mask = np.arange(gids.shape[1])[None, :] < (oc + og)[:, None]
    assert np.array_equal(np.where(mask, gids, -1), np.where(mask, oids, -1)), "lists differ (order or members)"
    return float((oc + og).mean())


@pytest.mark.timeout(900)
def test_config2_bar_stage_by_stage(tmp_path):
    dk = deck.config2()
    path = str(tmp_path / "bar.bin")
    dk.write_bin(path)
    n = dk.np_
    print("config2 bar: %d particles" % n, flush=True)
    assert n == 2094492
    o = Oracle(path, ncap=100, gcap=64)
    assert o.initial() == 0
    with gpu.Solver(path) as s:
        s.initial()
        rows, _ = s.download_state()
        errs = _row_errors(rows, o.rows(n))
        assert max(errs.values()) <= TOL, errs
        for step in (1, 2):
            for st in STAGES:
                s.stage(st)
                assert o.stage(st) == 0, st
                if st == "neighbours":
                    mean = _assert_lists_equal(s, o, n)
                    print("step %d: lists equal, mean length %.2f" % (step, mean), flush=True)
                rows, sc = s.download_state()
                orow = o.rows(n)
                errs = _row_errors(rows, orow)
                bad = {k: v for k, v in errs.items() if v > TOL}
                assert not bad, (step, st, bad)
                osc = o.scalars()
                for k in ("dt", "dtold", "ptime", "critts", "hmax"):
                    assert abs(getattr(sc, k) - osc[k]) <= TOL * abs(osc[k]), (step, st, k)
                if st in ("calctotale",):
                    assert abs(sc.kinetice - osc["kinetice"]) <= 1e-9 * abs(osc["kinetice"])
                    assert abs(sc.internale - osc["internale"]) <= 1e-9 * abs(osc["internale"])
            print("step %d: every stage within %.0e of the oracle (max %.2e)" % (
                step, TOL, max(_row_errors(rows, orow).values())), flush=True)
    del o


@pytest.mark.timeout(600)
def test_config2_bar_fused_steps_are_deterministic(tmp_path):
    dk = deck.config2()
    path = str(tmp_path / "bar.bin")
    dk.write_bin(path)
    out = []
    for _ in range(2):
        with gpu.Solver(path) as s:
            s.initial()
            s.step(6)
            rows, sc = s.download_state()
            out.append((rows, sc.ptime, sc.kinetice, sc.internale))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


@pytest.mark.timeout(600)
def test_plain_bar_at_scale_conserves_momentum(tmp_path):
    """Plain SPH, 2.1 M particles, no symmetry plane (no ghosts), the two
    halves of the bar driven into each other: the pair forces are
    antisymmetric, so the momentum stage's sum of m a vanishes up to
    rounding."""
    dk = deck.taylor_bar(dx=7.3e-5, model="f3dm3", sym=False, fit_text=False)
    z = dk.x[:, 2]
    dk.v[:, 2] = np.where(z > 0.5 * (z.min() + z.max()), -190.0, 190.0)
    path = str(tmp_path / "plain.bin")
    dk.write_bin(path)
    with gpu.Solver(path) as s:
        s.initial()
        s.step(3)
        for st in ("neighbours", "strain", "update_h", "constitutive", "dt_update", "momentum"):
            s.stage(st)
        m = s.field(gpu.F_MASS)
        a = s.field(gpu.F_A).reshape(-1, 3)
    ma = m[:, None] * a
    assert np.abs(ma.sum(axis=0)).max() <= 1e-12 * np.abs(ma).sum(axis=0).max()


@pytest.mark.timeout(600)
def test_batched_256_bars_sample_equals_runs_alone(tmp_path):
    """The batched bench workload (256 bars of 27 k - 200 k particles, 28.9 M
    in all): after 4 batched steps, every 32nd run equals its deck stepped
    alone for 4 steps, bit for bit."""
    decks, params = deck.batch_bars(256, seed=3)
    print("256 bars, %d particles" % sum(p["np"] for p in params), flush=True)
    sample = list(range(0, 256, 32)) + [255]
    with gpu.Batch(decks) as b:
        b.initial()
        b.step(4)
        got = {k: b.run(k).download_state()[0] for k in sample}
    for k in sample:
        with gpu.Solver(decks[k]) as s:
            s.initial()
            s.step(4)
            rows, _ = s.download_state()
        assert np.array_equal(rows, got[k]), "run %d differs from its deck stepped alone" % k
