"""Batched runs (mcm_gpu_batch_*): independent decks stepped together in one
launch set.  Each run keeps its own lists, arithmetic and scalars, so every
run of a batch must give the state of the same deck stepped alone with
mcm_gpu_step, bit for bit (and therefore the same parity with the oracle as
the single-run path, which test_gpu_parity checks).
"""
import os

import numpy as np
import pytest

import cases
from dumpio import Dump
from mcmtop_amd import deck, gpu

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")
STATE = ("x", "v", "rho", "h", "p", "e", "sigma", "alfa", "efps", "q", "c", "vabs", "pcut", "fail")


def _dump(run: gpu.Solver, tmp_path, tag) -> Dump:
    p = str(tmp_path / ("b_%s.bin" % tag))
    run.dump(p, 0, 0)
    return Dump(p)


def _alone(path, nsteps, tmp_path, tag, **opts) -> Dump:
    with gpu.Solver(path, **opts) as s:
        s.initial()
        s.step(nsteps)
        return _dump(s, tmp_path, tag)


def _assert_same_run(got: Dump, ref: Dump, what: str):
    for k in STATE:
        np.testing.assert_array_equal(got.f(k), ref.f(k), err_msg="%s: %s" % (what, k))
    np.testing.assert_array_equal(got.nbr, ref.nbr, err_msg=what + ": lists")
    np.testing.assert_array_equal(got.gnbr, ref.gnbr, err_msg=what + ": ghost lists")
    for k in ("ptime", "kinetice", "internale"):
        assert getattr(got, k) == getattr(ref, k), (what, k)


@pytest.mark.parametrize("pair_mapping", [0, 1])
def test_batch_runs_equal_runs_stepped_alone(pair_mapping, tmp_path):
    """All golden decks in one batch: both discretisations, one and two
    materials, no plane up to three planes and a periodic pair, every model
    and EOS of the path.  7 + 16 steps (an odd call runs one step outside the
    captured pair) against 23 steps of each deck alone."""
    names = sorted(cases.CASES)
    paths = [os.path.join(DATA, n + ".bin") for n in names]
    with gpu.Batch(paths, pair_mapping=pair_mapping) as b:
        assert len(b) == len(names)
        b.initial()
        b.step(7)
        b.step(16)
        got = [_dump(b.run(k), tmp_path, "b%d" % k) for k in range(len(names))]
    for k, (name, path) in enumerate(zip(names, paths)):
        print("batch run", k, name, flush=True)
        ref = _alone(path, 23, tmp_path, "a%d" % k, pair_mapping=pair_mapping)
        _assert_same_run(got[k], ref, name)


def test_batch_run_that_outgrows_its_ghost_capacity(tmp_path):
    """plain_f3dm3 makes 24 ghosts at the start and 36 a few steps later: with
    24 ghost slots it fits the initial search, overflows inside a batched
    step, grows and finishes the call's steps on its own; it and the other
    runs still equal their runs alone (default capacities), and the batch
    steps on afterwards.  Runs are added as binary deck bytes."""
    names = ["plain_f3dm3", "corr_f3dm10_eos4_hopt1", "plain_f3dm3"]
    decks = [deck.Deck.read_bin(os.path.join(DATA, n + ".bin")) for n in names]
    with gpu.Batch(ghost_capacity=24) as b:
        for dk in decks:
            b.add(dk)
        b.initial()
        assert b.run(0).scalars().ngp == 24
        b.step(20)
        assert b.run(0).scalars().ngp > 24
        b.step(6)
        got = [_dump(b.run(k), tmp_path, "g%d" % k) for k in range(len(names))]
        steps = [b.run(k).scalars().timestep for k in range(len(names))]
    assert len(set(steps)) == 1
    for k, name in enumerate(names):
        ref = _alone(os.path.join(DATA, name + ".bin"), 26, tmp_path, "r%d" % k)
        _assert_same_run(got[k], ref, name)


def test_run_from_deck_bytes_equals_run_from_file(tmp_path):
    path = os.path.join(DATA, "corr_f3dm10_eos4_hopt1.bin")
    a = _alone(path, 10, tmp_path, "file")
    with gpu.Solver(deck.Deck.read_bin(path)) as s:
        s.initial()
        s.step(10)
        b = _dump(s, tmp_path, "bytes")
    _assert_same_run(b, a, "bytes")


def test_batched_bars_of_uneven_size(tmp_path):
    """Four bars of the batched workload's generator (sizes, shapes,
    materials and speeds differ), scaled to small counts, with the mapping
    autotuner and the per-stage timing call in the run: each run equals its
    bar alone after 16 steps."""
    decks, params = deck.batch_bars(count=4, seed=11, nmin=2000, nmax=12000, dx_range=(0.3e-3, 0.5e-3))
    assert len({p["np"] for p in params}) == 4
    with gpu.Batch(decks) as b:
        b.initial()
        b.step(4)
        mapping = b.tune(2)          # 8 steps, both mappings
        b.step(2)
        ms = b.time_steps(2)
        got = [_dump(b.run(k), tmp_path, "u%d" % k) for k in range(4)]
    assert mapping in (0, 1)
    assert np.all(ms > 0)
    for k, dk in enumerate(decks):
        ref = _alone(dk, 16, tmp_path, "v%d" % k)
        _assert_same_run(got[k], ref, "bar %d" % k)
