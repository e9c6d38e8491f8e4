"""Reverse-mode derivative of the plain-SPH step (mcm_gpu_adj_*), checked
against central finite differences of the CPU oracle, which restates the
reference step (tests/test_oracle_golden.py pins it to the reference's own
state dumps).

Case: an f3dm10 + eqos4 Taylor bar (the differentiation workload's material
and EOS, 800 particles at dx = 0.8 mm), 30 steps, loss = identification loss
of mcmtop_amd.adjoint (final length, mushroom radius, 8 sampled surface
displacement histories) against a target from hidden parameters.  The
gradient is taken at parameters away from the hidden ones, with respect to
the four model constants (G, sigma_y, Eh, S1) and to single initial-velocity
components.

The initial velocities carry a seeded 0.5 % jitter.  With a uniform start
every particle has the same critical step, dt = tssf min(critts) is then not
differentiable with respect to one particle's velocity (a central difference
sees half of a one-sided slope) and only the parameter gradients, which move
all tied particles alike, would be comparable.  Eh barely changes the loss in
30 steps: its central difference is taken with a relative step of 1e-5 (at
1e-6 rounding of the loss difference dominates) and held to 1e-4.
"""
import os

import numpy as np
import pytest
import torch

from mcmtop_amd import adjoint, deck, gpu
from oracle import Oracle

pytestmark = pytest.mark.gpu

NSTEPS = 30
DX = 0.8e-3
RADIUS = 3.2e-3
NU = 0.35
FD_REL = (1.0e-6, 1.0e-6, 1.0e-5, 1.0e-6)   # relative steps of the central differences (G, sigma_y, Eh, S1)
TOL = (1.0e-5, 1.0e-5, 1.0e-4, 1.0e-5)      # gradient agreement (relative)
FD_V = 1.0e-3                               # m/s, step of the initial-velocity differences


def _bar():
    dk = deck.taylor_bar(dx=DX, radius=RADIUS, length=12.8e-3, model="f3dm10", tcrit_opt=1, seed=5)
    dk.v = dk.v * (1.0 + np.random.default_rng(7).uniform(-0.005, 0.005, size=dk.v.shape))
    return dk


def _model_params(dk):
    m = dk.materials[0]
    return np.array([m.card[0], m.card[1], m.card[2], m.eoscard[1]])


def _with(dk, params=None, v0=None):
    import copy
    d2 = copy.deepcopy(dk)
    if params is not None:
        d2.materials[0].card[0:3] = params[0:3]
        d2.materials[0].eoscard[1] = params[3]
    if v0 is not None:
        d2.v = np.array(v0, dtype=np.float64).reshape(-1, 3)
    return d2


def _loss_np(x, hist, x0s, target):
    length = x[:, 2].max() - x[:, 2].min()
    radius = np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2).max()
    disp = hist - x0s[None]
    return (((length - target["length"]) / target["length"]) ** 2 + ((radius - target["radius"]) / target["radius"]) ** 2
            + ((disp - target["disp"]) ** 2).sum() / (target["disp"] ** 2).sum())


def _oracle_loss(dk, sample, target, tmp_path):
    p = str(tmp_path / "fd.bin")
    dk.write_bin(p)
    o = Oracle(p)
    assert o.initial() == 0
    n = dk.np_
    hist = np.zeros((NSTEPS, len(sample), 3))
    for k in range(NSTEPS):
        assert o.step() == 0
        hist[k] = o.positions(n)[sample]
    x = o.positions(n)
    o.close()
    return _loss_np(x, hist, dk.x[sample], target)


@pytest.fixture(scope="module")
def problem():
    dk = _bar()
    sample = adjoint.surface_sample(dk.x, RADIUS, DX, 8, seed=2)
    hidden = _model_params(dk)
    # evaluation point away from the hidden parameters
    p_eval = hidden * np.array([1.05, 0.9, 1.2, 1.03])
    with gpu.Solver(dk) as s:
        s.set_material_params(0, hidden)
        s.initial()
        rec = s.record(NSTEPS, NSTEPS, sample)
        x = s.field(gpu.F_X)
        hist = rec.hist.copy()
        rec.close()
    target = {"length": x[:, 2].max() - x[:, 2].min(), "radius": np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2).max(),
              "disp": hist - dk.x[sample][None]}
    return dk, sample, p_eval, target


def _gpu_grad(dk, sample, params, target, ckpt):
    tt = {k: torch.as_tensor(v) for k, v in target.items()}
    with gpu.Solver(dk) as s:
        p = torch.tensor(params.reshape(1, 4), dtype=torch.float64, requires_grad=True)
        v = torch.tensor(dk.v, dtype=torch.float64, requires_grad=True)
        x, hist = adjoint.StepRun.apply(p, v, s, NSTEPS, ckpt, sample)
        L = adjoint.identification_loss(x, hist, torch.from_numpy(dk.x[sample]), tt)
        L.backward()
        return L.item(), p.grad.numpy()[0].copy(), v.grad.numpy().copy()


def test_parameter_gradients_match_oracle_differences(problem, tmp_path):
    dk, sample, p_eval, target = problem
    L, gp, gv = _gpu_grad(dk, sample, p_eval, target, 7)
    print("loss %.12e, nonfinite v0 gradients %d" % (L, int(np.sum(~np.isfinite(gv)))))
    L_orc = _oracle_loss(_with(dk, p_eval), sample, target, tmp_path)
    assert abs(L - L_orc) <= 1e-9 * abs(L_orc)
    fds = []
    for k in range(4):
        h = FD_REL[k] * abs(p_eval[k])
        pp, pm = p_eval.copy(), p_eval.copy()
        pp[k] += h
        pm[k] -= h
        fd = (_oracle_loss(_with(dk, pp), sample, target, tmp_path) -
              _oracle_loss(_with(dk, pm), sample, target, tmp_path)) / (pp[k] - pm[k])
        print("param %d: adjoint %.12e  fd %.12e  rel %.2e" % (k, gp[k], fd, abs(gp[k] - fd) / abs(fd)))
        fds.append(fd)
    for k in range(4):
        assert abs(gp[k] - fds[k]) <= TOL[k] * abs(fds[k]), (k, gp[k], fds[k])


def test_initial_velocity_gradients_match_oracle_differences(problem, tmp_path):
    dk, sample, p_eval, target = problem
    _, _, gv = _gpu_grad(dk, sample, p_eval, target, 7)
    # the largest components, and two sampled particles' z velocity
    assert np.all(np.isfinite(gv))
    flat = np.argsort(-np.abs(gv).ravel())[:3].tolist() + [3 * int(sample[0]) + 2, 3 * int(sample[3]) + 0]
    scale = np.abs(gv).max()
    res = []
    for f in flat:
        i, c = divmod(f, 3)
        h = FD_V
        vp, vm = dk.v.copy(), dk.v.copy()
        vp[i, c] += h
        vm[i, c] -= h
        p_dk = _with(dk, p_eval)
        fd = (_oracle_loss(_with(p_dk, v0=vp), sample, target, tmp_path) -
              _oracle_loss(_with(p_dk, v0=vm), sample, target, tmp_path)) / (vp[i, c] - vm[i, c])
        print("v0[%d,%d]: adjoint %.12e  fd %.12e" % (i, c, gv[i, c], fd))
        res.append((i, c, gv[i, c], fd))
    for i, c, g, fd in res:
        assert abs(g - fd) <= 1.0e-5 * max(abs(fd), 1e-3 * scale), (i, c, g, fd)


def test_checkpoint_interval_does_not_change_the_gradient(problem):
    dk, sample, p_eval, target = problem
    ref = _gpu_grad(dk, sample, p_eval, target, 1)
    for ck in (7, NSTEPS):
        got = _gpu_grad(dk, sample, p_eval, target, ck)
        assert got[0] == ref[0]
        np.testing.assert_array_equal(got[1], ref[1])
        np.testing.assert_array_equal(got[2], ref[2])


def test_reverse_kernel_timing_leaves_the_gradient_unchanged(problem, capfd, monkeypatch):
    """MCM_ADJ_TIMING=1 at recording time: the reverse steps time their kernels
    with events, the totals are printed when the recording is released (with
    its run, at the end of _gpu_grad), and the gradient is the same bit for bit."""
    dk, sample, p_eval, target = problem
    ref = _gpu_grad(dk, sample, p_eval, target, 7)
    capfd.readouterr()
    monkeypatch.setenv("MCM_ADJ_TIMING", "1")
    got = _gpu_grad(dk, sample, p_eval, target, 7)
    monkeypatch.delenv("MCM_ADJ_TIMING")
    err = capfd.readouterr().err
    assert got[0] == ref[0]
    np.testing.assert_array_equal(got[1], ref[1])
    np.testing.assert_array_equal(got[2], ref[2])
    lines = [ln for ln in err.splitlines() if ln.startswith("adjoint timing:")]
    assert len(lines) == 15, err
    assert "%d steps, checkpoint every 7, %d reverse steps" % (NSTEPS, NSTEPS) in lines[0]
    for name in ("k_adj_fill", "k_adj_sort", "k_adj_const", "k_adj_mom", "k_adj_strain", "k_adj_fold_end"):
        assert any(ln.split()[2] == name for ln in lines[1:]), name
    assert all(float(ln.split()[3]) > 0.0 for ln in lines[1:])


def test_unsupported_decks_are_rejected():
    dk = deck.taylor_bar(dx=DX, radius=RADIUS, length=12.8e-3, model="f3dm10", tcrit_opt=1, corrected=True)
    with gpu.Solver(dk) as s:
        s.initial()
        with pytest.raises(gpu.McmGpuError, match="plain SPH"):
            s.record(4, 2, [0])


def _instance(seed, length):
    dk = deck.taylor_bar(dx=DX, radius=RADIUS, length=length, model="f3dm10", tcrit_opt=1, seed=seed)
    dk.v = dk.v * (1.0 + np.random.default_rng(seed + 100).uniform(-0.005, 0.005, size=dk.v.shape))
    sample = adjoint.surface_sample(dk.x, RADIUS, DX, 8, seed=seed)
    ident = adjoint.Identification(dk, NSTEPS, 1, sample, NU)
    ident.make_target([117e9, 400e6, 100e6, 1.49])
    return ident


def test_identification_fits_run_concurrently_and_reduce_the_loss():
    """Two bars of different length fitted at once (one thread and stream
    each) from parameters 10-15 % off: the loss falls, and each instance's
    iterates are bit-identical to the same fit run alone."""
    starts = [np.array([117e9 * 1.1, 400e6 * 0.87, 100e6, 1.49 * 1.12]),
              np.array([117e9 * 0.9, 400e6 * 1.15, 100e6, 1.49 * 0.9])]
    idents = [_instance(11, 12.8e-3), _instance(12, 9.6e-3)]
    try:
        together = adjoint.fit_many(idents, starts, 8, lr=0.05, threads=2)
        alone = [adjoint.fit(idn, th, 8, lr=0.05) for idn, th in zip(idents, starts)]
    finally:
        for idn in idents:
            idn.close()
    for a, b in zip(together, alone):
        np.testing.assert_array_equal(np.array(a["loss"]), np.array(b["loss"]))
        np.testing.assert_array_equal(a["final"], b["final"])
        print("loss %.3e -> %.3e" % (a["loss"][0], a["loss"][-1]))
        assert a["loss"][-1] < 0.2 * a["loss"][0]


# ---- planes: corner ghosts made from ghosts, periodic ghosts
def _quarter_deck():
    import cases
    dx = 0.4e-3
    pts = cases._quarter_bar(dx, 1.2e-3, 7)
    rho = 8930.0
    m = deck.material(10, rho, rho * dx ** 3 * len(pts), 1.2 * dx, deck.f3dm10(45e9, 400e6, 100e6), eos=4,
                      eoscard={0: 3940.0, 1: 1.49, 4: 2.0, 5: 0.47})
    rng = np.random.default_rng(21)
    v = np.zeros((len(pts), 3))
    v[:, 2] = -150.0
    v += rng.uniform(-3.0, 3.0, size=v.shape)
    return deck.make_deck(pts, v, [m], tcrit_opt=1, jitter_dx=0.01 * dx, sym_flag=(1, 0, 1, 0, 1, 0))


def _periodic_deck():
    dx = 0.4e-3
    nx = 5
    pts = deck.block(dx, nx, 4, 5)
    rho = 2700.0
    m = deck.material(10, rho, rho * dx ** 3 * len(pts), 1.2 * dx, deck.f3dm10(26e9, 300e6, 500e6), eos=1,
                      eoscard={1: 70e9, 2: 30e9, 8: 1.0})
    rng = np.random.default_rng(22)
    v = np.zeros((len(pts), 3))
    v[:, 2] = -100.0
    v[:, 0] = 20.0 * np.sin(2 * np.pi * pts[:, 2] / (5 * dx))
    v += rng.uniform(-2.0, 2.0, size=v.shape)
    return deck.make_deck(pts, v, [m], jitter_dx=0.01 * dx, tcrit_opt=1,
                          sym_flag=(2, 2, 0, 0, 1, 0), sym_x=(0.0, nx * dx, 0, 0, 0, 0))


def _oracle_functional(dk, w, nsteps, tmp_path):
    p = str(tmp_path / "fdp.bin")
    dk.write_bin(p)
    o = Oracle(p)
    assert o.initial() == 0
    for _ in range(nsteps):
        assert o.step() == 0
    x = o.positions(dk.np_)
    o.close()
    return float((w * x).sum())


@pytest.mark.parametrize("which", ["corner", "periodic"])
def test_gradients_with_mirrored_and_periodic_ghosts(which, tmp_path):
    """L = sum w . x after 20 steps (random weights): parameter and
    initial-velocity gradients through ghost-of-ghost corners (three
    symmetry planes) and periodic ghosts against central differences of the
    oracle."""
    nsteps = 20
    dk = _quarter_deck() if which == "corner" else _periodic_deck()
    w = np.random.default_rng(5).standard_normal((dk.np_, 3))
    p0 = _model_params(dk)
    with gpu.Solver(dk) as s:
        s.initial()
        rec = s.record(nsteps, 7, [])
        v0b, pb = rec.backward(xbar=w)
        rec.close()
    gp = pb[0]
    for k in range(4):
        if p0[k] == 0.0:
            continue
        h = 1e-6 * abs(p0[k])
        pp, pm = p0.copy(), p0.copy()
        pp[k] += h
        pm[k] -= h
        fd = (_oracle_functional(_with(dk, pp), w, nsteps, tmp_path) -
              _oracle_functional(_with(dk, pm), w, nsteps, tmp_path)) / (pp[k] - pm[k])
        print("%s param %d: adjoint %.10e fd %.10e" % (which, k, gp[k], fd))
        assert abs(gp[k] - fd) <= 1e-5 * abs(fd) + 1e-9 * np.abs(gp).max(), (k, gp[k], fd)
    # the particle nearest the planes' common corner (or edge), and the largest components
    near = int(np.argmin(dk.x.sum(axis=1)))
    comps = [(near, 0), (near, 1), (near, 2)] + [divmod(int(f), 3) for f in np.argsort(-np.abs(v0b).ravel())[:2]]
    scale = np.abs(v0b).max()
    for i, c in comps:
        vp, vm = dk.v.copy(), dk.v.copy()
        vp[i, c] += 1e-3
        vm[i, c] -= 1e-3
        fd = (_oracle_functional(_with(dk, v0=vp), w, nsteps, tmp_path) -
              _oracle_functional(_with(dk, v0=vm), w, nsteps, tmp_path)) / 2e-3
        print("%s v0[%d,%d]: adjoint %.10e fd %.10e" % (which, i, c, v0b[i, c], fd))
        assert abs(v0b[i, c] - fd) <= 1e-5 * max(abs(fd), 1e-3 * scale), (i, c, v0b[i, c], fd)


def _two_material_deck():
    """f3dm10 + eqos4 bar with an f3dm1 (elastic) cap, in kernel contact
    (contacttype 0: one list for both materials), on a symmetry plane."""
    dx = 0.4e-3
    pts = deck.block(dx, 5, 5, 9)
    rho1, rho2 = 8930.0, 7850.0
    cap = pts[:, 2] > 6 * dx
    n1, n2 = int((~cap).sum()), int(cap.sum())
    m1 = deck.material(10, rho1, rho1 * dx ** 3 * n1, 1.2 * dx, deck.f3dm10(45e9, 400e6, 100e6), eos=4,
                       eoscard={0: 3940.0, 1: 1.49, 4: 2.0, 5: 0.47})
    m2 = deck.material(1, rho2, rho2 * dx ** 3 * n2, 1.2 * dx, deck.f3dm1(200e9, 0.3))
    rng = np.random.default_rng(31)
    v = np.zeros((len(pts), 3))
    v[:, 2] = -150.0
    v += rng.uniform(-3.0, 3.0, size=v.shape)
    return deck.make_deck(pts, v, [m1, m2], matid=np.where(cap, 2, 1), tcrit_opt=1, jitter_dx=0.01 * dx,
                          sym_flag=(0, 0, 0, 0, 1, 0))


def test_two_materials_with_an_elastic_part(tmp_path):
    """Model 10 and model 1 in one run: the gradient with respect to G,
    sigma_y, Eh, S1 of the first material and E, nu of the second (whose
    constants and frozen sound speed derive from them) against central
    differences of the oracle; L = sum w . x after 20 steps."""
    import copy
    nsteps = 20
    dk = _two_material_deck()
    w = np.random.default_rng(6).standard_normal((dk.np_, 3))
    with gpu.Solver(dk) as s:
        p_all = np.array([s.material_params(0), s.material_params(1)])
        s.initial()
        rec = s.record(nsteps, 5, [])
        _, pb = rec.backward(xbar=w)
        rec.close()
    np.testing.assert_array_equal(p_all[1], [200e9, 0.3, 0.0, 0.0])
    checks = [(0, k, (0, k) if k < 3 else None) for k in range(4)] + [(1, 0, (1, 0)), (1, 1, (1, 5))]
    for mat, k, where in checks:
        base = p_all[mat][k]
        h = 1e-6 * abs(base)
        vals = []
        for sgn in (1.0, -1.0):
            d2 = copy.deepcopy(dk)
            if where is None:
                d2.materials[0].eoscard[1] = base + sgn * h
            else:
                d2.materials[where[0]].card[where[1]] = base + sgn * h
            vals.append(_oracle_functional(d2, w, nsteps, tmp_path))
        fd = (vals[0] - vals[1]) / (2 * h)
        print("material %d param %d: adjoint %.10e fd %.10e" % (mat, k, pb[mat][k], fd))
        assert abs(pb[mat][k] - fd) <= 1e-5 * abs(fd), (mat, k, pb[mat][k], fd)
