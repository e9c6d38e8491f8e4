#!/usr/bin/env python3
"""Benchmark: FP64 particle-steps/s of the MCM SPH time step on one MI355X.

Workload (one GPU): the normalised (corrected) SPH Taylor bar at scale --
dx = 0.073 mm on r = 3.2 mm, L = 25.4 mm (2,094,492 particles), f3dm10 +
eqos4 (the reference rejects f3dm3 with an EOS), variable h (h_opt = 1),
symmetry plane at z = 0, v0 = 190 m/s, CFL 0.3, seeded +-0.01 dx jitter.
A "step" is one complete loop body of mcm_solution (neighbour search with
ghosts, strain, update_h, constitutive + EOS + AV + energy, dt, momentum,
boundary conditions, energy totals, velocity and position update), with
inputs resident in HBM; dt/ptime advance on the device.

Batched form (the "batched vs plain" part of the metric): 256 independent
bars of uneven size (27 k to 200 k particles, every parameter drawn from one
seed, mcmtop_amd.deck.batch_bars) stepped together in one launch set
(mcm_gpu_batch_step), against the same bars stepped one at a time, each with
its own captured step (mcm_gpu_step).  A sample of the runs is compared with
its run alone after the same steps.

Prints one JSON line (see the keys below).  --gpus N > 1 exits: the build
runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic cost of the two pair kernels of the corrected step (f3dm10 +
# eqos4, h_opt 1, tcrit_opt 1); nbar = mean list length (self and ghosts
# included).  Bytes: every per-particle word the kernel must read or write,
# once (cache-ideal gathers), plus the 4-byte list entries.
#   strain pass : reads the strain record {x, v, m/rho, h} (64 B), the list
#                 length (4 B) and the list; writes rod (6), spin (3),
#                 tracerod, x after the pass (3) and the momentum record's
#                 x, y (2): 188 + 4 nbar bytes
#   momentum    : reads the momentum record {x, h, m, m/rho, A(6)} (96 B), two
#                 list lengths (8 B), the list, v (3), |v|, e, dispbc, material
#                 (48 B); writes v, x, |v| (56 B): 208 + 4 nbar bytes
# Flops (add, mul, div, sqrt = 1; no FMA contraction): per list entry over the
# reference's three passes, and per particle (inverse of B, contractions,
# integration): strain 242 / 81, momentum 209 / 163.
# Algorithmic bytes (per particle, per list entry) and flops (per list
# entry, per particle) of the kernels that can dominate a step; keys
# (kernel, corrected passes).
KERNELS = {
    ("strain", 3): {"name": "k_strain_const<MODE_FSTRAIN, corrected>",
                    "symbol": "void k_strain_const<4, 1, false, 0>(DevState, int)",
                    "bytes": (188.0, 4.0), "flops": (242.0, 81.0)},
    ("momentum", 3): {"name": "k_momentum<MODE_FUSED, corrected>", "symbol": "void k_momentum<0, 1, false>(DevState, int)",
                      "bytes": (208.0, 4.0), "flops": (209.0, 163.0)},
    ("strain", 1): {"name": "k_strain_const<MODE_FSTRAIN, corrected one-pass>",
                    "symbol": "void k_strain_const<4, 2, false, 0>(DevState, int)",
                    "bytes": (188.0, 4.0), "flops": (100.0, 150.0)},
    ("momentum", 1): {"name": "k_momentum<MODE_FUSED, corrected one-pass>",
                      "symbol": "void k_momentum<0, 2, false>(DevState, int)",
                      "bytes": (208.0, 4.0), "flops": (130.0, 260.0)},
    ("list", 1): {"name": "k_list<one material>", "symbol": "void k_list<false, false>(DevState, int)",
                  "bytes": (56.0, 4.0), "flops": None},
    ("list", 3): {"name": "k_list<one material>", "symbol": "void k_list<false, false>(DevState, int)",
                  "bytes": (56.0, 4.0), "flops": None},
}
PROFILE = os.path.join(HERE, "profiles", "step_kernels_taylor_bar_corrected_2.1M.md")


def measured_traffic(symbol: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by the copy calibration), or None."""
    if not os.path.exists(PROFILE):
        return None
    in_table = False
    for line in open(PROFILE):
        if line.startswith("#"):
            in_table = "HBM traffic per dispatch" in line
            continue
        cells = [c.strip().strip("`") for c in line.split("|")]
        if in_table and len(cells) > 6 and cells[1] == symbol and cells[-2].replace(".", "").isdigit():
            return float(cells[-2]) * 1e9
    return None


def build_deck(path: str, dx: float, length: float):
    from mcmtop_amd import deck
    d = deck.taylor_bar(dx=dx, length=length, model="f3dm10", corrected=True, h_opt=1)
    d.write_bin(path)
    return d


def cpu_baseline(tmpdir: str, dx: float):
    """The reference Fortran (flang build of the reference sources, oracle/_ref)
    on a bounded sample of the same workload: the same bar cross-section and
    dx, 1.0 mm long (78,468 particles), 10 steps, one core."""
    probe = os.path.join(HERE, "oracle", "_ref", "ref_probe")
    path = os.path.join(tmpdir, "sample.bin")
    d = build_deck(path, dx, 1.0e-3)
    steps = 10
    kind = "reference"
    if os.path.exists(probe):
        out = subprocess.run([probe, path, tmpdir, str(steps), "0"], capture_output=True, text=True, timeout=600)
        secs = None
        for line in out.stdout.splitlines():
            if "loop_seconds=" in line:
                secs = float(line.split("loop_seconds=")[1])
        if secs is None:
            return None
    else:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        from oracle import Oracle
        kind = "port"
        o = Oracle(path)
        o.initial()
        t0 = time.perf_counter()
        o.run(steps)
        secs = time.perf_counter() - t0
    cpu = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": d.np_ * steps / secs, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "sample": "same bar section and dx (0.073 mm), 1.0 mm long (%d particles), %d steps, %s"
                      % (d.np_, steps, "flang -O2 build of the reference" if kind == "reference" else "C oracle"),
            "cpu": cpu, "host_threads": os.cpu_count()}


def kernel_ms(ms):
    return {"neighbours": ms[0], "list_build": ms[5], "strain_pass": ms[1], "constitutive": ms[2], "momentum": ms[3],
            "step": ms[4]}


def timed_run(deck_path: str, passes: int, steps: int, warmup: int, skin: float = 0.0):
    """warmup untimed steps, then `steps` timed steps of the bar (corrected SPH
    in `passes` passes, neighbour skin of width skin h_max) between two
    synchronisations of the run's stream, then per-kernel times of 10 more
    steps with HIP events."""
    from mcmtop_amd import gpu
    with gpu.Solver(deck_path, corrected_one_pass=passes == 1, nbr_skin=skin) as s:
        s.initial()
        s.step(warmup)
        # timed region: the step call enqueues on the context stream and
        # returns after synchronising it
        t0 = time.perf_counter()
        s.step(steps)
        t1 = time.perf_counter()
        sc = s.scalars()
        builds_timed = sc.skin_builds
        ms = s.time_steps(10)
        counts, gcounts, _ = s.neighbour_lists()
        nbar = float((counts + gcounts).mean())
    secs = t1 - t0
    print("bench: %d-pass corrected, skin %g, %.3f ms/step" % (passes, skin, 1e3 * secs / steps), file=sys.stderr,
          flush=True)
    return {"value": sc.np * steps / secs, "secs": secs, "ms_per_step": 1e3 * secs / steps, "ms": ms, "nbar": nbar,
            "sc": sc, "skin_builds": builds_timed}


def batched_leg(nruns: int, seed: int, steps: int, warmup: int, check_every: int, plain: bool = True,
                passes: int = 1):
    """Throughput of nruns uneven bars stepped as one batch and (plain) stepped
    one run at a time, both with the headline's corrected passes; equality of
    sampled runs between the two."""
    import numpy as np
    from mcmtop_amd import deck, gpu
    t = time.perf_counter()
    decks, params = deck.batch_bars(nruns, seed=seed)
    sizes = np.array([p["np"] for p in params])
    total = int(sizes.sum())
    print("bench: %d batch decks, %d particles in %.1f s" % (nruns, total, time.perf_counter() - t),
          file=sys.stderr, flush=True)
    t = time.perf_counter()
    b = gpu.Batch(decks, corrected_one_pass=passes == 1)
    b.initial()
    print("bench: batch initialised in %.1f s" % (time.perf_counter() - t), file=sys.stderr, flush=True)
    b.step(warmup)
    mapping = b.tune(2)
    tune_steps = 2 * (2 + 2)
    t0 = time.perf_counter()
    b.step(steps)
    secs_b = time.perf_counter() - t0
    ms_b = b.time_steps(2)
    print("bench: batched %.2f ms/step (pair mapping %d)" % (1e3 * secs_b / steps, mapping),
          file=sys.stderr, flush=True)
    sample = list(range(0, nruns, check_every))
    got = {k: (b.run(k).field(gpu.F_X), b.run(k).field(gpu.F_V), b.run(k).field(gpu.F_SIGMA)) for k in sample}
    b.close()
    secs_p = 0.0
    worst = 0.0
    for k, dk in enumerate(decks if plain else []):
        with gpu.Solver(dk, pair_mapping=mapping, corrected_one_pass=passes == 1) as s:
            s.initial()
            s.step(warmup + tune_steps)
            t0 = time.perf_counter()
            s.step(steps)
            secs_p += time.perf_counter() - t0
            if k in got:
                # the batch ran 2 more steps (its timing call): catch up
                s.step(2)
                for a, ref in zip(got[k], (s.field(gpu.F_X), s.field(gpu.F_V), s.field(gpu.F_SIGMA))):
                    nb = np.linalg.norm(ref)
                    worst = max(worst, float(np.linalg.norm(a - ref) / nb) if nb > 0 else float(np.linalg.norm(a)))
        if k % 32 == 31:
            print("bench: plain runs %d/%d" % (k + 1, nruns), file=sys.stderr, flush=True)
    v_b = total * steps / secs_b
    v_p = total * steps / secs_p if plain else None
    return {"runs": nruns, "particles": total, "seed": seed,
            "run_particles": {"min": int(sizes.min()), "median": float(np.median(sizes)), "max": int(sizes.max())},
            "steps": steps, "warmup": warmup, "pair_mapping": mapping, "corrected_passes": passes,
            "value": v_b, "unit": "particle-steps/s", "ms_per_step": 1e3 * secs_b / steps,
            "plain_value": v_p, "plain_ms_per_step": 1e3 * secs_p / steps if plain else None,
            "speedup": v_b / v_p if plain else None,
            "checked_runs": len(sample) if plain else 0, "max_rel_l2_vs_plain": worst if plain else None,
            "kernel_ms": {"neighbours": ms_b[0], "strain_pass": ms_b[1], "constitutive": ms_b[2],
                          "momentum": ms_b[3], "step": ms_b[4]},
            "workload": "taylor_bar_corrected_f3dm10_eqos4_hopt1 x %d, r U[2,5] mm, L/D U[3,8], dx U[0.05,0.2] mm, "
                        "v0 U[100,250] m/s, sigma_y U[200,600] MPa, E U[70,210] GPa" % nruns}


def adjoint_leg(steps: int, ckpt: int, dx: float, nsample: int, fd: bool):
    """Differentiation workload: gradients of the Taylor-bar identification
    loss (final length, mushroom radius, sampled surface displacement
    histories) with respect to (E, sigma_y, E_tan, S1) and the initial
    velocities, by the device reverse sweep; times of the plain forward, the
    recorded forward and the reverse sweep; central differences of the GPU
    forward for the four parameters and two velocity components."""
    import numpy as np
    import torch
    from mcmtop_amd import adjoint, deck, gpu
    radius, nu = 3.2e-3, 0.35
    dk = deck.taylor_bar(dx=dx, radius=radius, length=25.4e-3, model="f3dm10", tcrit_opt=1, seed=1,
                         fit_text=False)
    # seeded 0.5 % velocity jitter: with a uniform start all critical steps tie
    # and dt = tssf min(critts) has no derivative per particle velocity
    dk.v = dk.v * (1.0 + np.random.default_rng(7).uniform(-0.005, 0.005, size=dk.v.shape))
    sample = adjoint.surface_sample(dk.x, radius, dx, nsample, seed=2)
    hidden = np.array([117e9, 400e6, 100e6, 1.49])
    theta = hidden * np.array([1.05, 0.9, 1.2, 1.03])
    ident = adjoint.Identification(dk, steps, ckpt, sample, nu)
    ident.make_target(hidden)
    s = ident.solver
    # plain forward (the fused step, no recording)
    s.set_material_params(0, adjoint.physical_to_model(torch.tensor(theta), nu).numpy())
    s.set_initial_velocity(dk.v)
    s.initial()
    s.step(2)
    s.initial()
    t0 = time.perf_counter()
    s.step(steps)
    t_fwd = time.perf_counter() - t0
    # recorded forward + reverse sweep, timed separately, with the
    # workload's checkpoint interval and with every step kept in HBM
    modes = {}
    for ck in (ckpt, 1):
        ident.ckpt_every = ck
        th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
        v0 = torch.from_numpy(dk.v.copy()).requires_grad_(True)
        t0 = time.perf_counter()
        x, hist = ident.run(th, v0)
        t_rec = time.perf_counter() - t0
        L = adjoint.identification_loss(x, hist, ident.x0s, ident.target)
        t0 = time.perf_counter()
        L.backward()
        t_bwd = time.perf_counter() - t0
        modes[ck] = (t_rec, t_bwd, L.item(), th.grad.numpy().copy(), v0.grad.numpy().copy())
        print("bench: adjoint %d particles, checkpoint every %d: forward %.3f s, recorded %.3f s, reverse %.3f s"
              % (dk.np_, ck, t_fwd, t_rec, t_bwd), file=sys.stderr, flush=True)
    t_rec, t_bwd, Lv, g_theta, g_v = modes[ckpt]
    g1 = modes[1]
    out = {"workload": "taylor_bar_plain_f3dm10_eqos4_tcrit1", "particles": dk.np_, "dx_mm": dx * 1e3,
           "steps": steps, "ckpt_every": ckpt, "sampled": len(sample), "loss": Lv,
           "forward_s": t_fwd, "recorded_forward_s": t_rec, "reverse_s": t_bwd,
           "ratio_reverse": t_bwd / t_fwd, "ratio_recorded_plus_reverse": (t_rec + t_bwd) / t_fwd,
           "every_step_kept": {"recorded_forward_s": g1[0], "reverse_s": g1[1], "ratio_reverse": g1[1] / t_fwd,
                               "ratio_recorded_plus_reverse": (g1[0] + g1[1]) / t_fwd,
                               "max_rel_diff_vs_checkpointed": float(max(
                                   np.max(np.abs(g1[3] - g_theta) / np.abs(g_theta)),
                                   np.max(np.abs(g1[4] - g_v)) / np.max(np.abs(g_v))))},
           "grad_theta": {"E": g_theta[0], "sigma_y": g_theta[1], "E_tan": g_theta[2], "S1": g_theta[3]},
           "grad_v0_norm": float(np.linalg.norm(g_v))}
    if fd:
        def loss_at(th_, v_):
            with torch.no_grad():
                xx, hh = ident.run(torch.tensor(th_), torch.from_numpy(v_))
                return adjoint.identification_loss(xx, hh, ident.x0s, ident.target).item()
        def param_fd(g):
            rel_, spread_ = {}, {}
            for k, name in enumerate(("E", "sigma_y", "E_tan", "S1")):
                fds = {}
                for step in (1e-5, 1e-6, 1e-7):
                    h = step * theta[k]
                    tp, tm = theta.copy(), theta.copy()
                    tp[k] += h
                    tm[k] -= h
                    fds[step] = (loss_at(tp, dk.v) - loss_at(tm, dk.v)) / (tp[k] - tm[k])
                rel_[name] = abs(g[k] - fds[1e-6]) / abs(fds[1e-6])
                spread_[name] = (max(fds.values()) - min(fds.values())) / abs(fds[1e-6])
            return rel_, spread_
        rel, spread = param_fd(g_theta)
        for f_ in np.argsort(-np.abs(g_v).ravel())[:2]:
            i, c = divmod(int(f_), 3)
            vp, vm = dk.v.copy(), dk.v.copy()
            vp[i, c] += 1e-3
            vm[i, c] -= 1e-3
            f = (loss_at(theta, vp) - loss_at(theta, vm)) / 2e-3
            rel["v0[%d,%d]" % (i, c)] = abs(g_v[i, c] - f) / abs(f)
        out["fd_rel_error"] = rel
        out["fd_spread"] = spread
        out["fd"] = ("central differences of the GPU forward, relative step 1e-6, velocity step 1e-3 m/s; "
                     "fd_spread: relative spread of the differences at steps 1e-5, 1e-6, 1e-7 (kinks of the "
                     "800-step loss: yield transitions, min/max switches)")
        # the same bar and loss over a shorter horizon, whose loss has fewer kinks
        short = min(100, steps)
        ident.nsteps, ident.ckpt_every = short, 1
        ident.make_target(hidden)
        th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
        x, hist = ident.run(th, torch.from_numpy(dk.v.copy()))
        adjoint.identification_loss(x, hist, ident.x0s, ident.target).backward()
        rel_s, spread_s = param_fd(th.grad.numpy().copy())
        out["fd_short"] = {"steps": short, "rel_error": rel_s, "spread": spread_s}
    ident.close()
    return out


def ident_leg(ninst: int, steps: int, iters: int, threads: int, seed: int):
    """Batched identification, reduced: ninst independent bars of uneven
    size and material (deck.batch_bars: plain SPH, f3dm10 + eqos4, fixed h),
    each with its own hidden (E, sigma_y) and S1 = 1.49, fitted from seeded
    start points 10-20 % off by Adam on the logarithms over (E, sigma_y, S1),
    every gradient by the device reverse sweep with every step kept; the
    instances run concurrently (one host thread and HIP stream each).  One
    iteration is also timed with the instances one after another."""
    import numpy as np
    from mcmtop_amd import adjoint, deck
    decks, params = deck.batch_bars(ninst, seed=seed, nmin=15000, nmax=40000, corrected=False, h_opt=0,
                                    dx_range=(0.15e-3, 0.3e-3))
    rng = np.random.default_rng(seed + 1)
    idents, starts, hidden = [], [], []
    for dk, pr in zip(decks, params):
        dk.v = dk.v * (1.0 + rng.uniform(-0.005, 0.005, size=dk.v.shape))
        sample = adjoint.surface_sample(dk.x, pr["radius"], pr["dx"], 32, seed=int(rng.integers(1 << 30)))
        idn = adjoint.Identification(dk, steps, 1, sample, 0.35)
        th = np.array([pr["E"], pr["sigma_y"], 100e6, 1.49])
        idn.make_target(th)
        idents.append(idn)
        hidden.append(th)
        starts.append(th * np.array([rng.uniform(0.8, 1.2), rng.uniform(0.8, 1.2), 1.0, rng.uniform(0.9, 1.1)]))
    adjoint.fit_many(idents, starts, 1, threads=threads)    # buffers allocated once per run
    t0 = time.perf_counter()
    adjoint.fit_many(idents, starts, 1, threads=1)
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = adjoint.fit_many(idents, starts, iters, lr=0.03, threads=threads)
    t_conc = (time.perf_counter() - t0) / iters
    for idn in idents:
        idn.close()
    total = int(sum(p["np"] for p in params))
    ratio = [r["loss"][-1] / r["loss"][0] for r in res]
    names = {"E": 0, "sigma_y": 1, "S1": 3}
    perr = {nm: {"start_median": float(np.median([abs(st[k] / h[k] - 1) for st, h in zip(starts, hidden)])),
                 "after_median": float(np.median([abs(r["final"][k] / h[k] - 1) for r, h in zip(res, hidden)]))}
            for nm, k in names.items()}
    print("bench: identification %d bars (%d particles), %d steps: %.2f s per iteration concurrent, %.2f s one "
          "after another" % (ninst, total, steps, t_conc, t_seq), file=sys.stderr, flush=True)
    return {"instances": ninst, "particles": total, "steps": steps, "iterations": iters, "threads": threads,
            "s_per_iteration_concurrent": t_conc, "s_per_iteration_one_by_one": t_seq,
            "speedup": t_seq / t_conc,
            "gradient_particle_steps_per_s": total * steps / t_conc,
            "loss_ratio_after": {"median": float(np.median(ratio)), "max": float(np.max(ratio))},
            "param_rel_error": perr,
            "workload": "batch_bars(%d, seed=%d), 15k-40k particles, plain SPH, f3dm10 + eqos4, tcrit_opt 1, "
                        "fit (E, sigma_y, S1) by Adam (lr 0.03) on the logarithms" % (ninst, seed)}


def with_elastic_cap(dk, dx: float, frac: float = 0.1):
    """The top frac of a bar's length as a second material, f3dm1 steel
    (E 200 GPa, nu 0.3): the elastic sub-part of the identification
    configuration, in kernel contact with the bar."""
    import numpy as np
    from mcmtop_amd import deck
    z = dk.x[:, 2]
    cap = z > z.min() + (1.0 - frac) * (z.max() - z.min())
    n2 = int(cap.sum())
    m1 = dk.materials[0]
    m1.mass = m1.mass * (dk.np_ - n2) / dk.np_
    rho2 = 7850.0
    m2 = deck.material(1, rho2, rho2 * dx ** 3 * n2, m1.h, deck.f3dm1(200e9, 0.3), av_q=m1.av_q, av_l=m1.av_l)
    dk.materials = [m1, m2]
    dk.mat = np.where(cap, 2, 1).astype(np.int32)
    return dk


def ident_config5(ninst: int, steps: int, ckpt: int, iters: int, threads: int, seed: int):
    """Batched identification at the configuration's size: ninst bars of
    about 100k particles (uneven radius, length, dx, v0, E, sigma_y), each an
    f3dm10 + eqos4 body with an f3dm1 cap, steps with a checkpoint every
    ckpt steps; (E, sigma_y, S1) of the body fitted by Adam from seeded
    start points.  Time per iteration over all instances (concurrent, each
    instance's buffers released when its fit ends)."""
    import numpy as np
    from mcmtop_amd import adjoint, deck
    t0 = time.perf_counter()
    decks, params = deck.batch_bars(ninst, seed=seed, nmin=90000, nmax=110000, corrected=False, h_opt=0,
                                    dx_range=(0.1e-3, 0.3e-3))
    rng = np.random.default_rng(seed + 1)
    idents, starts = [], []
    for dk, pr in zip(decks, params):
        dk = with_elastic_cap(dk, pr["dx"])
        dk.v = dk.v * (1.0 + rng.uniform(-0.005, 0.005, size=dk.v.shape))
        sample = adjoint.surface_sample(dk.x, pr["radius"], pr["dx"], 32, seed=int(rng.integers(1 << 30)))
        idn = adjoint.Identification(dk, steps, ckpt, sample, 0.35)
        th = np.array([pr["E"], pr["sigma_y"], 100e6, 1.49])
        idents.append((idn, th))
        starts.append(th * np.array([rng.uniform(0.8, 1.2), rng.uniform(0.8, 1.2), 1.0, rng.uniform(0.9, 1.1)]))
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda it: it[0].make_target(it[1]), idents))
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = adjoint.fit_many([i for i, _ in idents], starts, iters, lr=0.03, threads=threads, release=True)
    t_it = (time.perf_counter() - t0) / iters
    for idn, _ in idents:
        idn.close()
    total = int(sum(p["np"] for p in params))
    ratio = [r["loss"][-1] / r["loss"][0] for r in res]
    print("bench: identification at size, %d bars (%d particles), %d steps: %.1f s per iteration"
          % (ninst, total, steps, t_it), file=sys.stderr, flush=True)
    return {"instances": ninst, "particles": total, "steps": steps, "ckpt_every": ckpt, "iterations": iters,
            "threads": threads, "setup_s": t_setup, "s_per_iteration": t_it,
            "projected_200_iterations_min": 200 * t_it / 60.0,
            "gradient_particle_steps_per_s": total * steps / t_it,
            "loss_ratio_after": {"median": float(np.median(ratio)), "max": float(np.max(ratio))},
            "workload": "batch_bars(%d, seed=%d) at 90k-110k particles with an f3dm1 cap (top 10 %%), plain SPH, "
                        "f3dm10 + eqos4 body, tcrit_opt 1; (E, sigma_y, S1) of the body by Adam (lr 0.03)"
                        % (ninst, seed)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=7.3e-5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=256, help="runs of the batched leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-seed", type=int, default=3)
    ap.add_argument("--batch-plain", type=int, default=1,
                    help="1: also step the batch's bars one at a time (plain_value, speedup, equality check)")
    ap.add_argument("--adjoint-steps", type=int, default=800, help="steps of the differentiation leg (0: skip it)")
    ap.add_argument("--adjoint-ckpt", type=int, default=50)
    ap.add_argument("--adjoint-dx", type=float, default=2.54e-4)
    ap.add_argument("--adjoint-fd", type=int, default=1, help="1: check the gradients by central differences")
    ap.add_argument("--ident", type=int, default=16, help="bars of the batched identification leg (0: skip it)")
    ap.add_argument("--ident-steps", type=int, default=200)
    ap.add_argument("--ident-iters", type=int, default=10)
    ap.add_argument("--ident-threads", type=int, default=8)
    ap.add_argument("--ident-config5", type=int, default=0,
                    help="bars of the identification leg at the configuration's size (100k particles, 600 steps, "
                         "checkpoint every 40; 0: skip it)")
    ap.add_argument("--ident-config5-iters", type=int, default=2)
    ap.add_argument("--passes", type=int, default=1, choices=(1, 3),
                    help="corrected SPH of the headline run: 1, single-pass moments (FP64, within the "
                         "target's accuracy bar, tests/test_gpu_one_pass.py); 3, the reference's three "
                         "passes (bit-exact). The other mode is timed beside it.")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="neighbour skin of the headline run as a fraction of h_max (0: the reference's search "
                         "every step; > 0: skin lists with the same neighbour sets, tests/test_gpu_skin.py; "
                         "slower than the search over the first 200 steps of this workload, 5.2 against "
                         "4.4 ms per step at 0.1)")
    ap.add_argument("--other-steps", type=int, default=100,
                    help="timed steps of the other mode, the reference's three passes without a skin "
                         "(bit-exact; 0: skip it)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (a known-size kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the SPH step on one GPU; --gpus %d is not supported "
              "(multi-GPU slab decomposition is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 18   # float32 words
        src = torch.ones(n, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": args.copy_calibration, "bytes_each_way": 4 * n}), flush=True)
        return
    from mcmtop_amd import gpu
    tmpdir = tempfile.mkdtemp(prefix="mcmbench")
    deck_path = os.path.join(tmpdir, "cfg2.bin")
    t = time.perf_counter()
    d = build_deck(deck_path, args.dx, 25.4e-3)
    print("bench: deck %d particles in %.1f s" % (d.np_, time.perf_counter() - t), file=sys.stderr, flush=True)
    head = timed_run(deck_path, args.passes, args.steps, args.warmup, args.skin)
    n = d.np_
    value = head["value"]
    ms, nbar, sc1 = head["ms"], head["nbar"], head["sc"]
    # the dominant kernel: roofline against HBM (algorithmic bytes) and, for
    # the pair kernels, its FP64 rate (algorithmic flops)
    kms = {"list": ms[5], "strain": ms[1], "momentum": ms[3]}
    dom = max(kms, key=kms.get)
    K = KERNELS[(dom, args.passes)]
    t = kms[dom] * 1e-3
    bytes_launch = n * (K["bytes"][0] + K["bytes"][1] * nbar)
    achieved = bytes_launch / t / 1e9
    traffic = measured_traffic(K["symbol"])
    fp64 = None
    if K["flops"]:
        flops_launch = n * (K["flops"][0] * nbar + K["flops"][1])
        fp64 = {"achieved": flops_launch / t / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": flops_launch / t / 1e12 / FP64_PEAK_TFLOPS}
    other = None
    if args.other_steps > 0:
        o = timed_run(deck_path, 3, args.other_steps, args.warmup, 0.0)
        other = {"corrected_passes": 3, "skin": 0.0, "value": o["value"], "ms_per_step": o["ms_per_step"],
                 "steps": args.other_steps, "kernel_ms": kernel_ms(o["ms"]),
                 "accuracy": "bit-exact (the reference's three passes and list order)"}
    res = {
        "metric": "FP64 particle-steps/s, batched vs plain @\u22641e-7 L2; adjoint/forward time ratio",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic Taylor-bar deck (seeded lattice + jitter)",
        "config": {"workload": "taylor_bar_corrected_f3dm10_eqos4_hopt1", "particles": n, "dx_mm": args.dx * 1e3,
                   "corrected_passes": args.passes, "nbr_skin": args.skin,
                   "skin_builds": head["skin_builds"], "steps_since_initial": args.warmup + args.steps,
                   "accuracy": ("the reference's neighbour sets; per-step fields within 1e-10, 1000-step L2 within "
                                "1e-7, energies within 1e-9 of the reference (tests/test_gpu_one_pass.py, "
                                "tests/test_gpu_skin.py)")
                   if args.passes == 1 or args.skin > 0 else "bit-exact with the reference",
                   "mean_neighbours": round(nbar, 2), "ghosts": sc1.ngp, "ptime_end": sc1.ptime,
                   "pair_interactions_per_s": n * nbar * args.steps / head["secs"],
                   "kernel_ms": kernel_ms(ms)},
        "roofline": {"bound": "hbm", "kernel": K["name"], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": bytes_launch, "launch_ms": kms[dom], "fp64": fp64},
        "other_mode": other,
    }
    if args.batch > 0:
        res["batched"] = batched_leg(args.batch, args.batch_seed, args.batch_steps, 4, 16, bool(args.batch_plain),
                                     args.passes)
    res["adjoint_forward_ratio"] = None
    if args.ident_config5 > 0:
        res["identification_at_size"] = ident_config5(args.ident_config5, 600, 40, args.ident_config5_iters,
                                                      args.ident_threads, 9)
    if args.ident > 0:
        res["identification"] = ident_leg(args.ident, args.ident_steps, args.ident_iters, args.ident_threads, 5)
    if args.adjoint_steps > 0:
        res["adjoint"] = adjoint_leg(args.adjoint_steps, args.adjoint_ckpt, args.adjoint_dx, 32, bool(args.adjoint_fd))
        # the whole gradient: the recorded forward plus the reverse sweep, over one plain forward
        res["adjoint_forward_ratio"] = res["adjoint"]["ratio_recorded_plus_reverse"]
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(tmpdir, args.dx)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
