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


def kernel_bytes_per_particle(nbar: float, corrected: bool) -> float:
    """Algorithmic HBM bytes per particle of the strain+constitutive kernel
    (f3dm10 + eqos4, h_opt = 1, tcrit_opt = 1): every per-particle array it
    must read or write is moved once (cache-ideal neighbour gathers), plus
    the 4-byte neighbour ids.
    reads : x(3) v(3) mass rho h | sigma(6) p e c q pcut efps fail vabs h0 rho0 = 23 words
    writes: x_half(3) rho h hold | sigma(6) p e c q pcut p_cut efps fail | A(6) = 27 words
    """
    words = 23 + 27
    return 8.0 * words + 4.0 * (nbar + 1)


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
    return {"value": d.np_ * steps / secs, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "sample": "same bar section and dx (0.073 mm), 1.0 mm long (%d particles), %d steps, %s"
                      % (d.np_, steps, "flang -O2 build of the reference" if kind == "reference" else "C oracle")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=7.3e-5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--corrected-one-pass", action="store_true",
                    help="accumulate the corrected kernels' three passes in one (rounding differs)")
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
    s = gpu.Solver(deck_path, corrected_one_pass=args.corrected_one_pass)
    s.initial()
    s.step(args.warmup)
    sc = s.scalars()
    # timed region: the step call enqueues on the context stream and returns
    # after synchronising it
    t0 = time.perf_counter()
    s.step(args.steps)
    t1 = time.perf_counter()
    sc1 = s.scalars()
    secs = t1 - t0
    n = d.np_
    value = n * args.steps / secs
    # per-kernel-group durations with HIP events on the context stream
    ms = s.time_steps(10)
    sc2 = s.scalars()
    nbar = float(sc2.max_count)
    # mean neighbour count from the lists
    counts, gcounts, _ = s.neighbour_lists()
    nbar = float((counts + gcounts).mean())
    bpp = kernel_bytes_per_particle(nbar, True)
    achieved = bpp * n / (ms[1] * 1e-3) / 1e9
    res = {
        "metric": "FP64 particle-steps/s (plain per-instance launch)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic Taylor-bar deck (seeded lattice + jitter)",
        "config": {"workload": "taylor_bar_corrected_f3dm10_eqos4_hopt1", "particles": n, "dx_mm": args.dx * 1e3,
                   "corrected_passes": 1 if args.corrected_one_pass else 3,
                   "mean_neighbours": round(nbar, 2), "ghosts": sc1.ngp, "ptime_end": sc1.ptime,
                   "kernel_ms": {"neighbours": ms[0], "strain_constitutive": ms[1], "momentum_move": ms[2],
                                 "step": ms[3]}},
        "roofline": {"bound": "hbm", "kernel": "k_strain_const", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_particle": bpp},
    }
    s.close()
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(tmpdir, args.dx)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
