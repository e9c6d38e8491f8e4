"""The Fortran host (ISO_C_BINDING, mcmtop_amd/fortran) drives the GPU step
through the same-named stage subroutines and through the fused step; both
agree with the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from dumpio import Dump
from oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mcmtop_amd", "fortran", "mcm_gpu_run")
DATA = os.path.join(ROOT, "tests", "golden", "data")


@pytest.mark.parametrize("mode", ["stages", "fused"])
def test_fortran_host_runs_the_gpu_step(mode, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built (__graft_entry__.build())")
    deck = os.path.join(DATA, "corr_f3dm10_eos4_hopt1.bin")
    dump = str(tmp_path / "f.bin")
    out = subprocess.run([EXE, deck, "15", mode, dump], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mcm_gpu_run: np=96 steps=15" in out.stdout
    o = Oracle(deck)
    o.initial()
    o.run(15)
    od, gd = o.state(15, 10), Dump(dump)
    for k in ("x", "v", "sigma", "e", "rho"):
        assert np.linalg.norm(gd.f(k) - od.f(k)) <= 1e-10 * np.linalg.norm(od.f(k)), k
