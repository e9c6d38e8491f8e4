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
@pytest.mark.parametrize("name", ["corr_f3dm10_eos4_hopt1", "penalty_contact_f3dm1", "xsph_corr_f3dm10_eos4_sym"])
def test_fortran_host_runs_the_gpu_step(name, mode, tmp_path):
    """The loop of mcm_solution through the stage subroutines (mcm_rep_cont
    included) and the fused step; penalty contact and XSPH decks too."""
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built (__graft_entry__.build())")
    deck = os.path.join(DATA, name + ".bin")
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


def test_fortran_host_reads_the_reference_text_deck(tmp_path):
    """The host passes the reference's own version-2 input deck to the
    library, as mcm_getinput reads it."""
    import cases
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built (__graft_entry__.build())")
    txt = str(tmp_path / "bar.mcm")
    cases.build("corr_f3dm10_eos4_hopt1").write_mcm(txt)
    dump = str(tmp_path / "f.bin")
    out = subprocess.run([EXE, txt, "15", "fused", dump], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    o = Oracle(os.path.join(DATA, "corr_f3dm10_eos4_hopt1.bin"))
    o.initial()
    o.run(15)
    od, gd = o.state(15, 10), Dump(dump)
    for k in ("x", "v", "sigma", "e", "rho"):
        assert np.array_equal(gd.f(k), od.f(k)), k
