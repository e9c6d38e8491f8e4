"""CPU check of the derivatives the adjoint takes of the per-particle
physics: tests/dual/dual_check.hip (built here with hipcc, run on the host,
no GPU) runs the model-10 chain of one constitutive update (rhoupd,
stress_rot1, f3dm10, sueos, nodalav, hieupd, eqos, stress_up, critts) with
Dual<4> seeds and compares the Jacobian with Richardson-extrapolated central
differences of the same templated code in double, over random elastic,
plastic, compressed and tensile states.  It also requires the dual value
part to equal the double chain bit for bit, so the dual chain follows the
forward's branches.  (A wrong quotient rule in mcm_dual.h gives errors of
order 1e6 here.)
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "dual", "dual_check.hip")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_dual_jacobian_of_the_constitutive_chain(tmp_path):
    exe = str(tmp_path / "dual_check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950", SRC, "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    cases = re.findall(r"^(\S+) worst (\S+) checked (\d+)$", out, re.M)
    assert len(cases) == 4, out
    for name, worst, checked in cases:
        assert int(checked) > 2000, (name, checked)
        assert float(worst) < 1e-4, (name, worst)
    m = re.search(r"value_mismatch (\d+)", out)
    assert m and int(m.group(1)) == 0, out
