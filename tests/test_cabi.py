"""The C-ABI library loads on a machine without a GPU and exports exactly
the entry points include/mcm_gpu.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from mcmtop_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mcm_gpu.h")).read()
    return sorted(set(re.findall(r"\b(mcm_(?:gpu|vtk)_[a-z_0-9]+)\s*\(", text)))


def test_header_and_python_mirror_agree():
    assert _declared() == sorted(gpu.SYMBOLS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libmcm_gpu.so not built")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(gpu.McmGpuError):
        gpu._LIB = None
        gpu.load_library(str(tmp_path / "absent.so"))
    gpu._LIB = None


def test_bench_refuses_more_than_one_gpu():
    """The step runs on one GPU: bench.py --gpus 2 exits with a message
    before it touches a device."""
    import subprocess
    import sys
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr and out.stdout == ""
