"""The pair loops' fast path (mcmtop_amd/csrc/mcm_physics.h) replaces the
device's division by Divisor::quot with a shared reciprocal and its square
root by sqrt_noscale, without the scaling steps, inside operand ranges that
the path guarantees.  tests/gpu_checks/quot_check.hip compares both with
the device's own a / b and sqrt, bit for bit, on random operands spread over
those ranges (random exponents and significands, signed zeros)."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_checks", "libquot_check.so")
CASES = {
    0: "corrected quotients: {0} u +-[2^-502, 2^335] over [2^-200, 2^200]",
    1: "Divisor range: +-[2^-300, 2^300] and +-0 over [2^-300, 2^300]",
    2: "reciprocals of [2^-300, 2^300]",
    3: "square roots of 0 and [2^-767, 2^1023]",
}


@pytest.mark.parametrize("which", sorted(CASES))
def test_fast_path_arithmetic_equals_the_device_division_and_sqrt(which):
    if not os.path.exists(LIB):
        pytest.fail("tests/gpu_checks/libquot_check.so not built (__graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    lib.quot_check.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_double)]
    out = (ctypes.c_double * 5)()
    assert lib.quot_check(which, 1 << 26, 0x5eed + which, out) == 0
    assert out[0] == 0, "%s: %d mismatches, first a=%r b=%r got %r want %r" % (
        CASES[which], int(out[0]), out[1], out[2], out[3], out[4])


def test_cap_model_expf_on_the_device_equals_the_c_library():
    """The cap model's single-precision exp as the device evaluates it
    (cap_expf, mcmtop_amd/csrc/mcm_cap.h) against the host C library's expf,
    which the reference's flang build calls: equal bits (or both NaN) for all
    2^32 float bit patterns (tests/gpu_checks/expf_check.hip)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_checks", "libexpf_check.so")
    if not os.path.exists(path):
        pytest.fail("tests/gpu_checks/libexpf_check.so not built (__graft_entry__.build())")
    chk = ctypes.CDLL(path)
    chk.expf_check_device.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_ulonglong)]
    out = (ctypes.c_ulonglong * 2)()
    assert chk.expf_check_device(0, 1 << 32, 16, out) == 0
    assert out[0] == 0, "%d mismatches, first pattern 0x%08x" % (out[0], out[1])
