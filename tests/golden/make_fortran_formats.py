"""Fixture of Fortran formatted output for the VTK writer's number formats.

Compiles fortran_formats.f90 with flang (ROCm's, in this container), feeds
it seeded doubles of the magnitudes the writer meets (coordinates,
velocities, stresses, energies), edge cases (signed zeros, values that
round to zero, exact decimal ties, F15.8 overflow, three-digit exponents)
and stores the values with flang's output in data/fortran_formats.json.
"""
import json
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def values():
    rng = np.random.default_rng(2024)
    v = [0.0, -0.0, 0.125, -0.125, 1.03125, -1.09375, 2.53125, 0.5e-8, -1.5e-9, 4.999999995e-9,
         99999.99999999, 999999.99999999, 1.0e6, -1.0e5, -99999.9999, 1.0e150, -2.5e-120, 1.0e-99, 9.99995e99,
         9.999995, 0.999995, 123456.5, 1.0e-320, 3.4e38, 3.5e38, 1.0e-40,
         float("inf"), float("-inf"), float("nan")]
    v += list(rng.uniform(-0.06, 0.06, 300))
    v += list(rng.uniform(-300.0, 300.0, 200))
    v += list(np.sign(rng.uniform(-1, 1, 300)) * 10.0 ** rng.uniform(-3, 10, 300))
    v += list(np.sign(rng.uniform(-1, 1, 100)) * 10.0 ** rng.uniform(-200, 200, 100))
    return [float(x) for x in v]


def main():
    vals = values()
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "ff")
        subprocess.run([FLANG, "-O0", os.path.join(HERE, "fortran_formats.f90"), "-o", exe], check=True)
        inp = os.path.join(d, "in.bin")
        with open(inp, "wb") as f:
            f.write(struct.pack("<i", len(vals)))
            f.write(struct.pack("<%dd" % len(vals), *vals))
        out = subprocess.run([exe, inp], check=True, capture_output=True, text=True).stdout.split("\n")
    n = len(vals)
    rec = {"values_hex": [struct.pack("<d", x).hex() for x in vals],
           "f15_8": out[0:3 * n:3], "e24_5": out[1:3 * n:3], "e24_5_real4": out[2:3 * n:3],
           "i6": out[3 * n:5 * n:2], "i5": out[3 * n + 1:5 * n:2]}
    with open(os.path.join(HERE, "data", "fortran_formats.json"), "w") as f:
        json.dump(rec, f, indent=0)
    print("wrote", n, "values")


if __name__ == "__main__":
    main()
