"""Writes tests/golden/fmt/{values.npy, es_golden.txt}: values and their Fortran ES text.

The values are printed by es_golden.f90 compiled with the same Fortran compiler as the
reference build (amdflang), with the edit descriptors of the reference's writers: es15.7 (.out,
src/specsim3d.f90:1393), es13.5 (ASCII .vtk, f90:1411), es16.8 (.vti / .vtk headers,
f90:330,376), es14.6 (variogram table, f90:489-518).  The set covers exact decimal ties of
float32 and float64 values at each precision, neighbours of powers of ten and of rounding
boundaries, signed zeros, non-finite values, subnormals, 3-digit exponents and random values.

Usage: python tests/golden/fmt/make_fmt_golden.py
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "amdflang")


def values():
    rng = np.random.default_rng(20261203)
    v = [0.0, -0.0, np.nan, np.inf, -np.inf, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 1e-300,
         1e300, 1e-100, 1e100, 9.99999995, 9.999999949, 123456785.0, 0.5, 1.0, -1.0, 0.1, 0.3]
    # exact decimal ties: 1 + i/2^j has j decimals, so ties fall at several precisions
    for j in range(5, 12):
        for i in range(1, 2 ** j, max(1, 2 ** j // 64)):
            for t in (-20, -7, -1, 0, 3, 9):
                v.append((1.0 + i / 2.0 ** j) * 2.0 ** t)
                v.append(-(1.0 + i / 2.0 ** j) * 2.0 ** t)
    # neighbours of powers of ten and of the rounding boundaries d.ddddddd5 x 10^k
    for k in range(-25, 26):
        p = 10.0 ** k
        v += [p, np.nextafter(p, 0), np.nextafter(p, np.inf)]
        for P in (6, 7, 8, 9):
            b = (10.0 ** P - 0.5) * 10.0 ** (k - P + 1)
            v += [b, np.nextafter(b, 0), np.nextafter(b, np.inf)]
            b = 1.2345675 * p
            v += [b, np.nextafter(b, 0), np.nextafter(b, np.inf)]
    # random values: wide exponents, and float32-exact Gaussian values as FP32 fields hold
    v += list(rng.standard_normal(1500) * 10.0 ** rng.uniform(-30, 30, 1500))
    v += list(rng.standard_normal(1500).astype(np.float32).astype(np.float64))
    v += list((rng.standard_normal(200) * 10.0 ** rng.uniform(-320, -300, 200)))  # subnormals
    return np.array(v, dtype=np.float64)


def main():
    v = values()
    with tempfile.TemporaryDirectory() as wd:
        exe = os.path.join(wd, "es_golden")
        subprocess.run([FC, "-O2", os.path.join(HERE, "es_golden.f90"), "-o", exe], check=True)
        fin = os.path.join(wd, "values.bin")
        with open(fin, "wb") as fh:
            fh.write(np.int32(v.size).tobytes())
            fh.write(v.tobytes())
        out = subprocess.run([exe, fin], check=True, capture_output=True, text=True).stdout
    lines = out.splitlines()
    assert len(lines) == v.size
    np.save(os.path.join(HERE, "values.npy"), v)
    with open(os.path.join(HERE, "es_golden.txt"), "w") as fh:
        fh.write(out)
    print(f"{v.size} values")


if __name__ == "__main__":
    main()
