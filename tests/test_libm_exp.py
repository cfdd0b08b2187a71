"""The FP64 exp of the device covariance (specsim3d_amd/csrc/ss3d_exp.h) against glibc's exp.

The reference's covariance calls the C library's exp (src/specsim3d.f90:153-178); the device
covariance restates glibc's method so that the covariance grid, and hence the spectral density
of the exponential and Gaussian models, is the reference's bit for bit.  Checked here on the CPU:
the header built for the host gives glibc's results bit for bit on 4.4e7 arguments covering the
covariance's range (-3 h and -3 h^2 <= 0), the subnormal results, the rescaled path (|x| >= 512),
tiny arguments, positive arguments and the special values; and the 2^(k/128) table equals its
definition at 300 bits (mpmath).
"""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "specsim3d_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("libm_exp") / "libm_exp_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "libm_exp_check.cpp"), "-lm"], check=True)
    return exe


@pytest.mark.parametrize("n,lo,hi", [
    (12_000_000, -40.0, 0.0),          # exponential and Gaussian covariances of ordinary lags
    (4_000_000, -1.0, 0.0),
    (4_000_000, -512.0, -40.0),
    (4_000_000, -708.4, -512.0),       # rescaled path, normal results
    (4_000_000, -745.2, -708.3),       # subnormal results
    (4_000_000, -1100.0, -745.0),      # underflow to zero
    (4_000_000, -3e-16, 3e-16),        # tiny arguments
    (4_000_000, 0.0, 709.8),           # positive arguments (overflow at the top)
    (4_000_000, -0.01, 0.01),
])
def test_equals_glibc(checker, n, lo, hi):
    out = subprocess.run([checker, str(n), repr(lo), repr(hi)], check=True, capture_output=True, text=True).stdout
    m = re.search(r"n (\d+) mismatches (\d+) special (\d+)", out)
    assert m, out
    assert int(m.group(2)) == 0 and int(m.group(3)) == 0, out


def test_table_is_its_definition():
    mpmath = pytest.importorskip("mpmath")
    import sys

    sys.path.insert(0, CSRC)
    try:
        import gen_exp_table
    finally:
        sys.path.pop(0)
    with open(os.path.join(CSRC, "ss3d_exp_table.inc")) as fh:
        committed = [int(v, 16) for v in re.findall(r"0x([0-9a-f]{16})ull", fh.read())]
    expect = [v for pair in gen_exp_table.table() for v in pair]
    assert committed == expect
    assert mpmath  # the table was recomputed at 300 bits
