"""FP32-only spectra (SS3D_OPT_SPEC_FP32; the request loop's and the batch's FP32 work): the passes
after the x pass transform the columns kx <= npx/2 alone, by the same operations as a full build, so
the FP32 amplitudes -- and every FP32 field drawn from them -- equal a full build's bit for bit.  The
density of a real covariance is Hermitian, S(-k) = conj S(k): the maxima of |Re S| and |Im S| over the
kept columns are those of the whole grid, and the sums count the mirrored columns twice, so the
diagnostics agree with a full build's to rounding and sum(S) = C(0) still holds.  FP64 realizations
and saving need the FP64 amplitudes and are refused.
Periodic grids: 512^3 (the benchmark), 1024x1024x512 (the large job: x axis 2^10, 513 kept columns),
a 3-D Gaussian-model grid, npz = 1 (the y pass is the last one), npy = 1 (x then z), npx = 2 (no
column to leave out), npx = 4 and 32 (kept columns 3 and 17: rows of 4 and 24 values), a 1-D grid
(x pass alone) and the generic FP64 passes (full builds)."""
import sys

import numpy as np
import pytest

import specsim3d_amd as S

pytestmark = pytest.mark.gpu

BENCH = [(2, 0.7, 80.0, 40.0, 10.0, 30.0, 15.0, 0.0)]
GAUSS = [(3, 0.9, 12.0, 8.0, 5.0, 120.0, -20.0, 0.0)]
SPH = [(1, 0.8, 14.0, 9.0, 5.0, 30.0, 15.0, 0.0), (2, 0.1, 20.0, 20.0, 20.0, 0.0, 0.0, 0.0)]
LARGE = [(3, 1.0, 150.0, 60.0, 20.0, 120.0, -20.0, 0.0)]  # the 1024x1024x512 large-job model

# (structures, nugget, output grid, periodic grid, realizations, igauss, generic FP64 passes)
CASES = [
    (BENCH, 0.3, (256, 256, 256), (512, 512, 512), 2, 1, 0),
    (LARGE, 0.0, (512, 512, 256), (1024, 1024, 512), 1, 1, 0),
    (GAUSS, 0.1, (40, 30, 20), (64, 64, 32), 3, 0, 0),
    (SPH, 0.0, (64, 16, 1), (128, 128, 1), 4, 1, 0),
    (SPH, 0.2, (40, 1, 20), (128, 1, 64), 3, 1, 0),
    (SPH, 0.1, (2, 8, 8), (2, 16, 16), 2, 1, 0),
    (SPH, 0.1, (3, 8, 8), (4, 16, 16), 2, 0, 0),
    (GAUSS, 0.2, (20, 12, 10), (32, 16, 16), 3, 1, 0),
    (BENCH, 0.1, (40, 1, 1), (256, 1, 1), 2, 0, 0),
    (GAUSS, 0.0, (24, 20, 12), (64, 64, 32), 2, 1, 1),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    with S.Engine(0) as e:
        yield e


@pytest.mark.parametrize("sts,nugget,grid,npg,nreal,igauss,generic", CASES)
def test_fp32_only_spectrum_equals_full(eng, sts, nugget, grid, npg, nreal, igauss, generic):
    say(f"[spec32] grid {grid} periodic {npg} nreal {nreal} igauss {igauss} generic {generic}")
    eng.set_option(S.OPT_FP64_GENERIC, generic)
    try:
        full = eng.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
        eng.set_option(S.OPT_SPEC_FP32, 1)
        try:
            half = eng.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
        finally:
            eng.set_option(S.OPT_SPEC_FP32, 0)
    finally:
        eng.set_option(S.OPT_FP64_GENERIC, 0)
    try:
        f1, m1, v1, _ = eng.realize(full, 4711, igauss, 5, nreal, *grid, precision=S.FP32)
        f2, m2, v2, _ = eng.realize(half, 4711, igauss, 5, nreal, *grid, precision=S.FP32)
        assert np.array_equal(f1, f2) and np.array_equal(m1, m2) and np.array_equal(v1, v2)
        d1, d2 = full.diag, half.diag
        say(f"[spec32] full {d1}\n[spec32] fp32-only {d2}")
        assert d2["c0"] == d1["c0"]
        assert d2["max_real"] == pytest.approx(d1["max_real"], rel=1e-12)
        assert d2["spos"] == pytest.approx(d1["spos"], rel=1e-12)
        assert d2["sneg"] == pytest.approx(d1["sneg"], rel=1e-9, abs=1e-12 * d1["spos"])
        assert d2["spos"] + d2["sneg"] == pytest.approx(d1["c0"], abs=1e-9)
        # S(-k) = conj S(k): the mirrored columns hold the same magnitudes, up to rounding
        assert d2["max_imag"] == pytest.approx(d1["max_imag"], rel=1e-6, abs=1e-15 * d1["max_real"])
        with pytest.raises(S.SS3DError):
            eng.realize(half, 4711, igauss, 0, 1, *grid, precision=S.FP64)
    finally:
        full.close()
        half.close()


def test_fp32_only_spectrum_not_saved(eng, tmp_path):
    eng.set_option(S.OPT_SPEC_FP32, 1)
    try:
        sp = eng.spectrum(SPH, 0.1, 1.0, 1.0, 1.0, 64, 32, 16)
    finally:
        eng.set_option(S.OPT_SPEC_FP32, 0)
    try:
        with pytest.raises(S.SS3DError):
            sp.save(str(tmp_path / "s.spec"))
    finally:
        sp.close()
