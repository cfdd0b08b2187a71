"""The z-run layout of the FP32 passes (SS3D_OPT_ZRUN, on by default): where the z and y axes have
length 2^9, the z fill runs wave-local (its waves exchange without block barriers) and writes runs
of 8 z per pair of kx columns, the y pass reads and writes that layout, and the C2R pass reads
whole runs and moves them to rows through LDS.  Every column and row is transformed by the same
arithmetic as in the row layouts, so the fields equal those of SS3D_OPT_ZRUN = 0 bit for bit; the
statistics' block sums group other rows (a run of 8 z instead of consecutive y), so mean and
variance agree to rounding.  The cases cover both coefficient laws, the fused schedule (several
batches) and the separate fill and y launches (one batch), even and odd cropped nx (8-byte or
4-byte output stores), the C2R plans and lengths that hold whole runs (8 to 32 rows per block,
x axes of 256 to 2048 points), the two-stream pipeline and the 256^3 benchmark grid.
"""
import sys

import numpy as np
import pytest

import specsim3d_amd as S

pytestmark = pytest.mark.gpu

BENCH = [(2, 0.7, 80.0, 40.0, 10.0, 30.0, 15.0, 0.0)]  # exponential sill 0.7, ranges 80/40/10
SMALLX = [(1, 1.0, 10.0, 60.0, 60.0, 20.0, 0.0, 0.0)]  # short x range: a 256-point x axis
THIN = [(1, 1.0, 200.0, 60.0, 200.0, 0.0, 0.0, 0.0)]  # long y and z ranges: thin crops in 256x512x512
WIDE = [(1, 1.0, 200.0, 400.0, 200.0, 0.0, 0.0, 0.0)]  # and a long x range: 1024- and 2048-point x axes

# (structures, nugget, output grid, realizations, igauss, z/y radix plan, C2R radix plan, pipeline);
# plans None: the untuned default (largest radix 2^5).  Periodic grids 512^3 (the benchmark) and
# 256x512x512; nz 264 keeps more than half the z axis (no pruned last pass); 100x32x16 keeps two
# runs of z and one C2R tile of y; 300x32x16 and 1001x20x16 in 1024x512x512 and 2048x512x512 run
# the C2R pass at 2^9 and 2^10 (plan 5: 16 and 8 rows per block)
CASES = [
    (BENCH, 0.3, (256, 256, 256), 17, 1, 3, 5, 0),
    (SMALLX, 0.1, (128, 248, 256), 9, 0, None, None, 0),
    (SMALLX, 0.1, (127, 256, 200), 3, 1, 4, 3, 0),
    (SMALLX, 0.0, (130, 248, 264), 20, 1, 3, 4, 0),
    (SMALLX, 0.2, (128, 256, 256), 20, 0, 3, 5, 1),
    (THIN, 0.1, (100, 32, 16), 20, 1, 3, 5, 0),
    (WIDE, 0.1, (300, 32, 16), 20, 0, 3, 5, 0),
    (WIDE, 0.0, (1001, 20, 16), 18, 1, 4, 5, 0),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    with S.Engine(0) as e:
        yield e


@pytest.mark.parametrize("sts,nugget,grid,nreal,igauss,rmzy,rmx,pipe", CASES)
def test_zrun_equals_row_layout(eng, sts, nugget, grid, nreal, igauss, rmzy, rmx, pipe):
    nx, ny, nz = grid
    npg = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    assert npg[1] == npg[2] == 512 and npg[0] in (256, 512, 1024, 2048), npg
    say(f"[zrun] grid {grid} periodic {npg} nreal {nreal} igauss {igauss} plans {rmzy}/{rmx} pipeline {pipe}")
    sp = eng.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
    lx = int(np.log2(npg[0])) - 1
    plans = [(6, 9, rmzy), (7, 9, rmzy), (8, lx, rmx)]
    try:
        for kind, logn, rm in plans:
            eng.tuning_set(kind, logn, rm or 5)
        eng.set_option(S.OPT_PIPELINE, pipe)
        got = {}
        for opt in (1, 0):
            eng.set_option(S.OPT_ZRUN, opt)
            got[opt] = eng.realize(sp, 4711, igauss, 3, nreal, nx, ny, nz, precision=S.FP32)
    finally:
        eng.set_option(S.OPT_ZRUN, 1)
        eng.set_option(S.OPT_PIPELINE, 0)
        for kind, logn, _ in plans:
            eng.tuning_set(kind, logn, 5)
        sp.close()
    (f1, m1, v1, i1), (f0, m0, v0, i0) = got[1], got[0]
    assert np.array_equal(f1, f0)
    assert np.allclose(m1, m0, rtol=0, atol=1e-12) and np.allclose(v1, v0, rtol=1e-12, atol=0)
    assert np.array_equal(i1, i0)
    assert np.all(np.isfinite(f1)) and np.all(v1 > 0)


def test_zrun_matches_fp64(eng):
    """The z-run FP32 fields against the FP64 drop-in path on the same draws: <= 1e-5 relative RMS."""
    sts, nugget, (nx, ny, nz), _, igauss, _, _, _ = CASES[1]
    npg = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    say(f"[zrun] fp32 vs fp64 {npg}")
    sp = eng.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
    try:
        f32 = eng.realize(sp, 99, igauss, 0, 2, nx, ny, nz, precision=S.FP32)[0]
        f64 = eng.realize(sp, 99, igauss, 0, 2, nx, ny, nz, precision=S.FP64)[0]
    finally:
        sp.close()
    for r in range(2):
        rel = np.sqrt(np.mean((f32[r] - f64[r]) ** 2) / np.mean(f64[r] ** 2))
        say(f"[zrun] realization {r}: rel rms {rel:.3e}")
        assert rel <= 1e-5
