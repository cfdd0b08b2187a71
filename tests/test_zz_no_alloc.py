"""ss3d_realize_device_noalloc (the realization path with no device allocation) against
ss3d_realize_device, through the C ABI: fields and statistics bit for bit, FP32 and FP64, both
coefficient laws, two calls in a row (the second on a smaller grid and fewer realizations, so the
scratch array keeps the first call's contents), the z-run layout of the FP32 passes, the
statistics-free form, and the error paths
(a grid whose one realization exceeds the scratch array; random-stream tables not built)."""
import sys

import numpy as np
import pytest

import specsim3d_amd as S
from hipmem import DevBuf

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = S.Engine(0)
    yield e
    e.close()


def run_both(eng, sp, iseed, igauss, r0, nreal, grid, prec, with_stats=True):
    nx, ny, nz = grid
    n = nx * ny * nz
    esz, dt = (8, np.float64) if prec == S.FP64 else (4, np.float32)
    outs = []
    for fn in (eng.realize_device, eng.realize_device_noalloc):
        with DevBuf(nreal * n * esz) as out, DevBuf(nreal * 3 * 8) as st:
            fn(sp, iseed, igauss, r0, nreal, nx, ny, nz, prec, out.ptr, st.ptr if with_stats else None)
            eng.synchronize()
            outs.append((out.get(dt, nreal * n), st.get(np.float64, nreal * 3) if with_stats else None))
    return outs


CASES = [
    # (model, nugget, grid, ipad, nreal), the second of each pair smaller than the first
    ([(2, 0.7, 20, 10, 5, 30, 15, 0)], 0.3, (60, 50, 50), 1, 40),
    ([(3, 1.0, 8, 6, 4, 120, -20, 0)], 0.0, (30, 20, 12), 1, 7),
]


@pytest.mark.parametrize("prec", [S.FP32, S.FP64])
@pytest.mark.parametrize("igauss", [0, 1])
def test_noalloc_equals_allocating_path(eng, prec, igauss):
    for k, (sts, nug, grid, ipad, nreal) in enumerate(CASES):
        npg = S.periodic_grid(sts, 1, 1, 1, *grid, ipad)
        say(f"[noalloc] prec {prec} igauss {igauss} call {k + 1} grid {grid} periodic {npg} nreal {nreal}")
        sp = eng.spectrum(sts, nug, 1, 1, 1, *npg)
        try:
            (fa, sa), (fb, sb) = run_both(eng, sp, 4242 + k, igauss, 3, nreal, grid, prec)
        finally:
            sp.close()
        assert np.array_equal(fa, fb), np.abs(fa - fb).max()
        assert np.array_equal(sa, sb)


# crops that keep few rows of 256x512x512 periodic grids (z and y axes of 2^9): the FP32 passes run
# in the z-run layout (wave-local fill, runs of 8 z per kx pair) inside the scratch array's windows;
# two calls in a row, the second smaller
ZRUN_STS = [(1, 1.0, 200.0, 60.0, 200.0, 0.0, 0.0, 0.0)]
ZRUN_CASES = [((100, 32, 16), 20), ((130, 20, 24), 9)]


def test_noalloc_zrun_layout(eng):
    for k, (grid, nreal) in enumerate(ZRUN_CASES):
        npg = S.periodic_grid(ZRUN_STS, 1, 1, 1, *grid, 1)
        assert npg == (256, 512, 512), npg
        say(f"[noalloc] z-run layout, call {k + 1} grid {grid} periodic {npg} nreal {nreal}")
        sp = eng.spectrum(ZRUN_STS, 0.1, 1, 1, 1, *npg)
        try:
            (fa, sa), (fb, sb) = run_both(eng, sp, 777 + k, 1, 2, nreal, grid, S.FP32)
        finally:
            sp.close()
        assert np.array_equal(fa, fb), np.abs(fa - fb).max()
        assert np.array_equal(sa, sb)


def test_noalloc_without_statistics(eng):
    sts, nug, grid, ipad, nreal = CASES[1]
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, ipad)
    sp = eng.spectrum(sts, nug, 1, 1, 1, *npg)
    try:
        (fa, _), (fb, _) = run_both(eng, sp, 99, 1, 0, nreal, grid, S.FP32, with_stats=False)
    finally:
        sp.close()
    assert np.array_equal(fa, fb)


def test_noalloc_refuses_what_it_cannot_hold(eng):
    # 200^3 in 256^3, FP64: one realization's two pass buffers need ~374 MB (> 256 MiB)
    sts = [(1, 1.0, 10, 10, 10, 0, 0, 0)]
    grid = (200, 200, 200)
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 0)
    assert npg == (256, 256, 256)
    sp = eng.spectrum(sts, 0.0, 1, 1, 1, *npg)
    try:
        with DevBuf(200 ** 3 * 8) as out:
            # random-stream tables not built yet for this spectrum
            with pytest.raises(S.SS3DError) as ei:
                eng.realize_device_noalloc(sp, 1, 0, 0, 1, *grid, S.FP64, out.ptr)
            assert ei.value.code == 1
            eng.realize_device(sp, 1, 0, 0, 1, *grid, S.FP64, out.ptr)  # builds them
            eng.synchronize()
            with pytest.raises(S.SS3DError) as ei:
                eng.realize_device_noalloc(sp, 1, 0, 0, 1, *grid, S.FP64, out.ptr)
            assert ei.value.code == 2
            # the FP32 passes of the same grid fit
            eng.realize_device_noalloc(sp, 1, 0, 0, 1, *grid, S.FP32, out.ptr)
            eng.synchronize()
    finally:
        sp.close()
