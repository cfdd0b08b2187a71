"""Device variograms (ss3d_variogram) against the oracle's vario_lag (f90:405-431) and the
reference's own --vario tables, on host FP64 fields and device FP32 fields."""
import ctypes
import sys

import numpy as np
import pytest

import golden_cases as G
import oracle_ffi as O
import specsim3d_amd as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_gam(f, dirs):
    nreal = f.shape[0]
    nlagmax = max(1, max(d[3] for d in dirs))
    g = np.zeros((nreal, len(dirs), nlagmax))
    for r in range(nreal):
        for k, (sx, sy, sz, nlag) in enumerate(dirs):
            for l in range(1, nlag + 1):
                g[r, k, l - 1] = O.vario_lag(f[r], l * sx, l * sy, l * sz)
    return g


@pytest.fixture(scope="module")
def eng():
    with S.Engine(0) as e:
        yield e


@pytest.mark.parametrize("name", sorted(G.vario_cases()))
def test_fp64_field_matches_oracle_and_reference_table(eng, name):
    say(f"[vario] {name}")
    c = G.vario_cases()[name]
    f = G.vario_field(name)
    gam = eng.variogram(f, c["dirs"])
    ref = oracle_gam(f, c["dirs"])
    assert np.allclose(gam, ref, rtol=1e-12, atol=0.0), np.abs(gam - ref).max()
    # the reference's table: mean over realizations, and per-realization columns
    _, rows = G.vario_table(name)
    for row in rows:
        d, il = int(row[0]), int(row[4])
        gm = gam[:, d - 1, il - 1]
        assert gm.mean() == pytest.approx(float(row[8]), rel=6e-7)
        if c["iper"]:
            for r in range(c["nreal"]):
                assert gm[r] == pytest.approx(float(row[10 + r]), rel=6e-7)


def test_device_fp32_field(eng):
    sts = [(2, 0.8, 14, 9, 5, 30, 15, 0)]
    grid = (48, 40, 24)
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 1)
    say("[vario] device fp32", grid)
    sp = eng.spectrum(sts, 0.2, 1, 1, 1, *npg)
    f32, _, _, _ = eng.realize(sp, 4242, 1, 0, 3, *grid, S.FP32)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), f32.nbytes) == 0
    try:
        assert hip.hipMemcpy(p, f32.ctypes.data, f32.nbytes, 1) == 0
        dirs = [(1, 0, 0, 40), (0, 1, 0, 39), (0, 0, 1, 23), (1, 1, -1, 20), (-2, 1, 0, 12)]
        gam = eng.variogram(p.value, dirs, device_shape=f32.shape, precision=S.FP32)
    finally:
        hip.hipFree(p)
        sp.close()
    ref = oracle_gam(f32.astype(np.float64), dirs)
    assert np.allclose(gam, ref, rtol=1e-12, atol=0.0), np.abs(gam - ref).max()
    host = eng.variogram(f32, dirs)
    assert np.array_equal(host, gam)


def test_edge_cases_and_symmetry(eng):
    rng = np.random.default_rng(3)
    f = rng.standard_normal((2, 6, 7, 9))
    dirs = [(0, 0, 0, 5), (1, 0, 0, 0), (9, 0, 0, 3), (0, 7, 1, 2), (1, -1, 1, 8), (-1, 1, -1, 8), (3, 2, 1, 4)]
    gam = eng.variogram(f, dirs)
    assert np.all(gam[:, 0] == 0) and np.all(gam[:, 1] == 0) and np.all(gam[:, 2] == 0) and np.all(gam[:, 3] == 0)
    assert np.allclose(gam[:, 4], gam[:, 5], rtol=1e-14)
    assert np.allclose(gam, oracle_gam(f, dirs), rtol=1e-12, atol=0.0)
    assert np.array_equal(gam, eng.variogram(f, dirs))  # fixed summation order
    # one plane, one row, one column
    for shape in [(1, 1, 5, 8), (1, 1, 1, 8), (1, 6, 1, 1)]:
        g = rng.standard_normal(shape)
        d = [(1, 0, 0, 9), (0, 1, 0, 4), (0, 0, 1, 7), (1, 1, 1, 3)]
        assert np.allclose(eng.variogram(g, d), oracle_gam(g, d), rtol=1e-12, atol=0.0)
