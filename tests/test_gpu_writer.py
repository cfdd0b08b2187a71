"""Streaming output writer on the GPU (ss3d_writer_realize): files written while realizations are
computed must equal, byte for byte, the files the same writer makes from ss3d_realize's host
fields (whose host path is pinned to the reference's own files by tests/test_writer.py), in FP64
and FP32, with every file kind, with the device cast to float32 (no text output) and with the
all-realizations copy (GSLIB output), over many calls so the device and host slot rings are
reused.  Statistics equal ss3d_realize's; variograms of the device chunks equal those of the
host field."""
import os
import sys

import numpy as np
import pytest

import specsim3d_amd as S

pytestmark = pytest.mark.gpu

STS = [(2, 0.8, 14, 9, 5, 30, 15, 0), (3, 0.2, 30, 20, 10, 0, 0, 0)]
NUGGET = 0.05
GRID = (24, 20, 12)
ORG, SPC = (0.5, -1.0, 2.0), (1.0, 1.5, 0.5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    with S.Engine(0) as e:
        yield e


@pytest.fixture(scope="module")
def spec(eng):
    npg = S.periodic_grid(STS, *SPC, *GRID, 1)
    sp = eng.spectrum(STS, NUGGET, *SPC, *npg)
    yield sp
    sp.close()


def paths(d, base, kinds, nreal):
    p = {}
    if "bin" in kinds:
        p["bin"] = os.path.join(d, base + ".bin")
    if "vti" in kinds:
        p["vti"] = os.path.join(d, base + ".vti")
    if "vtk" in kinds:
        p["vtk_base"] = os.path.join(d, base)
        p["vtk_count"] = nreal
    if "out" in kinds:
        p["gslib"] = os.path.join(d, base + ".out")
    if "vtk_ascii" in kinds:
        p["vtk_ascii"] = os.path.join(d, base + "_1.vtk")
    return p


def files_of(p, nreal):
    f = [p[k] for k in ("bin", "vti", "gslib", "vtk_ascii") if k in p]
    f += [f"{p['vtk_base']}_{r + 1}.vtk" for r in range(p.get("vtk_count", 0))]
    return f


@pytest.mark.parametrize("prec,kinds,calls,iseed,igauss", [
    (S.FP64, ("bin", "vti", "vtk"), (1, 1, 1, 1, 1, 1, 1), 4711, 0),       # device cast, rings reused
    (S.FP64, ("bin", "vti", "vtk", "out", "vtk_ascii"), (2, 3, 2), 99, 1),  # all-realizations copy
    (S.FP64, ("out",), (7,), -5, 1),
    (S.FP32, ("bin", "vti", "vtk", "out"), (1, 2, 1, 3), 1234, 1),
    (S.FP32, ("bin", "vtk_ascii"), (3, 3, 1), 77, 0),
])
def test_streamed_files_equal_host_written(tmp_path, eng, spec, prec, kinds, calls, iseed, igauss):
    nreal = sum(calls)
    say(f"[writer] prec={prec} kinds={kinds} calls={calls}")
    field, mean, var, mi = eng.realize(spec, iseed, igauss, 0, nreal, *GRID, precision=prec)
    ref = paths(str(tmp_path), "host", kinds, nreal)
    with S.Writer(None, ref, *GRID, nreal, prec, ORG, SPC) as w:
        w.put(field)
    got = paths(str(tmp_path), "dev", kinds, nreal)
    m, v, x = [], [], []
    with S.Writer(eng, got, *GRID, nreal, prec, ORG, SPC) as w:
        assert w.slot >= max(calls)
        for c in calls:
            a, b, e, dev = w.realize(spec, iseed, igauss, c)
            assert dev
            m += list(a)
            v += list(b)
            x += list(e)
    assert np.array_equal(np.array(m), mean) and np.array_equal(np.array(v), var)
    assert np.array_equal(np.array(x), mi)
    for fr, fg in zip(files_of(ref, nreal), files_of(got, nreal)):
        assert open(fr, "rb").read() == open(fg, "rb").read(), os.path.basename(fg)
    if "bin" in kinds:
        raw = np.fromfile(got["bin"], dtype="<f4")
        assert np.array_equal(raw, field.astype(np.float32).ravel())


@pytest.mark.parametrize("prec", [S.FP64, S.FP32])
def test_variogram_of_device_chunks(tmp_path, eng, spec, prec):
    nreal, calls = 5, (2, 2, 1)
    dirs = [(1, 0, 0, 6), (0, 1, 1, 4), (1, -1, 0, 3)]
    field = eng.realize(spec, 31, 1, 0, nreal, *GRID, precision=prec)[0]
    want = eng.variogram(field, dirs)
    with S.Writer(eng, {"bin": str(tmp_path / "v.bin")}, *GRID, nreal, prec, ORG, SPC) as w:
        parts = []
        for c in calls:
            dev = w.realize(spec, 31, 1, c)[3]
            parts.append(eng.variogram(dev, dirs, device_shape=(c, GRID[2], GRID[1], GRID[0]), precision=prec))
    got = np.concatenate(parts)
    assert np.array_equal(got, want)


def test_writer_argument_checks(tmp_path, eng, spec):
    w = S.Writer(eng, {"bin": str(tmp_path / "c.bin")}, *GRID, 3, S.FP32)
    with pytest.raises(S.SS3DError):
        w.realize(spec, 1, 0, w.slot + 1)  # more than a slot
    w.realize(spec, 1, 0, 2)
    with pytest.raises(S.SS3DError):
        w.realize(spec, 1, 0, 2)  # past nreal
    w.realize(spec, 1, 0, 1)
    w.close()
    f = eng.realize(spec, 1, 0, 0, 3, *GRID, precision=S.FP32)[0]
    assert np.array_equal(np.fromfile(str(tmp_path / "c.bin"), dtype="<f4"), f.ravel())


def test_spectrum_destroyed_after_its_context():
    """A spectrum destroyed after its context (garbage-collection order) touches no freed context:
    the context detaches its spectra when it is destroyed; a detached spectrum is refused by
    other calls.  Engine.close() closes its writers, servers, batches and spectra first."""
    npg = S.periodic_grid(STS, *SPC, *GRID, 1)
    e = S.Engine(0)
    sp = e.spectrum(STS, NUGGET, *SPC, *npg)
    h = sp.handle
    assert S.lib().ss3d_context_destroy(e.handle) == 0  # C order: context first
    e.handle = None
    amp = np.empty(int(np.prod(npg)))
    assert S.lib().ss3d_spectrum_amplitude(h, amp.ctypes.data_as(S._dp)) == 1
    assert "destroyed" in S.last_error()
    assert S.lib().ss3d_spectrum_destroy(h) == 0
    sp.handle = None
    e2 = S.Engine(0)
    sp2 = e2.spectrum(STS, NUGGET, *SPC, *npg)
    w = S.Writer(e2, {}, *GRID, 2, S.FP32)
    w.realize(sp2, 5, 1, 1)
    e2.close()  # closes the writer (1 of 2 realizations given: its error is dropped) and the spectrum
    assert sp2.handle is None and w.handle is None
    with S.Engine(0) as e3:  # the device is still usable
        sp3 = e3.spectrum(STS, NUGGET, *SPC, *npg)
        f = e3.realize(sp3, 5, 1, 0, 1, *GRID, precision=S.FP32)[0]
        assert np.isfinite(f).all()
