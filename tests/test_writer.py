"""Output writers on the CPU (no GPU needed: a writer opened without a context is fed host fields).

  - ss3d_format_es against the Fortran runtime's ES text of 9,565 values (tests/golden/fmt,
    written by es_golden.f90 with the reference build's compiler), and against Python's correctly
    rounded conversion on random float64 and float32-exact values;
  - the reference's own output files of the tiny format runs (tests/golden/formats, written by
    the reference executable): the writer fed the oracle's FP64 field of the same run must
    reproduce every file byte for byte (.out, .bin, .vti, binary and ASCII .vtk), in one call
    or realization by realization;
  - float32 input, error paths (unwritable path, early close, bad arguments).
"""
import os

import numpy as np
import pytest

import golden_cases as G
import oracle_ffi as O
import specsim3d_amd as S

FMTDIR = os.path.join(G.GOLDEN, "fmt")
FORMATS = os.path.join(G.GOLDEN, "formats")


def fortran_es(x, w, d):
    """ESw.d from Python's correctly rounded %E (ties to even) in the Fortran runtime's layout."""
    if np.isnan(x):
        t = "NaN"
    elif np.isinf(x):
        t = "-Inf" if x < 0 else "Inf"
    else:
        m, e = ("%.*E" % (d, abs(x))).split("E")
        e = int(e)
        sign = "-" if np.signbit(x) else ""
        t = sign + m + (("E%+03d" % e) if abs(e) <= 99 else ("%+04d" % e))
    return t.rjust(w) if len(t) <= w else "*" * w


def test_format_es_matches_fortran_fixture():
    v = np.load(os.path.join(FMTDIR, "values.npy"))
    lines = open(os.path.join(FMTDIR, "es_golden.txt")).read().splitlines()
    assert len(lines) == v.size
    cols = list(zip(*[ln.split("|") for ln in lines]))
    for (w, d), want in zip([(15, 7), (13, 5), (16, 8), (14, 6)], cols):
        got = S.format_es(v, w, d)
        bad = [(float(v[i]).hex(), got[i], want[i]) for i in range(v.size) if got[i] != want[i]]
        assert not bad, (w, d, bad[:5])


def test_format_es_ties_round_to_even():
    # 1 + 2^-8 = 1.00390625 and 1 + 3*2^-8 = 1.01171875 are exact ties at 8 digits
    assert S.format_es([1.00390625, 1.01171875, -1.00390625], 15, 7) == \
        ["  1.0039062E+00", "  1.0117188E+00", " -1.0039062E+00"]
    assert S.format_es([9.99999995, 1e100, 1e-310], 15, 7) == ["  9.9999999E+00", "  1.0000000+100", "  1.0000000-310"]


@pytest.mark.parametrize("kind", ["f64_wide", "f64_unit", "f32"])
def test_format_es_random_against_python(kind):
    rng = np.random.default_rng({"f64_wide": 1, "f64_unit": 2, "f32": 3}[kind])
    n = 200000
    if kind == "f64_wide":
        v = rng.standard_normal(n) * 10.0 ** rng.uniform(-40, 40, n)
    elif kind == "f64_unit":
        v = rng.standard_normal(n)
    else:  # float32 fields: many exact decimal ties at 8 digits
        v = (rng.standard_normal(n) * 3).astype(np.float32).astype(np.float64)
    for w, d in [(15, 7), (13, 5), (16, 8)]:
        got = S.format_es(v, w, d)
        idx = rng.choice(n, 20000, replace=False)
        bad = [(float(v[i]).hex(), got[i]) for i in idx if got[i] != fortran_es(v[i], w, d)]
        assert not bad, (w, d, bad[:5])


def fmt_case():
    return G.manifest()["formats"]["case"]


def oracle_field(case):
    o = O.simulate([tuple(s) for s in case["structures"]], case["nugget"], *case["spacing"], *case["grid"],
                   case["ipad"], case["iseed"], case["nreal"], case["igauss"])
    return o["res"]


def write_case(tmp_path, base, files, res, chunks, precision=S.FP64):
    case = fmt_case()
    nx, ny, nz = case["grid"]
    paths = {}
    if "bin" in files:
        paths["bin"] = str(tmp_path / f"{base}.bin")
    if "vti" in files:
        paths["vti"] = str(tmp_path / f"{base}.vti")
    if "vtk" in files:
        paths["vtk_base"] = str(tmp_path / base)
        paths["vtk_count"] = files["vtk"]
    if "out" in files:
        paths["gslib"] = str(tmp_path / f"{base}.out")
    if "vtk_ascii" in files:
        paths["vtk_ascii"] = str(tmp_path / f"{base}_1.vtk")
    with S.Writer(None, paths, nx, ny, nz, res.shape[0], precision, case["origin"], case["spacing"]) as w:
        r = 0
        for c in chunks:
            w.put(res[r:r + c])
            r += c


@pytest.mark.parametrize("chunks", [(2,), (1, 1)])
def test_writer_reproduces_reference_files(tmp_path, chunks):
    res = oracle_field(fmt_case())
    # new-format run (codes 1-4): .out, .bin, .vti and binary .vtk of every realization
    write_case(tmp_path, "fmt_new", {"out": 1, "bin": 1, "vti": 1, "vtk": 2}, res, chunks)
    for f in ["fmt_new.out", "fmt_new.bin", "fmt_new.vti", "fmt_new_1.vtk", "fmt_new_2.vtk"]:
        assert (tmp_path / f).read_bytes() == open(os.path.join(FORMATS, f), "rb").read(), f
    # legacy run: .out and the ASCII .vtk of realization 1
    write_case(tmp_path, "fmt_legacy", {"out": 1, "vtk_ascii": 1}, res, chunks)
    for f in ["fmt_legacy.out", "fmt_legacy_1.vtk"]:
        assert (tmp_path / f).read_bytes() == open(os.path.join(FORMATS, f), "rb").read(), f
    # legacy run with --bin: .bin and the binary .vtk of realization 1
    write_case(tmp_path, "fmt_legbin", {"bin": 1, "vtk": 1}, res, chunks)
    for f in ["fmt_legbin.bin", "fmt_legbin_1.vtk"]:
        assert (tmp_path / f).read_bytes() == open(os.path.join(FORMATS, f), "rb").read(), f
    assert not (tmp_path / "fmt_legbin_2.vtk").exists()


def test_writer_ascii_vtk_without_gslib(tmp_path):
    """The ASCII .vtk alone keeps only realization 1 (no all-realizations copy)."""
    res = oracle_field(fmt_case())
    write_case(tmp_path, "fmt_legacy", {"vtk_ascii": 1}, res, (1, 1))
    f = "fmt_legacy_1.vtk"
    assert (tmp_path / f).read_bytes() == open(os.path.join(FORMATS, f), "rb").read()


def test_writer_float32_fields(tmp_path):
    """FP32 fields: .bin holds them as they are; .out prints their exact value in es15.7."""
    rng = np.random.default_rng(7)
    nx, ny, nz, nreal = 7, 3, 5, 4
    f = rng.standard_normal((nreal, nz, ny, nx)).astype(np.float32)
    paths = {"bin": str(tmp_path / "a.bin"), "gslib": str(tmp_path / "a.out"), "vti": str(tmp_path / "a.vti")}
    with S.Writer(None, paths, nx, ny, nz, nreal, S.FP32) as w:
        w.put(f[:3])
        w.put(f[3:])
    assert np.array_equal(np.fromfile(paths["bin"], dtype="<f4"), f.ravel())
    lines = open(paths["gslib"]).read().splitlines()
    assert lines[:2 + nreal] == ["SPECSIM3D realizations", "   4", "sim1", "sim2", "sim3", "sim4"]
    body = lines[2 + nreal:]
    cells = f.reshape(nreal, -1).T
    assert len(body) == cells.shape[0]
    for p in (0, 17, cells.shape[0] - 1):
        assert body[p] == "".join(fortran_es(float(x), 15, 7) for x in cells[p])
    # .vti: header, then per realization a UInt64 byte count and the float32 block
    raw = open(paths["vti"], "rb").read()
    k = raw.index(b"   _") + 4
    nb = nx * ny * nz * 4
    for r in range(nreal):
        off = k + r * (nb + 8)
        assert int.from_bytes(raw[off:off + 8], "little") == nb
        assert np.array_equal(np.frombuffer(raw[off + 8:off + 8 + nb], dtype="<f4"), f[r].ravel())
    assert raw[k + nreal * (nb + 8):] == b"\n  </AppendedData>\n</VTKFile>\n"


def test_writer_errors(tmp_path):
    with pytest.raises(S.SS3DError) as e:
        S.Writer(None, {"bin": str(tmp_path / "no_such_dir" / "a.bin")}, 2, 2, 2, 1)
    assert e.value.code == 6 and "no_such_dir" in str(e.value)
    w = S.Writer(None, {"bin": str(tmp_path / "b.bin")}, 2, 2, 2, 2)
    w.put(np.zeros((1, 2, 2, 2)))
    with pytest.raises(S.SS3DError) as e:
        w.close()
    assert e.value.code == 1 and "1 of 2" in str(e.value)
    with pytest.raises(S.SS3DError):
        S.Writer(None, {"vtk_base": str(tmp_path / "c"), "vtk_count": 3}, 2, 2, 2, 2)
    w = S.Writer(None, {"bin": str(tmp_path / "d.bin")}, 2, 2, 2, 1)
    with pytest.raises(S.SS3DError):
        w.put(np.zeros((2, 2, 2, 2)))
    w.put(np.ones((1, 2, 2, 2)))
    w.close()
    assert np.array_equal(np.fromfile(str(tmp_path / "d.bin"), dtype="<f4"), np.ones(8, dtype=np.float32))
    w2 = S.Writer(None, {}, 2, 2, 2, 1)  # opened without a context: realize is refused
    assert S.lib().ss3d_writer_realize(w2.handle, None, 1, 0, 1, None, None, None, None) == 1
    assert "without a context" in S.last_error()
    w2.put(np.zeros((1, 2, 2, 2)))
    w2.close()
