"""The Fortran driver end to end on the GPU, checked the way the reference's own scripts check
the reference executable, plus byte-level comparisons with the reference's output files.

  - tests/golden/formats: the reference's .out/.bin/.vti/.vtk files for a tiny run; the
    driver's files must be byte-identical (GSLIB text included);
  - tests/test_gauss_amplitudes.py:55-67 (16^3 legacy grid, 300 realizations): fixed
    amplitudes give a deterministic sample variance, Gaussian coefficients do not, same mean
    variance and variogram, Gaussian kurtosis, --gauss == fourth seed-line value;
  - tests/test_conditional.py:96-124 (unconditional part): legacy command line and new
    in-file format give byte-identical .bin; the .vti arrays equal the .bin;
  - golden realizations: the driver's float32 .bin equals the reference's float32 cast.
"""
import json
import os
import shutil
import sys

import numpy as np
import pytest

import golden_cases as G
from driver_util import EXE, par_text, run, strip_times

pytestmark = pytest.mark.gpu
FMT = os.path.join(G.GOLDEN, "formats")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fmt_manifest():
    return G.manifest()["formats"]


@pytest.mark.parametrize("base,args,files,key", [
    ("fmt_new", [], ["fmt_new.out", "fmt_new.bin", "fmt_new.vti", "fmt_new_1.vtk", "fmt_new_2.vtk"], "stdout_new"),
    ("fmt_legacy", [], ["fmt_legacy.out", "fmt_legacy_1.vtk"], "stdout_legacy"),
    ("fmt_legbin", ["--bin"], ["fmt_legbin.bin", "fmt_legbin_1.vtk"], "stdout_legbin"),
])
def test_output_files_byte_identical(tmp_path, base, args, files, key):
    say(f"[driver] formats {base} {args}")
    shutil.copy(os.path.join(FMT, base + ".par"), tmp_path / (base + ".par"))
    r = run(EXE, str(tmp_path), base + ".par", *args)
    assert r.returncode == 0, r.stdout + r.stderr
    for f in files:
        got = (tmp_path / f).read_bytes()
        want = open(os.path.join(FMT, f), "rb").read()
        assert got == want, f
    # max|Im| of a realization is rounding noise of the inverse transform: compare up to it
    cut = lambda lines: [ln.split("max|Im|")[0] for ln in lines]
    assert cut(strip_times(r.stdout)) == cut([ln for ln in fmt_manifest()[key] if ln.strip()])


def run_gauss(tmp_path, base, fourth="", flag=None, n=16, nreal=300):
    text = ("Parameters for SPECSIM3D\n************************\n\nSTART OF PARAMETERS:\n"
            f"0 0 0\n1 1 1\n{n} {n} {n}\n0.0 1\n2 1.0 5 5 5 0 0 0\n4711 {nreal} 0 {fourth}\n{base}\n1\n3\n0\n")
    (tmp_path / (base + ".par")).write_text(text)
    r = run(EXE, str(tmp_path), base + ".par", *([flag] if flag else []))
    assert r.returncode == 0, r.stdout + r.stderr
    f = np.fromfile(tmp_path / (base + ".bin"), dtype=np.float32).reshape(nreal, n, n, n).astype(np.float64)
    return f, r.stdout


def vario_x(f, kmax=5):
    return np.array([0.5 * np.mean((f[:, :, :, k:] - f[:, :, :, :-k]) ** 2) for k in range(1, kmax + 1)])


def test_gauss_amplitudes_contract(tmp_path):
    """tests/test_gauss_amplitudes.py of the reference, as assertions."""
    say("[driver] gauss amplitudes 16^3 x 300")
    fa, _ = run_gauss(tmp_path, "t_gauss_a")
    fb, out = run_gauss(tmp_path, "t_gauss_b", fourth="1")
    fc, _ = run_gauss(tmp_path, "t_gauss_c", flag="--gauss")
    va = fa.reshape(len(fa), -1).var(axis=1)
    vb = fb.reshape(len(fb), -1).var(axis=1)
    assert va.std() < 1e-5
    assert vb.std() > 0.03
    se = vb.std() / np.sqrt(len(vb))
    assert abs(vb.mean() - va.mean()) < 4 * se
    assert np.abs(vario_x(fa) - vario_x(fb)).max() < 0.04
    z = fb.ravel() / fb.std()
    assert abs(np.mean(z ** 4) - 3.0) < 0.1
    assert np.array_equal(fb, fc)
    assert "Gaussian" in out


def test_legacy_and_new_format_identical(tmp_path):
    """Legacy file + --bin and the new in-file outputs give the same .bin; the .vti arrays are the
    .bin bytes with UInt64 sizes (tests/test_conditional.py:96-119)."""
    say("[driver] legacy vs new format")
    nx, ny, nz, nreal = 64, 64, 16, 4
    kw = dict(origin=(0, 0, 0), spacing=(1, 1, 1), grid=(nx, ny, nz), nugget=0.0,
              structures=[(1, 1.0, 20, 10, 5, 45, 0, 0)], iseed=31791191, nreal=nreal, ipad=1)
    (tmp_path / "leg.par").write_text(par_text(**kw, base="leg"))
    (tmp_path / "new.par").write_text(par_text(**kw, base="new", codes=[3, 2, 4]))
    r1 = run(EXE, str(tmp_path), "leg.par", "--bin")
    r2 = run(EXE, str(tmp_path), "new.par")
    assert r1.returncode == 0 and r2.returncode == 0, r1.stdout + r2.stdout
    a = (tmp_path / "leg.bin").read_bytes()
    assert a == (tmp_path / "new.bin").read_bytes()
    assert (tmp_path / f"new_{nreal}.vtk").exists()
    raw = (tmp_path / "new.vti").read_bytes()
    head = raw[:raw.index(b"<AppendedData")].decode()
    assert 'WholeExtent="0 %d 0 %d 0 %d"' % (nx - 1, ny - 1, nz - 1) in head and head.count("<DataArray") == nreal
    pos = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    nb = 4 * nx * ny * nz
    ref = np.frombuffer(a, dtype="<f4").reshape(nreal, -1)
    for r in range(nreal):
        off = pos + r * (nb + 8)
        assert int.from_bytes(raw[off:off + 8], "little") == nb
        assert np.array_equal(np.frombuffer(raw[off + 8:off + 8 + nb], dtype="<f4"), ref[r])


@pytest.mark.parametrize("name", sorted(G.cases()))
def test_driver_bin_matches_reference_cast(tmp_path, name):
    c = G.cases()[name]
    say(f"[driver] golden {name}")
    (tmp_path / "g.par").write_text(par_text(c["origin"], c["spacing"], c["grid"], c["nugget"], c["structures"],
                                             c["iseed"], c["nreal"], c["ipad"], c["igauss"], base="g", codes=[3]))
    r = run(EXE, str(tmp_path), "g.par")
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(tmp_path / "g.bin", dtype="<f4")
    want = G.field(name).astype(np.float32).ravel()
    # FP64 fields agree to ~1e-15; their float32 casts differ only where a value sits on a
    # float32 rounding tie (expected count well below one per case)
    mism = np.count_nonzero(got != want)
    assert mism <= max(1, got.size // 100000), mism
    assert np.abs(got - want).max() <= np.abs(want).max() * 2 ** -23
    # the stdout lines (grid, diagnostics, per-realization statistics) are the reference's
    want_lines = [ln for ln in c["stdout"] if "realization" in ln or "grid" in ln or "C(0)" in ln]
    got_lines = [ln for ln in strip_times(r.stdout) if "realization" in ln or "grid" in ln or "C(0)" in ln]
    assert [ln.split("max|Im|")[0] for ln in got_lines] == [ln.split("max|Im|")[0] for ln in want_lines]


def test_fp32_option(tmp_path):
    c = G.cases()["exp_nug_gauss"]
    (tmp_path / "g.par").write_text(par_text(c["origin"], c["spacing"], c["grid"], c["nugget"], c["structures"],
                                             c["iseed"], c["nreal"], c["ipad"], c["igauss"], base="g", codes=[3]))
    r = run(EXE, str(tmp_path), "g.par", "--fp32")
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(tmp_path / "g.bin", dtype="<f4").astype(np.float64)
    want = G.field("exp_nug_gauss").ravel()
    assert np.sqrt(np.mean((got - want) ** 2) / np.mean(want ** 2)) < 1e-5


@pytest.mark.parametrize("name", sorted(G.vario_cases()))
def test_vario_table_and_report(tmp_path, name):
    """--vario: the <base>_vario.csv table and the stdout report of vario_report (f90:434-529)
    against the reference's own, for FP64 realizations of the same run."""
    c = G.vario_cases()[name]
    say(f"[driver] vario {name}")
    (tmp_path / "v.par").write_text(par_text(c["origin"], c["spacing"], c["grid"], c["nugget"], c["structures"],
                                             c["iseed"], c["nreal"], c["ipad"], c["igauss"], base=name, codes=[3]))
    shutil.copy(os.path.join(G.VARIO, name + "_v.par"), tmp_path / "vv.par")
    r = run(EXE, str(tmp_path), "v.par", "--vario", "vv.par")
    assert r.returncode == 0, r.stdout + r.stderr
    assert any(ln.startswith(" time variograms           :") for ln in r.stdout.splitlines())
    head, rows = G.vario_table(name)
    lines = (tmp_path / (name + "_vario.csv")).read_text().splitlines()
    assert lines[0].split(",") == head
    got = [ln.split(",") for ln in lines[1:]]
    assert len(got) == len(rows)
    for g, w in zip(got, rows):
        assert g[:5] == w[:5] and g[6] == w[6]  # dir, step, lag, pairs
        assert g[5] == w[5] and g[7] == w[7]    # h and the model variogram: same host arithmetic
        for a, b in zip(g[8:], w[8:]):          # ensemble and per-realization values
            assert float(a) == pytest.approx(float(b), rel=2e-6, abs=1e-12), (g, w)
    want = [ln for ln in c["stdout"] if "vario" in ln or ln.startswith("   dir")]
    have = [ln for ln in strip_times(r.stdout) if "vario" in ln or ln.startswith("   dir")]
    assert len(have) == len(want)
    for h, w in zip(have, want):
        if h.startswith("   dir"):
            assert h.split()[:6] == w.split()[:6]
            assert [float(v) for v in h.split()[6:]] == pytest.approx([float(v) for v in w.split()[6:]], abs=2e-5)
        else:
            assert h == w
