"""The Fortran driver's command line and parameter-file checks (no GPU needed: every case stops
before the engine is opened).  Messages and exit codes are those of the reference program
(src/specsim3d.f90:1129-1220), compared directly with the reference build where it exists."""
import os

import pytest

from driver_util import EXE, REF_EXE, par_text, run

GOOD = dict(origin=(0, 0, 0), spacing=(1, 1, 1), grid=(8, 8, 8), nugget=0.0,
            structures=[(1, 1.0, 5, 5, 5, 0, 0, 0)], iseed=1, nreal=2, ipad=0)

CASES = {
    "nstruct": ("0 0 0\n1 1 1\n8 8 8\n0.0 11\n", [], "nstruct out of range", 0),
    "ranges": ("0 0 0\n1 1 1\n8 8 8\n0.0 1\n1 1 0 1 1 0 0 0\n", [], "ranges must be positive", 0),
    "igauss": ("0 0 0\n1 1 1\n8 8 8\n0.0 1\n1 1 5 5 5 0 0 0\n1 2 0 5\nx\n", [], "igauss must be 0 or 1", 1),
    "code": (par_text(**GOOD, base="x", codes=[9]), [], "output code", 1),
    "nout": (par_text(**GOOD, base="x") + "12\n", [], "nout must be in 1..9", 1),
    "icond": (par_text(**GOOD, base="x", codes=[3], icond=2), [], "icond must be 0 or 1", 1),
    "option": (par_text(**GOOD, base="x"), ["--foo"], "unknown option: --foo", 1),
    "missing": (None, [], "Parameter file not found", 1),
}


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.fail(f"driver not built: {EXE}")
    return EXE


@pytest.mark.parametrize("case", sorted(CASES))
def test_errors_match_reference(exe, case, tmp_path):
    text, args, msg, rc = CASES[case]
    if text is not None:
        (tmp_path / "p.par").write_text(text)
    r = run(exe, str(tmp_path), "p.par", *args)
    out = r.stdout + r.stderr
    assert msg in out, out
    assert r.returncode == rc, (r.returncode, out)
    if os.path.exists(REF_EXE):
        q = run(REF_EXE, str(tmp_path), "p.par", *args)
        assert q.returncode == r.returncode
        assert [ln for ln in (q.stdout + q.stderr).splitlines() if "IEEE" not in ln] == \
               [ln for ln in out.splitlines() if "IEEE" not in ln]


def test_grid_cap_kept(exe, tmp_path):
    """The reference stops above 2^27 periodic nodes (f90:1253); the driver does too by default."""
    big = dict(GOOD, grid=(1024, 1024, 256), nreal=1)
    (tmp_path / "p.par").write_text(par_text(**big, base="x", codes=[3]))
    r = run(exe, str(tmp_path), "p.par")
    assert "periodic grid too large (>2^27 nodes)" in r.stdout + r.stderr


COND_BLOCK = "d.dat\n1 2 3 4 0\n-1e21 1e21\n0 0.0\n-5 5\n1 0\n1 0\n{r}\n0 0 0\n{nd}\n"
COND_CASES = {  # conditioning block errors (f90:734-761): message, exit code, how the block is given
    "ndmax": (COND_BLOCK.format(r="10 10 10", nd=0), "cond: ndmax must be in 1..100", 0, "file"),
    "radii": (COND_BLOCK.format(r="10 0 10", nd=8), "cond: search radii must be positive", 0, "file"),
    "datafile": (COND_BLOCK.format(r="10 10 10", nd=8), "Data file not found: d.dat", 1, "inline"),
    "condfile": (None, "Conditioning parameter file not found: missing.cpar", 1, "cli"),
}


@pytest.mark.parametrize("case", sorted(COND_CASES))
def test_conditioning_block_errors(exe, case, tmp_path):
    """The driver reads the conditioning block before the simulation, the reference after it: the
    error line and the exit code are the same, the reference has printed its realizations first."""
    block, msg, rc, how = COND_CASES[case]
    if how == "inline":
        (tmp_path / "p.par").write_text(par_text(**GOOD, base="x", codes=[3], icond=1) + block)
        args = []
    else:
        (tmp_path / "p.par").write_text(par_text(**GOOD, base="x"))
        if block is not None:
            (tmp_path / "c.cpar").write_text(block)
        args = ["--bin", "--cond", "c.cpar" if block is not None else "missing.cpar"]
    r = run(exe, str(tmp_path), "p.par", *args)
    out = r.stdout + r.stderr
    assert r.returncode == rc and msg in out, (r.returncode, out)
    assert "realization" not in out  # stopped before the engine was opened
    if os.path.exists(REF_EXE):
        q = run(REF_EXE, str(tmp_path), "p.par", *args)
        assert q.returncode == r.returncode and msg in q.stdout + q.stderr
