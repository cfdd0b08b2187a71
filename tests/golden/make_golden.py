"""Write the golden fixtures of the reference program (run in the build container only).

The reference executables are compiled from /root/reference/src/specsim3d.f90 by oracle/Makefile:
  oracle/_ref/specsim3d      -- unmodified program (float32 .bin, .out, .vti, .vtk, stdout)
  oracle/_ref/specsim3d_f64  -- same program whose code-3 output holds the FP64 realizations

For every case below this script writes a parameter file into a scratch directory, runs the
reference there and stores:
  tests/golden/cases.json            parameters, periodic grid, stdout, SHA-256 of the source
  tests/golden/<case>.npy            FP64 realizations, shape (nreal, nz, ny, nx)
  tests/golden/formats/*             the reference's own output files for two tiny runs
                                     (new-format codes 1-4 and the legacy default outputs)
Fixtures are data (inputs and the reference's outputs); no reference source is stored.

Usage:  make -C oracle ref && python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SRC = "/root/reference/src/specsim3d.f90"
EXE = os.path.join(ROOT, "oracle", "_ref", "specsim3d")
EXE64 = os.path.join(ROOT, "oracle", "_ref", "specsim3d_f64")

# name: (origin, spacing, grid, nugget, structures, iseed, nreal, ipad, igauss)
CASES = {
    # 64x64x16 spherical 20/10/5 az 45, ipad=1, igauss=0, seed 69069, one realization
    "c1_sph_pad": ((0, 0, 0), (1, 1, 1), (64, 64, 16), 0.0, [(1, 1.0, 20, 10, 5, 45, 0, 0)], 69069, 1, 1, 0),
    "exp_nug_gauss": ((0, 0, 0), (1, 1, 1), (24, 20, 12), 0.3, [(2, 0.7, 8, 4, 2, 30, 15, 0)], 4711, 3, 1, 1),
    "gau_rot": ((0.5, 0.5, 0.5), (1, 1, 1), (16, 16, 8), 0.0, [(3, 1.0, 6, 3, 2, 120, -20, 10)], 12345, 2, 1, 0),
    "exp_2d": ((0, 0, 0), (1, 1, 1), (40, 30, 1), 0.3, [(2, 0.7, 10, 10, 10, 45, 0, 0)], 777, 3, 1, 0),
    "legacy_exp_gauss": ((0, 0, 0), (1, 1, 1), (16, 16, 16), 0.0, [(2, 1.0, 5, 5, 5, 0, 0, 0)], 4711, 3, 0, 1),
    "nested_sph_gau": ((0, 0, 0), (2, 1.5, 1), (20, 12, 6), 0.1,
                       [(1, 0.5, 12, 6, 3, 60, 10, 0), (3, 0.4, 5, 5, 2, 10, 0, 30)], 99991, 2, 1, 1),
    "odd_legacy_sph": ((0, 0, 0), (1, 1, 1), (13, 7, 5), 0.0, [(1, 1.0, 6, 4, 3, 15, 5, 5)], 31791191, 2, 0, 0),
    "line_pad_exp": ((0, 0, 0), (1, 1, 1), (50, 1, 1), 0.0, [(2, 1.0, 10, 10, 10, 0, 0, 0)], 2024, 2, 1, 1),
    "line_legacy": ((0, 0, 0), (1, 1, 1), (32, 1, 1), 0.0, [(1, 1.0, 8, 8, 8, 0, 0, 0)], 5, 2, 0, 0),
    "nx1_legacy": ((0, 0, 0), (1, 1, 1), (1, 16, 8), 0.2, [(2, 0.8, 4, 4, 2, 0, 0, 0)], 8, 2, 0, 1),
    "az300_negseed": ((0, 0, 0), (1, 1, 0.5), (12, 12, 10), 0.0, [(2, 1.0, 9, 3, 2, 300, 25, -40)], -271828, 2, 1, 0),
    "gau_2d_pad": ((0, 0, 0), (1, 1, 1), (32, 24, 1), 0.0, [(3, 1.0, 8, 4, 4, 75, 0, 0)], 1357, 2, 1, 1),
}


def par_text(case, base, codes=(3,), legacy=False):
    (o, d, g, nug, sts, iseed, nreal, ipad, igauss) = case
    lines = ["Parameters for SPECSIM3D", "************************", "", "START OF PARAMETERS:",
             "%r %r %r" % tuple(float(v) for v in o), "%r %r %r" % tuple(float(v) for v in d),
             "%d %d %d" % g, "%r %d" % (float(nug), len(sts))]
    for s in sts:
        lines.append("%d %r %r %r %r %r %r %r" % (s[0], *map(float, s[1:])))
    lines.append("%d %d %d %d" % (iseed, nreal, ipad, igauss))
    lines.append(base)
    if not legacy:
        lines.append(str(len(codes)))
        lines += [str(c) for c in codes]
        lines.append("0")
    return "\n".join(lines) + "\n"


def run(exe, wd, par, args=()):
    r = subprocess.run([exe, par, *args], cwd=wd, capture_output=True, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    if r.returncode != 0:
        sys.exit(f"reference failed on {par}:\n{r.stdout}\n{r.stderr}")
    return r.stdout


def main():
    for exe in (EXE, EXE64):
        if not os.path.exists(exe):
            sys.exit(f"missing {exe}: run `make -C oracle ref` first")
    sha = hashlib.sha256(open(REF_SRC, "rb").read()).hexdigest()
    manifest = {"reference_source_sha256": sha, "generator": "tests/golden/make_golden.py", "cases": {}}
    wd = tempfile.mkdtemp(prefix="ss3d_golden_")
    try:
        for name, case in CASES.items():
            (o, d, g, nug, sts, iseed, nreal, ipad, igauss) = case
            with open(os.path.join(wd, name + ".par"), "w") as fh:
                fh.write(par_text(case, name))
            out64 = run(EXE64, wd, name + ".par")
            f64 = np.fromfile(os.path.join(wd, name + ".bin"), dtype="<f8")
            assert f64.size == nreal * g[0] * g[1] * g[2], name
            out32 = run(EXE, wd, name + ".par")
            f32 = np.fromfile(os.path.join(wd, name + ".bin"), dtype="<f4")
            assert np.array_equal(f32, f64.astype(np.float32)), name
            np.save(os.path.join(HERE, name + ".npy"), f64.reshape(nreal, g[2], g[1], g[0]))
            grid_line = [ln for ln in out32.splitlines() if "periodic grid" in ln][0]
            manifest["cases"][name] = {
                "origin": list(o), "spacing": list(d), "grid": list(g), "nugget": nug,
                "structures": [list(s) for s in sts], "iseed": iseed, "nreal": nreal, "ipad": ipad,
                "igauss": igauss, "periodic_grid": [int(v) for v in grid_line.split()[-3:]],
                "stdout": [ln for ln in out64.splitlines() if "time" not in ln],
            }
            print(name, manifest["cases"][name]["periodic_grid"])
        # reference output files in every format for two tiny runs
        fdir = os.path.join(HERE, "formats")
        os.makedirs(fdir, exist_ok=True)
        tiny = ((0.5, 1.0, -2.0), (1.0, 2.0, 0.5), (6, 5, 4), 0.1, [(2, 0.9, 4, 3, 2, 30, 10, 0)], 424242, 2, 1, 0)
        with open(os.path.join(wd, "fmt_new.par"), "w") as fh:
            fh.write(par_text(tiny, "fmt_new", codes=(1, 2, 3, 4)))
        out_new = run(EXE, wd, "fmt_new.par")
        with open(os.path.join(wd, "fmt_legacy.par"), "w") as fh:
            fh.write(par_text(tiny, "fmt_legacy", legacy=True))
        out_leg = run(EXE, wd, "fmt_legacy.par")
        with open(os.path.join(wd, "fmt_legbin.par"), "w") as fh:
            fh.write(par_text(tiny, "fmt_legbin", legacy=True))
        out_lb = run(EXE, wd, "fmt_legbin.par", ("--bin",))
        files = ["fmt_new.par", "fmt_new.out", "fmt_new.bin", "fmt_new.vti", "fmt_new_1.vtk", "fmt_new_2.vtk",
                 "fmt_legacy.par", "fmt_legacy.out", "fmt_legacy_1.vtk",
                 "fmt_legbin.par", "fmt_legbin.bin", "fmt_legbin_1.vtk"]
        for f in files:
            shutil.copy(os.path.join(wd, f), os.path.join(fdir, f))
        strip = lambda s: [ln for ln in s.splitlines() if "time" not in ln]
        manifest["formats"] = {"case": {"origin": list(tiny[0]), "spacing": list(tiny[1]), "grid": list(tiny[2]),
                                        "nugget": tiny[3], "structures": [list(s) for s in tiny[4]],
                                        "iseed": tiny[5], "nreal": tiny[6], "ipad": tiny[7], "igauss": tiny[8]},
                               "files": files, "stdout_new": strip(out_new), "stdout_legacy": strip(out_leg),
                               "stdout_legbin": strip(out_lb)}
    finally:
        shutil.rmtree(wd)
    with open(os.path.join(HERE, "cases.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
