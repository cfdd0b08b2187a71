"""Write the variogram fixtures of the reference program (run in the build container only).

For each case the reference executables built by oracle/Makefile run a parameter file with
`--vario <file>` in a scratch directory:
  oracle/_ref/specsim3d_f64  -- FP64 realizations (code-3 file) and the variogram table
  oracle/_ref/specsim3d      -- the unmodified program, whose stdout and table must be the same
and this script stores under tests/golden/vario/:
  <case>.npy          FP64 realizations, shape (nreal, nz, ny, nx)
  <case>_v.par        the variogram parameter file
  <case>_vario.csv    the reference's table (header, one row per lag with pairs)
  vario.json          parameters and the reference's stdout lines (timers removed)
Fixtures are data (inputs and the reference's outputs); no reference source is stored.

Usage:  make -C oracle ref && python tests/golden/make_vario_golden.py
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from make_golden import EXE, EXE64, par_text  # noqa: E402

OUT = os.path.join(HERE, "vario")

# name: (case as in make_golden, iperreal, [(sx, sy, sz, nlag), ...])
CASES = {
    # 64x64x16 spherical 20/10/5 N45E, five realizations, the directions of the reference's own check
    "yao_dirs": (((0, 0, 0), (1, 1, 1), (64, 64, 16), 0.0, [(1, 1.0, 20, 10, 5, 45, 0, 0)], 31791191, 5, 1, 0),
                 0, [(1, 1, 0, 40), (1, -1, 0, 40), (0, 0, 1, 40)]),
    # nested exponential + gaussian, per-realization columns, axis, diagonal and long steps; some
    # lags run past the grid (no pairs: no row)
    "nested_iper": (((0, 0, 0), (2, 1.5, 1), (20, 12, 6), 0.1,
                     [(2, 0.5, 12, 6, 3, 60, 10, 0), (3, 0.4, 5, 5, 2, 10, 0, 30)], 99991, 3, 1, 1),
                    1, [(1, 0, 0, 25), (0, 1, 0, 8), (0, 0, 1, 7), (2, -1, 1, 6), (-1, 0, 0, 4)]),
    # 2D grid (nz = 1)
    "flat_2d": (((0, 0, 0), (1, 1, 1), (40, 30, 1), 0.3, [(2, 0.7, 10, 10, 10, 45, 0, 0)], 777, 2, 1, 0),
                1, [(1, 0, 0, 30), (0, 1, 0, 20), (1, 1, 0, 15), (0, 0, 1, 3)]),
}


def run(exe, wd, args):
    r = subprocess.run([exe, *args], cwd=wd, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
    if r.returncode != 0:
        sys.exit(f"reference failed: {args}\n{r.stdout}\n{r.stderr}")
    return r.stdout


def main():
    for exe in (EXE, EXE64):
        if not os.path.exists(exe):
            sys.exit(f"missing {exe}: run `make -C oracle ref` first")
    os.makedirs(OUT, exist_ok=True)
    manifest = {"generator": "tests/golden/make_vario_golden.py", "cases": {}}
    wd = tempfile.mkdtemp(prefix="ss3d_vario_")
    try:
        for name, (case, iper, dirs) in CASES.items():
            (o, d, g, nug, sts, iseed, nreal, ipad, igauss) = case
            with open(os.path.join(wd, name + ".par"), "w") as fh:
                fh.write(par_text(case, name))
            vpar = f"{len(dirs)} {iper}\n" + "".join(f"{a} {b} {c} {n}\n" for a, b, c, n in dirs)
            with open(os.path.join(wd, name + "_v.par"), "w") as fh:
                fh.write(vpar)
            out64 = run(EXE64, wd, [name + ".par", "--vario", name + "_v.par"])
            f64 = np.fromfile(os.path.join(wd, name + ".bin"), dtype="<f8")
            csv64 = open(os.path.join(wd, name + "_vario.csv")).read()
            out32 = run(EXE, wd, [name + ".par", "--vario", name + "_v.par"])
            csv32 = open(os.path.join(wd, name + "_vario.csv")).read()
            assert csv32 == csv64, name  # the table is computed from the FP64 realizations in both
            strip = lambda s: [ln for ln in s.splitlines() if "time" not in ln]
            assert strip(out32) == strip(out64), name
            np.save(os.path.join(OUT, name + ".npy"), f64.reshape(nreal, g[2], g[1], g[0]))
            shutil.copy(os.path.join(wd, name + "_v.par"), os.path.join(OUT, name + "_v.par"))
            shutil.copy(os.path.join(wd, name + "_vario.csv"), os.path.join(OUT, name + "_vario.csv"))
            manifest["cases"][name] = {
                "origin": list(o), "spacing": list(d), "grid": list(g), "nugget": nug,
                "structures": [list(s) for s in sts], "iseed": iseed, "nreal": nreal, "ipad": ipad,
                "igauss": igauss, "iper": iper, "dirs": [list(x) for x in dirs], "stdout": strip(out32)}
            print(name, len(csv32.splitlines()) - 1, "rows")
    finally:
        shutil.rmtree(wd)
    with open(os.path.join(OUT, "vario.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
