"""Write sampled fixtures of the reference program on large grids (build container only).

Workloads:
  c2       the 256^3 benchmark grid (512^3 periodic, ipad=1), exponential model sill 0.7 plus
           nugget 0.3, ranges 80/40/10, azimuth 30, dip 15, igauss=1, seed 69069, 2 realizations;
  gau_mid  128x128x64 (periodic 256x256x128), gaussian model sill 1, ranges 50/20/8, azimuth 120,
           dip -20, igauss=1, seed 4711, 2 realizations: the Gaussian spectrum's FP64 noise floor
           on a grid of 2^23 nodes.

The FP64 fields are large, so only a sample is stored:
  tests/golden/c2/<name>_sample.npz
    idx      int32 (NS,)            linear indices (x fastest) into one realization
    val      float64 (nreal, NS)    the reference's FP64 values at idx
    sums     float64 (nreal, 2)     sum and sum of squares of the whole realization (numpy pairwise)
    rowx     float64 (nreal, nx)    the whole x row at (y, z) = (0, 0)
    colz     float64 (nreal, nz)    the whole z column at (x, y) = (17, ny - 56)
  tests/golden/c2/<name>.json       parameters, periodic grid and the reference's stdout lines

The reference is oracle/_ref/specsim3d_f64 (the reference source compiled by oracle/Makefile with
its code-3 output holding FP64 values).  Fixtures are data; no reference source is stored.

Usage:  make -C oracle ref && python tests/golden/make_large_golden.py
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
from make_golden import EXE64, par_text  # noqa: E402

OUT = os.path.join(HERE, "c2")
# name: (origin, spacing, grid, nugget, structures, iseed, nreal, ipad, igauss)
CASES = {
    "c2": ((0, 0, 0), (1, 1, 1), (256, 256, 256), 0.3, [(2, 0.7, 80, 40, 10, 30, 15, 0)], 69069, 2, 1, 1),
    "gau_mid": ((0, 0, 0), (1, 1, 1), (128, 128, 64), 0.0, [(3, 1.0, 50, 20, 8, 120, -20, 0)], 4711, 2, 1, 1),
}
NS = 1 << 15


def write(name, case):
    (o, d, g, nug, sts, iseed, nreal, ipad, igauss) = case
    n = g[0] * g[1] * g[2]
    wd = tempfile.mkdtemp(prefix="ss3d_c2_")
    try:
        with open(os.path.join(wd, name + ".par"), "w") as fh:
            fh.write(par_text(case, name))
        r = subprocess.run([EXE64, name + ".par"], cwd=wd, capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 1)))
        if r.returncode != 0:
            sys.exit(f"reference failed:\n{r.stdout}\n{r.stderr}")
        f64 = np.fromfile(os.path.join(wd, name + ".bin"), dtype="<f8")
        assert f64.size == nreal * n
        f = f64.reshape(nreal, g[2], g[1], g[0])
        idx = np.sort(np.random.default_rng(20241204).choice(n, NS, replace=False)).astype(np.int32)
        flat = f.reshape(nreal, n)
        os.makedirs(OUT, exist_ok=True)
        np.savez(os.path.join(OUT, name + "_sample.npz"), idx=idx, val=flat[:, idx],
                 sums=np.stack([flat.sum(axis=1), (flat * flat).sum(axis=1)], axis=1),
                 rowx=f[:, 0, 0, :].copy(), colz=f[:, :, g[1] - 56, 17].copy())
        grid_line = [ln for ln in r.stdout.splitlines() if "periodic grid" in ln][0]
        meta = {"generator": "tests/golden/make_large_golden.py", "origin": list(o), "spacing": list(d), "grid": list(g),
                "nugget": nug, "structures": [list(s) for s in sts], "iseed": iseed, "nreal": nreal, "ipad": ipad,
                "igauss": igauss, "periodic_grid": [int(v) for v in grid_line.split()[-3:]],
                "stdout": [ln for ln in r.stdout.splitlines() if "time" not in ln]}
        with open(os.path.join(OUT, name + ".json"), "w") as fh:
            json.dump(meta, fh, indent=1)
        print("\n".join(meta["stdout"]))
    finally:
        shutil.rmtree(wd)


def main():
    if not os.path.exists(EXE64):
        sys.exit(f"missing {EXE64}: run `make -C oracle ref` first")
    for name, case in CASES.items():
        write(name, case)


if __name__ == "__main__":
    main()
