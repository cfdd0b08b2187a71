"""Writes tests/golden/cond/: the reference program's own conditional simulations (FP64).

For each case: the parameter file, the conditioning block (a separate file for the --cond
command line, or inside the parameter file), the GSLIB data file, and what the reference
executable built from its source (oracle/_ref/specsim3d_f64: the reference with its .bin holding
the FP64 realizations) writes: the conditioned FP64 realizations (<case>.npy, shape
(nreal, nz, ny, nx)), <base>_conddata.dat, <base>_nscore.trn (normal-score cases) and its
stdout without the timer lines (cond.json).

Cases cover: 3D simple kriging with every datum in the neighbourhood; a 2D lognormal field with
the normal-score transform, weights, duplicates merged in one node, data outside the grid and
outside the trimming limits, power lower tail and hyperbolic upper tail; a rotated anisotropic
search with few neighbours (ndmax < data in range), data off the nodes, a non-zero simple-kriging
mean and Gaussian coefficients; and the conditioning block inside the parameter file.

Usage: python tests/golden/make_cond_golden.py   (needs /root/reference, builds oracle/_ref)
"""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "cond")
EXE = os.path.join(ROOT, "oracle", "_ref", "specsim3d_f64")


def data_text(names, rows):
    s = "cond data\n%d\n" % len(names) + "\n".join(names) + "\n"
    return s + "".join(" ".join("%.8g" % v for v in r) + "\n" for r in rows)


def cases():
    rng = np.random.default_rng(20261204)
    out = {}
    # 1) 3D simple kriging, all data in the neighbourhood (the reference's test_gaussian, smaller)
    nx, ny, nz = 24, 20, 8
    idx = rng.choice(nx * ny * nz, 40, replace=False)
    i, j, k = np.unravel_index(idx, (nx, ny, nz))
    rows = np.c_[i, j, k, rng.normal(size=40)]
    out["sk3d"] = dict(
        par=f"0 0 0\n1 1 1\n{nx} {ny} {nz}\n0.0 1\n1 1.0 12 8 6 30 0 0\n12345 3 1\nsk3d\n",
        cpar="sk3d.dat\n1 2 3 4 0\n-1e21 1e21\n0 0.0\n-5 5\n1 0\n1 0\n100 100 100\n0 0 0\n100\n",
        data=data_text(["x", "y", "z", "v"], rows), grid=(nx, ny, nz), nreal=3, flags=["--bin"])
    # 2) 2D lognormal with normal scores, weights, duplicates, outside data, trimming, tails 2 and 4
    nx, ny = 32, 32
    idx = rng.choice(nx * ny, 70, replace=False)
    x, y = (idx % nx).astype(float), (idx // nx).astype(float)
    v = np.exp(rng.normal(1.0, 0.8, 70))
    w = rng.uniform(0.5, 2.0, 70)
    dup = np.c_[x[:6] + 0.2, y[:6] - 0.3, np.exp(rng.normal(1.0, 0.8, 6)), rng.uniform(0.5, 2.0, 6)]
    outside = np.c_[[-3.0, 40.0, 5.0], [4.0, 2.0, 33.7], [2.0, 3.0, 4.0], [1.0, 1.0, 1.0]]
    trimmed = np.c_[[3.0, 7.0], [9.0, 11.0], [-1.0, 500.0], [1.0, 1.0]]
    rows = np.r_[np.c_[x, y, v, w], dup, outside, trimmed]
    out["ns2d"] = dict(
        par=f"0 0 0\n1 1 1\n{nx} {ny} 1\n0.0 1\n2 1.0 15 10 15 20 0 0\n777 3 1\nns2d\n",
        cpar="ns2d.dat\n1 2 0 3 4\n0 100\n1 0.0\n0.0 60.0\n2 1.5\n4 1.5\n20 15 10\n30 0 0\n12\n",
        data=data_text(["x", "y", "v", "w"], rows), grid=(nx, ny, 1), nreal=3, flags=["--bin"])
    # 3) rotated anisotropic search, few neighbours, off-node data, mean 0.7, Gaussian coefficients
    nx, ny, nz = 20, 18, 10
    pts = np.c_[rng.uniform(0.5, 2 * nx - 1, 90), rng.uniform(0.5, 2 * ny - 1, 90), rng.uniform(0, nz - 1, 90)]
    rows = np.c_[pts, 0.7 + rng.normal(size=90)]
    out["rot3d"] = dict(
        par=f"0.5 0.5 0\n2 2 1\n{nx} {ny} {nz}\n0.1 1\n3 0.9 14 9 5 45 20 10\n4242 2 1 1\nrot3d\n",
        cpar="rot3d.dat\n1 2 3 4 0\n-1e21 1e21\n0 0.7\n-5 5\n1 0\n1 0\n16 10 6\n45 20 10\n8\n",
        data=data_text(["x", "y", "z", "v"], rows), grid=(nx, ny, nz), nreal=2, flags=["--bin"])
    # 4) the conditioning block inside the parameter file (output code 3 = .bin)
    c = out["sk3d"]
    block = c["cpar"].replace("sk3d.dat", "inf.dat")
    out["inf"] = dict(par=c["par"].replace("sk3d\n", "inf\n1\n3\n1\n") + block, cpar=None,
                      data=c["data"], grid=c["grid"], nreal=3, flags=[])
    return out


def main():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)
    os.makedirs(OUT, exist_ok=True)
    manifest = {}
    for name, c in cases().items():
        wd = tempfile.mkdtemp(prefix="ss3d_cond_")
        try:
            files = {f"{name}.par": c["par"], f"{name}.dat": c["data"]}
            if c["cpar"]:
                files[f"{name}.cpar"] = c["cpar"]
            for f, t in files.items():
                with open(os.path.join(wd, f), "w") as fh:
                    fh.write(t)
            args = [EXE, f"{name}.par", *c["flags"]] + (["--cond", f"{name}.cpar"] if c["cpar"] else [])
            r = subprocess.run(args, cwd=wd, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="4"))
            assert r.returncode == 0, r.stdout + r.stderr
            nx, ny, nz = c["grid"]
            f64 = np.fromfile(os.path.join(wd, f"{name}.bin"), dtype="<f8")
            assert f64.size == c["nreal"] * nx * ny * nz, name
            np.save(os.path.join(OUT, f"{name}.npy"), f64.reshape(c["nreal"], nz, ny, nx))
            for f in files:
                shutil.copy(os.path.join(wd, f), os.path.join(OUT, f))
            outs = [f"{name}_conddata.dat"] + ([f"{name}_nscore.trn"] if os.path.exists(
                os.path.join(wd, f"{name}_nscore.trn")) else [])
            for f in outs:
                shutil.copy(os.path.join(wd, f), os.path.join(OUT, f))
            manifest[name] = {"grid": list(c["grid"]), "nreal": c["nreal"], "flags": c["flags"],
                              "inputs": sorted(files), "outputs": outs,
                              "stdout": [ln for ln in r.stdout.splitlines() if "time" not in ln]}
            print(name, "ok")
        finally:
            shutil.rmtree(wd)
    with open(os.path.join(OUT, "cond.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)


if __name__ == "__main__":
    main()
