"""Conditional simulation through the Fortran driver (the reference's process contract) on the GPU.

  - tests/golden/cond: the reference executable's own runs.  The driver writes byte-identical
    <base>_conddata.dat and <base>_nscore.trn, prints the reference's lines (realizations, data
    counts, normal-score statistics, super blocks, neighbours), and its float32 .bin is the float32
    cast of the reference's FP64 conditioned fields (to 1 ulp, every model);
  - the contracts of the reference's tests/test_conditional.py (restated here): data reproduced at
    their nodes; the mean and variance over 400 realizations equal exact simple kriging within the
    Monte Carlo error; the normal-score table equals the standard normal quantiles and the back
    transform reproduces lognormal data; the conditioning block in the parameter file gives the
    same bytes as the --cond command line, with the VTK outputs of the new format.
"""
import os
import shutil
import sys

import numpy as np
import pytest
from scipy.stats import norm

import golden_cases as G
from driver_util import EXE, run, strip_times

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cut(lines):
    return [ln.split("max|Im|")[0] for ln in lines if ln.strip()]


@pytest.mark.parametrize("name", ["sk3d", "ns2d", "rot3d", "inf"])
def test_reference_cases(tmp_path, name):
    say(f"[driver cond] {name}")
    m = G.cond_manifest()[name]
    for f in m["inputs"]:
        shutil.copy(os.path.join(G.COND, f), tmp_path / f)
    args = list(m["flags"]) + (["--cond", f"{name}.cpar"] if f"{name}.cpar" in m["inputs"] else [])
    r = run(EXE, str(tmp_path), f"{name}.par", *args)
    assert r.returncode == 0, r.stdout + r.stderr
    for f in m["outputs"]:
        assert (tmp_path / f).read_bytes() == open(os.path.join(G.COND, f), "rb").read(), f
    assert cut(strip_times(r.stdout)) == cut(m["stdout"])
    c = G.cond_case(name)
    nx, ny, nz = c["grid"]
    got = np.fromfile(tmp_path / f"{c['base']}.bin", dtype="<f4").reshape(c["nreal"], nz, ny, nx)
    g32 = G.cond_field(name).astype(np.float32)
    assert np.all(np.abs(got - g32) <= np.spacing(np.abs(g32)))


def write_data(path, names, rows):
    with open(path, "w") as fh:
        fh.write("test data\n%d\n" % len(names) + "\n".join(names) + "\n")
        for r in rows:
            fh.write(" ".join("%.8g" % v for v in r) + "\n")


NX, NY, NZ, NREAL = 30, 30, 10, 400


def gaussian_case(wd):
    rng = np.random.default_rng(3)
    nd = 40
    idx = rng.choice(NX * NY * NZ, nd, replace=False)
    i, j, k = np.unravel_index(idx, (NX, NY, NZ))
    v = rng.normal(size=nd)
    write_data(os.path.join(wd, "tc_data.dat"), ["x", "y", "z", "v"], np.c_[i, j, k, v])
    with open(os.path.join(wd, "tc.par"), "w") as fh:
        fh.write(f"0 0 0\n1 1 1\n{NX} {NY} {NZ}\n0.0 1\n1 1.0 12 8 6 30 0 0\n12345 {NREAL} 1\ntc\n")
    with open(os.path.join(wd, "tc.cpar"), "w") as fh:
        fh.write("tc_data.dat\n1 2 3 4 0\n-1e21 1e21\n0 0.0\n-5 5\n1 0\n1 0\n100 100 100\n0 0 0\n100\n")
    return i, j, k, v


def test_simple_kriging_statistics(tmp_path):
    """Data reproduced; realization mean = SK estimate and variance = SK variance within Monte Carlo."""
    say("[driver cond] SK statistics, 400 realizations")
    i, j, k, v = gaussian_case(str(tmp_path))
    r = run(EXE, str(tmp_path), "tc.par", "--bin", "--cond", "tc.cpar")
    assert r.returncode == 0, r.stdout + r.stderr
    sims = np.fromfile(tmp_path / "tc.bin", dtype="<f4").reshape(NREAL, NZ, NY, NX)
    sims = sims.transpose(3, 2, 1, 0)  # (x, y, z, realization)
    az = np.radians(30.0)
    u = np.array([np.sin(az), np.cos(az), 0.0])
    w = np.array([np.cos(az), -np.sin(az), 0.0])

    def cov(d):
        h = np.sqrt((d @ u / 12.0) ** 2 + (d @ w / 8.0) ** 2 + (d[..., 2] / 6.0) ** 2)
        return np.where(h < 1, 1 - 1.5 * h + 0.5 * h ** 3, 0.0)

    xyz = np.c_[i, j, k].astype(float)
    K = cov(xyz[:, None] - xyz[None]) + 1e-10 * np.eye(len(v))
    g = np.meshgrid(*(np.arange(n, dtype=float) for n in (NX, NY, NZ)), indexing="ij")
    P = np.c_[g[0].ravel(), g[1].ravel(), g[2].ravel()]
    kv = cov(xyz[None] - P[:, None])
    W = np.linalg.solve(K, kv.T).T
    est = (W @ v).reshape(NX, NY, NZ)
    var = (1 - (W * kv).sum(1)).reshape(NX, NY, NZ)
    err_data = np.abs(sims[i, j, k, :] - v[:, None]).max()
    rms_mean = np.sqrt(np.mean((sims.mean(-1) - est) ** 2))
    mc = np.sqrt(var.mean() / NREAL)
    ratio = sims.var(-1, ddof=1).mean() / var.mean()
    say(f"data error {err_data:.2e}, RMS(mean-SK) {rms_mean:.4f} (MC {mc:.4f}), variance ratio {ratio:.4f}")
    assert err_data < 1e-5
    assert rms_mean < 1.3 * mc
    assert abs(ratio - 1) < 0.02


def test_normal_score_and_back_transform(tmp_path):
    say("[driver cond] normal scores")
    rng = np.random.default_rng(5)
    nd = 200
    idx = rng.choice(64 * 64, nd, replace=False)
    x, y = (idx % 64).astype(float), (idx // 64).astype(float)
    v = np.exp(rng.normal(1.0, 0.8, nd))
    write_data(str(tmp_path / "tn_data.dat"), ["x", "y", "v"], np.c_[x, y, v])
    (tmp_path / "tn.par").write_text("0 0 0\n1 1 1\n64 64 1\n0.0 1\n2 1.0 15 15 15 0 0 0\n777 10 1\ntn\n")
    (tmp_path / "tn.cpar").write_text("tn_data.dat\n1 2 0 3 0\n0 1e21\n1 0.0\n0.0 200.0\n1 0\n1 0\n"
                                      "30 30 30\n0 0 0\n16\n")
    r = run(EXE, str(tmp_path), "tn.par", "--bin", "--cond", "tn.cpar")
    assert r.returncode == 0, r.stdout + r.stderr
    table = np.loadtxt(tmp_path / "tn_nscore.trn", skiprows=4)
    assert np.abs(table[:, 1] - norm.ppf((np.arange(nd) + 0.5) / nd)).max() < 1e-6
    sims = np.fromfile(tmp_path / "tn.bin", dtype="<f4").reshape(10, 64, 64)
    assert np.abs(sims[:, idx // 64, idx % 64] - v[None]).max() / v.max() < 1e-5


def test_block_in_parameter_file(tmp_path):
    """Codes 3, 2, 4 and the conditioning block in the file give the --cond run's bytes."""
    say("[driver cond] block in the parameter file")
    gaussian_case(str(tmp_path))
    r = run(EXE, str(tmp_path), "tc.par", "--bin", "--cond", "tc.cpar")
    assert r.returncode == 0, r.stdout + r.stderr
    base = (tmp_path / "tc.par").read_text().splitlines()
    new = base[:-1] + ["tc_new", "3", "3", "2", "4", "1"] + (tmp_path / "tc.cpar").read_text().splitlines()
    (tmp_path / "tc_new.par").write_text("\n".join(new) + "\n")
    r2 = run(EXE, str(tmp_path), "tc_new.par")
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert (tmp_path / "tc.bin").read_bytes() == (tmp_path / "tc_new.bin").read_bytes()
    assert (tmp_path / f"tc_new_{NREAL}.vtk").exists()
    raw = (tmp_path / "tc_new.vti").read_bytes()
    pos = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    nb = 4 * NX * NY * NZ
    ref = np.fromfile(tmp_path / "tc_new.bin", dtype="<f4").reshape(NREAL, -1)
    for rr in (0, NREAL // 2, NREAL - 1):
        off = pos + rr * (nb + 8)
        assert int.from_bytes(raw[off:off + 8], "little") == nb
        assert np.array_equal(np.frombuffer(raw[off + 8:off + 8 + nb], dtype="<f4"), ref[rr])


def test_no_data_inside_the_grid(tmp_path):
    say("[driver cond] no data inside")
    write_data(str(tmp_path / "o.dat"), ["x", "y", "z", "v"], [[50, 50, 1, 1.0], [-9, 2, 1, 0.5]])
    (tmp_path / "o.par").write_text("0 0 0\n1 1 1\n8 8 4\n0.0 1\n1 1.0 5 5 5 0 0 0\n1 2 1\no\n")
    (tmp_path / "o.cpar").write_text("o.dat\n1 2 3 4 0\n-1e21 1e21\n0 0.0\n-5 5\n1 0\n1 0\n10 10 10\n0 0 0\n8\n")
    r = run(EXE, str(tmp_path), "o.par", "--bin", "--cond", "o.cpar")
    assert "cond: no conditioning data inside the grid" in r.stdout + r.stderr
    assert r.returncode == 0  # a STOP with a message, as the reference's
