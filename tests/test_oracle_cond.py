"""The oracle's conditioning (oracle/oracle_cond.c, a restatement of cond_mod,
src/specsim3d.f90:548-1083) pinned on the CPU:

  - against the reference executable's own conditional simulations (tests/golden/cond, FP64
    realizations written by the reference build): conditioned fields, the counts it prints, the
    used data (<base>_conddata.dat) and the normal-score table (<base>_nscore.trn);
  - against mathematics that fixes the result: the data are reproduced at their nodes; with every
    datum in the neighbourhood the kriged residual equals exact simple kriging solved by numpy;
    the normal scores of equal-weight data are the standard normal quantiles (scipy); gauinv and
    gcum against scipy's normal ppf/cdf; back transform of a table value is that value; the
    heapsort permutation sorts its keys.
"""
import os
import re

import numpy as np
import pytest
from scipy.stats import norm

import golden_cases as G
import oracle_ffi as O

CASES = ["sk3d", "ns2d", "rot3d", "inf"]


def run_case(name):
    c = G.cond_case(name)
    res = O.simulate(c["sts"], c["nugget"], *c["spacing"], *c["grid"], c["ipad"], c["iseed"], c["nreal"],
                     c["igauss"])["res"].copy()
    unc = res.copy()
    out = O.condition(res, c["origin"], c["spacing"], c["sts"], c["nugget"], O.cond_params(**c["params"]), c["rows"])
    return c, unc, res, out


@pytest.mark.parametrize("name", CASES)
def test_conditioned_fields_equal_reference(name):
    c, unc, res, out = run_case(name)
    ref = G.cond_field(name)
    assert res.shape == ref.shape
    assert np.array_equal(res, ref), np.abs(res - ref).max()


@pytest.mark.parametrize("name", CASES)
def test_counts_and_files_equal_reference(name):
    c, unc, res, out = run_case(name)
    info = out["info"]
    lines = G.cond_manifest()[name]["stdout"]
    cd = [ln for ln in lines if "conditioning data:" in ln][0].replace("(", " ").replace(",", " ").split()
    assert (info["nd"], info["nmerged"], info["nout"]) == (int(cd[2]), int(cd[5]), int(cd[7]))
    sb = [ln for ln in lines if "super blocks:" in ln][0].split()
    assert (info["nsbx"], info["nsby"], info["nsbz"]) == (int(sb[2]), int(sb[3]), int(sb[4]))
    mn = re.search(r"mean neighbours per node:\s+([\d.]+)\s+\(nodes without data: (\d+), data nodes: (\d+)\)",
                   [ln for ln in lines if "mean neighbours" in ln][0])
    nx, ny, nz = c["grid"]
    assert "%6.1f" % (info["nsum"] / (nx * ny * nz)) == "%6.1f" % float(mn.group(1))
    assert info["nnone"] == int(mn.group(2)) and info["nexact"] == int(mn.group(3))
    assert not any("WARNING" in ln for ln in lines) and info["nfail"] == 0
    # <base>_conddata.dat: x y z value gauss of every used node (es17.8)
    with open(os.path.join(G.COND, f"{c['base']}_conddata.dat")) as fh:
        body = np.array([list(map(float, ln.split())) for ln in fh.read().splitlines()[7:]])
    ijk = out["ijk"]
    xyz = np.array(c["origin"]) + (ijk - 1) * np.array(c["spacing"])
    assert np.allclose(body[:, :3], xyz, rtol=1e-8, atol=1e-12)
    assert np.allclose(body[:, 3], out["dval"], rtol=1e-8, atol=0) and np.allclose(body[:, 4], out["dg"], rtol=1e-8,
                                                                                    atol=1e-12)
    if c["params"]["itrans"] == 1:
        with open(os.path.join(G.COND, f"{c['base']}_nscore.trn")) as fh:
            tab = np.array([list(map(float, ln.split())) for ln in fh.read().splitlines()[4:]])
        assert np.allclose(tab[:, 0], out["vrt"], rtol=1e-8) and np.allclose(tab[:, 1], out["gtt"], rtol=1e-8,
                                                                              atol=1e-12)


@pytest.mark.parametrize("name", ["sk3d", "rot3d"])
def test_data_reproduced(name):
    c, unc, res, out = run_case(name)
    i, j, k = (out["ijk"] - 1).T
    want = out["dg"] + c["params"]["skmean"]
    assert np.abs(res[:, k, j, i] - want[None]).max() < 1e-12


def test_simple_kriging_exact():
    """Every datum in the neighbourhood: conditioned = unconditional + SK(data - unconditional at data)."""
    c, unc, res, out = run_case("sk3d")
    st = c["sts"][0]
    rot, _ = O.setrot(st)
    ijk = out["ijk"] - 1
    xyz = np.array(c["origin"]) + ijk * np.array(c["spacing"])

    def cov(d):
        r = d @ rot.T
        h = np.sqrt((r ** 2).sum(-1)) / st[2]
        return np.where(h < 1, st[1] * (1 - 1.5 * h + 0.5 * h ** 3), 0.0)

    K = cov(xyz[:, None] - xyz[None]) + 1e-10 * np.eye(len(xyz))
    nx, ny, nz = c["grid"]
    g = np.meshgrid(*(np.arange(n, dtype=float) for n in (nx, ny, nz)), indexing="ij")
    P = np.c_[g[0].ravel(order="F"), g[1].ravel(order="F"), g[2].ravel(order="F")] * np.array(c["spacing"]) + \
        np.array(c["origin"])
    W = np.linalg.solve(K, cov(xyz[None] - P[:, None]).T).T
    for r in range(c["nreal"]):
        u = unc[r].ravel()
        resid = out["dg"] - u[ijk[:, 0] + nx * (ijk[:, 1] + ny * ijk[:, 2])]
        want = (u + W @ resid)
        at = ijk[:, 0] + nx * (ijk[:, 1] + ny * ijk[:, 2])
        want[at] = out["dg"]
        assert np.abs(res[r].ravel() - want).max() < 1e-8


def test_normal_scores_and_back_transform():
    rng = np.random.default_rng(1)
    v = np.exp(rng.normal(size=300))
    g, vr, gt = O.nscore(v, np.ones(300))
    assert np.array_equal(vr, np.sort(v))
    assert np.abs(gt - norm.ppf((np.arange(300) + 0.5) / 300)).max() < 1e-12
    assert np.array_equal(g[np.argsort(v)], gt)
    for p in [1e-20, 1e-9, 0.01, 0.02425, 0.3, 0.5, 0.9, 0.97575, 0.999]:
        pc = max(p, 1e-15)
        assert abs(O.gauinv(p) - norm.ppf(pc)) < 1e-12 * max(1, abs(norm.ppf(pc)))
    # near 1 the Halley step evaluates gcum(x) - p with cancellation (the reference's algorithm):
    # about 1e-9 relative at 1 - 1e-12
    assert abs(O.gauinv(1 - 1e-12) - norm.ppf(1 - 1e-12)) < 1e-7 * norm.ppf(1 - 1e-12)
    for x in [-8, -2, -0.3, 0, 0.7, 3, 9]:
        assert abs(O.gcum(x) - norm.cdf(x)) < 1e-15
    L = O.lib()
    vr_, gt_ = np.ascontiguousarray(vr), np.ascontiguousarray(gt)
    for q in (0, 17, 299):
        z = L.or_backtr(gt[q], 300, O._p(vr_), O._p(gt_), 0.0, 100.0, 1, 1.0, 1, 1.0)
        assert z == pytest.approx(vr[q], rel=1e-15)


def test_heapsort_permutation():
    rng = np.random.default_rng(2)
    key = rng.integers(0, 20, 101).astype(float)
    idx = O.heapsort_idx(key)
    assert sorted(idx.tolist()) == list(range(1, 102))
    assert np.all(np.diff(key[idx - 1]) >= 0)


@pytest.mark.parametrize("case", ["escalated", "unsolved"])
def test_ridge_rule(case):
    """Negative nuggets make the kriging matrix indefinite: just below zero the reference's rule
    (ridge 1e-10 C(0), then 1e3 and 1e6 times larger, f90:989-1000) solves with the second ridge,
    which numpy reproduces; far below, no ridge solves and the nodes keep z_s."""
    rng = np.random.default_rng(9)
    nx, ny, nz = 12, 10, 4
    idx = rng.choice(nx * ny * nz, 14, replace=False)
    i, j, k = np.unravel_index(idx, (nx, ny, nz))
    v = rng.normal(size=14)
    rows = np.c_[i, j, k, v].astype(float)
    sts = [(1, 1.0, 30.0, 30.0, 30.0, 0.0, 0.0, 0.0)]
    rot, _ = O.setrot(sts[0])
    xyz = np.c_[i, j, k].astype(float)

    def cov(d):
        h = np.sqrt(((d @ rot.T) ** 2).sum(-1)) / 30.0
        return np.where(h < 1, 1 - 1.5 * h + 0.5 * h ** 3, 0.0)

    K = cov(xyz[:, None] - xyz[None])
    nug = -np.linalg.eigvalsh(K).min() - 5e-9 if case == "escalated" else -0.5
    params = dict(ixc=1, iyc=2, izc=3, ivc=4, iwc=0, tmin=-1e21, tmax=1e21, itrans=0, skmean=0.0, zlo=-5, zhi=5,
                  ltail=1, ltpar=0.0, utail=1, utpar=0.0, rmax=100.0, rmed=100.0, rvert=100.0, sa1=0.0, sa2=0.0,
                  sa3=0.0, ndmax=100)
    f = np.random.default_rng(1).standard_normal((2, nz, ny, nx))
    res = f.copy()
    out = O.condition(res, (0, 0, 0), (1, 1, 1), sts, nug, O.cond_params(**params), rows)
    assert out["info"]["nfail"] == nx * ny * nz - 14 and out["info"]["nexact"] == 14
    at = i + nx * (j + ny * k)
    mask = np.ones(nx * ny * nz, bool)
    mask[at] = False
    if case == "unsolved":
        assert np.array_equal(res.reshape(2, -1)[:, mask], f.reshape(2, -1)[:, mask])
    else:
        c0 = 1.0 + nug
        A = K + (nug + 1e-7 * c0) * np.eye(14)
        g = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        P = np.c_[g[0].ravel(order="F"), g[1].ravel(order="F"), g[2].ravel(order="F")].astype(float)
        W = np.linalg.solve(A, cov(xyz[None] - P[:, None]).T).T
        for r in range(2):
            u = f[r].ravel()
            want = u + W @ (v - u[at])
            # the 1e3-fold ridge leaves the system nearly singular (weights ~1e5): relative check
            assert np.abs(res[r].ravel()[mask] - want[mask]).max() < 1e-7 * np.abs(want[mask]).max()
    assert np.abs(res.reshape(2, -1)[:, at] - v[None]).max() < 1e-15


@pytest.mark.parametrize("name", CASES)
def test_library_data_preparation_equals_oracle(name):
    """ss3d_cond_prepare (the product's host code) gives the oracle's used data bit for bit."""
    import specsim3d_amd as S

    c, unc, res, out = run_case(name)
    p = c["params"]
    got = S.cond_prepare(c["rows"], c["grid"], c["origin"], c["spacing"], p["ixc"], p["iyc"], p["izc"], p["ivc"],
                         p["iwc"], p["tmin"], p["tmax"], p["itrans"], p["skmean"])
    info = out["info"]
    assert (got["nd"], got["nmerged"], got["nout"]) == (info["nd"], info["nmerged"], info["nout"])
    assert np.array_equal(got["ijk"], out["ijk"])
    assert np.array_equal(got["dval"], out["dval"]) and np.array_equal(got["dg"], out["dg"])
    if p["itrans"] == 1:
        assert np.array_equal(got["vr"], out["vrt"]) and np.array_equal(got["gt"], out["gtt"])
