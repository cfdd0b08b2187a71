"""Conditioning on the GPU (ss3d_cond_*) against the oracle and the reference's own conditional
simulations (tests/golden/cond):

  - the device conditions its FP64 realizations; the result equals the oracle conditioning the
    same unconditional field within 1e-10 (max-norm, relative), and the reference's FP64
    conditioned fields within 1e-10 for every model (the covariance uses the exp of the reference
    build's libm, so the spectra are the reference's bit for bit);
  - node counts (neighbours, nodes without data, data nodes, regularized systems) equal the
    oracle's and the reference's;
  - FP32 realizations: within 1e-5 relative RMS of the FP64 conditioned fields, data reproduced;
  - the ridge rule: systems that need a larger ridge, and systems that cannot be solved, are
    counted and treated as the oracle treats them (negative nugget makes them so);
  - the writer conditions each chunk before writing it: its files equal those written from the
    conditioned host fields.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import golden_cases as G
import oracle_ffi as O
import specsim3d_amd as S

pytestmark = pytest.mark.gpu
CASES = ["sk3d", "ns2d", "rot3d", "inf"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Dev:
    """A device copy of a host array (plumbing for the tests)."""

    def __init__(self, arr):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.arr = np.ascontiguousarray(arr)
        self.p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.p), self.arr.nbytes) == 0
        assert self.hip.hipMemcpy(self.p, self.arr.ctypes.data, self.arr.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.arr)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


@pytest.fixture(scope="module")
def eng():
    with S.Engine(0) as e:
        yield e


def prepared(c, field64):
    """Oracle conditioning of a copy of field64, and the prepared data it used."""
    ref = field64.copy()
    out = O.condition(ref, c["origin"], c["spacing"], c["sts"], c["nugget"], O.cond_params(**c["params"]), c["rows"])
    return ref, out


def device_cond(eng, c, out):
    p = c["params"]
    table = (out["vrt"], out["gtt"]) if p["itrans"] == 1 else None
    return eng.conditioning(c["sts"], c["nugget"], c["grid"], c["origin"], c["spacing"], out["ijk"], out["dg"],
                            p["rmax"], p["rmed"], p["rvert"], (p["sa1"], p["sa2"], p["sa3"]), p["ndmax"], p["itrans"],
                            p["skmean"], table, p["zlo"], p["zhi"], p["ltail"], p["ltpar"], p["utail"], p["utpar"])


def unconditional(eng, c, precision):
    npg = S.periodic_grid(c["sts"], *c["spacing"], *c["grid"], c["ipad"])
    sp = eng.spectrum(c["sts"], c["nugget"], *c["spacing"], *npg)
    f = eng.realize(sp, c["iseed"], c["igauss"], 0, c["nreal"], *c["grid"], precision=precision)[0]
    return sp, f


@pytest.mark.parametrize("name", CASES)
def test_fp64_conditioning_matches_oracle_and_reference(eng, name):
    say(f"[cond] fp64 {name}")
    c = G.cond_case(name)
    sp, f = unconditional(eng, c, S.FP64)
    want, out = prepared(c, f)
    k = device_cond(eng, c, out)
    d = Dev(f)
    try:
        k.apply(d.p.value, S.FP64, c["nreal"])
        eng.synchronize()  # apply is ordered on the context's stream; the copy back is not
        got = d.get()
    finally:
        d.free()
    info = k.info()
    k.close()
    sp.close()
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-10
    oi = out["info"]
    assert (info["nsbx"], info["nsby"], info["nsbz"], info["nsum"], info["nnone"], info["nexact"], info["nfail"]) == \
        (oi["nsbx"], oi["nsby"], oi["nsbz"], oi["nsum"], oi["nnone"], oi["nexact"], oi["nfail"])
    gold = G.cond_field(name)
    assert np.abs(got - gold).max() / np.abs(gold).max() < 1e-10
    # data nodes hold the data
    i, j, kk = (out["ijk"] - 1).T
    if c["params"]["itrans"] == 0:
        assert np.abs(got[:, kk, j, i] - (out["dg"] + c["params"]["skmean"])[None]).max() < 1e-12


@pytest.mark.parametrize("name", ["sk3d", "ns2d"])
def test_fp32_conditioning(eng, name):
    say(f"[cond] fp32 {name}")
    c = G.cond_case(name)
    sp, f32 = unconditional(eng, c, S.FP32)
    _, out = prepared(c, f32.astype(np.float64))
    k = device_cond(eng, c, out)
    d = Dev(f32)
    try:
        k.apply(d.p.value, S.FP32, c["nreal"])
        eng.synchronize()
        got = d.get()
    finally:
        d.free()
    k.close()
    sp.close()
    gold = G.cond_field(name)
    assert np.sqrt(np.mean((got - gold) ** 2) / np.mean(gold ** 2)) < 1e-5
    i, j, kk = (out["ijk"] - 1).T
    dat = G.cond_case(name)["rows"]
    if c["params"]["itrans"] == 0:
        assert np.abs(got[:, kk, j, i] - out["dg"].astype(np.float32)[None]).max() == 0
    else:  # back-transformed data values (merged nodes hold the weighted mean)
        assert np.abs(got[:, kk, j, i] - out["dval"][None]).max() / np.abs(dat[:, 2]).max() < 1e-5


@pytest.mark.parametrize("nugget,expect", [(None, "escalated"), (-0.5, "unsolved")])
def test_ridge_rule(eng, nugget, expect):
    """A negative nugget shifts the kriging matrix's spectrum: just below zero the reference's ridge
    rule solves with a 1e3-fold ridge; far below, no ridge solves and the node keeps z_s."""
    say(f"[cond] ridge {expect}")
    rng = np.random.default_rng(9)
    nx, ny, nz = 12, 10, 4
    idx = rng.choice(nx * ny * nz, 14, replace=False)
    i, j, k = np.unravel_index(idx, (nx, ny, nz))
    rows = np.c_[i, j, k, rng.normal(size=14)].astype(float)
    sts = [(1, 1.0, 30.0, 30.0, 30.0, 0.0, 0.0, 0.0)]
    if nugget is None:  # smallest eigenvalue of the data covariance matrix, less 5e-9
        rot, _ = O.setrot(sts[0])
        xyz = np.c_[i, j, k].astype(float)
        dd = (xyz[:, None] - xyz[None]) @ rot.T
        h = np.sqrt((dd ** 2).sum(-1)) / 30.0
        K = np.where(h < 1, 1 - 1.5 * h + 0.5 * h ** 3, 0.0)
        nugget = -np.linalg.eigvalsh(K).min() - 5e-9
    params = dict(ixc=1, iyc=2, izc=3, ivc=4, iwc=0, tmin=-1e21, tmax=1e21, itrans=0, skmean=0.0, zlo=-5, zhi=5,
                  ltail=1, ltpar=0.0, utail=1, utpar=0.0, rmax=100.0, rmed=100.0, rvert=100.0, sa1=0.0, sa2=0.0,
                  sa3=0.0, ndmax=100)
    c = dict(origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), grid=(nx, ny, nz), sts=sts, nugget=nugget,
             params=params, rows=rows, nreal=2)
    f = np.random.default_rng(1).standard_normal((2, nz, ny, nx))
    want, out = prepared(c, f)
    oi = out["info"]
    nodes = nx * ny * nz - oi["nexact"]
    assert oi["nfail"] == nodes  # every kriged node went through the ridge rule
    k = device_cond(eng, c, out)
    d = Dev(f)
    try:
        k.apply(d.p.value, S.FP64, 2)
        eng.synchronize()
        got = d.get()
    finally:
        d.free()
    info = k.info()
    k.close()
    assert info["nfail"] == oi["nfail"] and info["nexact"] == oi["nexact"]
    if expect == "unsolved":
        mask = np.ones(f.shape[1:], bool)
        mask[out["ijk"][:, 2] - 1, out["ijk"][:, 1] - 1, out["ijk"][:, 0] - 1] = False
        assert np.array_equal(got[:, mask], f[:, mask])  # unsolved nodes keep the unconditional value
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-8


def test_writer_conditions_chunks(tmp_path, eng):
    say("[cond] writer")
    c = G.cond_case("ns2d")
    sp, f = unconditional(eng, c, S.FP64)
    want, out = prepared(c, f)
    k = device_cond(eng, c, out)
    nx, ny, nz = c["grid"]
    files = {"bin": str(tmp_path / "d.bin"), "gslib": str(tmp_path / "d.out")}
    with S.Writer(eng, files, nx, ny, nz, c["nreal"], S.FP64, c["origin"], c["spacing"]) as w:
        w.condition(k)
        for _ in range(c["nreal"]):
            w.realize(sp, c["iseed"], c["igauss"], 1)
    # the same files from the conditioned host field of the device path
    d = Dev(f)
    try:
        k.apply(d.p.value, S.FP64, c["nreal"])
        eng.synchronize()
        cf = d.get()
    finally:
        d.free()
    ref = {"bin": str(tmp_path / "h.bin"), "gslib": str(tmp_path / "h.out")}
    with S.Writer(None, ref, nx, ny, nz, c["nreal"], S.FP64, c["origin"], c["spacing"]) as w:
        w.put(cf)
    for key in ("bin", "gslib"):
        assert open(files[key], "rb").read() == open(ref[key], "rb").read(), key
    assert np.abs(cf - want).max() / np.abs(want).max() < 1e-10
    k.close()
    sp.close()


@pytest.mark.parametrize("nreal", [19, 40])
def test_many_realizations_per_call(eng, nreal):
    """Calls with 16..31 and 32+ realizations sum 16 or 32 realizations per pass over the weights
    (plus a remainder): the same values as the oracle's."""
    say(f"[cond] {nreal} realizations per call")
    c = G.cond_case("ns2d")
    npg = S.periodic_grid(c["sts"], *c["spacing"], *c["grid"], c["ipad"])
    sp = eng.spectrum(c["sts"], c["nugget"], *c["spacing"], *npg)
    f = eng.realize(sp, 99, 1, 0, nreal, *c["grid"], precision=S.FP64)[0]
    sp.close()
    want, out = prepared(c, f)
    k = device_cond(eng, c, out)
    d = Dev(f)
    try:
        k.apply(d.p.value, S.FP64, nreal)
        eng.synchronize()
        got = d.get()
    finally:
        d.free()
    k.close()
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-10
