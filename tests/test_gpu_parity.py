"""Parity of the HIP engine (through the C ABI) with the reference and the oracle.

FP64 drop-in path: the reference's FP64 realizations (golden fixtures) to 1e-10 relative
(max-norm) for every model.  The spectral densities are the oracle's bit for bit: the covariance
grid uses the exp of the reference build's libm (ss3d_exp.h) and the forward transform repeats the
reference's butterflies.  This matters most for the Gaussian model, whose spectrum decays below
FP64 rounding: the sign of S at ~40% of its frequencies is rounding noise, so any other exp or
operation order moves those amplitudes (~1e-7 relative RMS in the field).
FP32 path: <= 1e-5 relative RMS against the FP64 field of the same draws.
"""
import sys

import numpy as np
import pytest

import golden_cases as G
from hipmem import DevBuf
import oracle_ffi as O
import specsim3d_amd as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = S.Engine(0)
    yield e
    e.close()


def rel_max(a, ref):
    return float(np.abs(a - ref).max() / np.abs(ref).max())


def rel_rms(a, ref):
    return float(np.sqrt(np.mean((a - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))


@pytest.mark.parametrize("name", sorted(G.cases()))
def test_fp64_matches_reference_fields(eng, name):
    c = G.cases()[name]
    say(f"[fp64] case {name} grid {c['grid']} periodic {c['periodic_grid']}")
    r = S.simulate(eng, *G.args(c))
    ref = G.field(name)
    assert tuple(r["np"]) == tuple(c["periodic_grid"])
    assert r["res"].shape == ref.shape
    assert rel_max(r["res"], ref) < 1e-10
    o = O.simulate(*G.args(c))
    assert np.allclose(r["mean"], o["mean"], rtol=0, atol=1e-9)
    assert np.allclose(r["var"], o["var"], rtol=1e-9, atol=1e-12)
    for (m, v), mo, vo in zip(G.stdout_stats(c), r["mean"], r["var"]):
        assert abs(m - mo) <= 6e-5 and abs(v - vo) <= 6e-5


@pytest.mark.parametrize("name", sorted(G.cases()))
def test_spectrum_matches_oracle(eng, name):
    c = G.cases()[name]
    say(f"[spectrum] case {name}")
    grid = tuple(c["periodic_grid"])
    sp = eng.spectrum(c["structures"], c["nugget"], *c["spacing"], *grid)
    amp = sp.amplitude()
    oamp, _, od = O.spectrum(c["structures"], c["nugget"], *c["spacing"], *grid)
    d = sp.diag
    sp.close()
    assert d["c0"] == od["c0"]
    assert d["spos"] + d["sneg"] == pytest.approx(od["spos"] + od["sneg"], rel=1e-12)
    assert d["max_real"] == pytest.approx(od["max_real"], rel=1e-12)
    # the covariance grid is the reference's (its libm's exp) and the transform repeats the
    # reference's butterflies: amplitudes bit for bit, for every model
    assert np.array_equal(amp, oamp), f"max |diff| {np.abs(amp - oamp).max():g}"


@pytest.mark.parametrize("name", sorted(G.cases()))
def test_fp32_close_to_fp64(eng, name):
    c = G.cases()[name]
    say(f"[fp32] case {name}")
    r32 = S.simulate(eng, *G.args(c), precision=S.FP32)
    ref = G.field(name)
    assert r32["res"].dtype == np.float32
    assert rel_rms(r32["res"].astype(np.float64), ref) < 1e-5
    o = O.simulate(*G.args(c))
    assert np.allclose(r32["mean"], o["mean"], rtol=0, atol=1e-5)
    assert np.allclose(r32["var"], o["var"], rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("precision", [S.FP64, S.FP32])
@pytest.mark.parametrize("igauss", [0, 1])
def test_realization_offsets_are_independent(eng, precision, igauss):
    """Realization r is the same whether computed alone or inside a batch (counter-based stream)."""
    say(f"[offsets] precision {precision} igauss {igauss}")
    sts = [(2, 0.7, 8, 4, 2, 30, 15, 0)]
    grid = S.periodic_grid(sts, 1, 1, 1, 24, 20, 12, 1)
    sp = eng.spectrum(sts, 0.3, 1, 1, 1, *grid)
    full, m, v, _ = eng.realize(sp, 4711, igauss, 0, 5, 24, 20, 12, precision)
    one, m1, v1, _ = eng.realize(sp, 4711, igauss, 3, 1, 24, 20, 12, precision)
    mid, _, _, _ = eng.realize(sp, 4711, igauss, 1, 3, 24, 20, 12, precision)
    sp.close()
    assert np.array_equal(full[3], one[0]) and m[3] == m1[0] and v[3] == v1[0]
    assert np.array_equal(full[1:4], mid)


# (output grid, realizations, C2R radix plan): 256^3 in 512^3 (8 members per block); 250x130x110
# in 512x512x256: 14300 rows in tiles of 8 (plan 3: 6 members per block, then 2) or 16 rows (plan
# 4: 3, 3 and 2 members), the last tile ragged; 251x130x110: an odd row length (4-byte stores)
C2R_LOOP = [((256, 256, 256), 8, 4), ((250, 130, 110), 8, 3), ((250, 130, 110), 8, 4), ((251, 130, 110), 8, 3)]


@pytest.mark.parametrize("grid,nreal,rm", C2R_LOOP)
def test_c2r_member_loop_equals_single_members(eng, grid, nreal, rm):
    """The C2R pass over several batch members per block (the rows of member b+1 are loaded while
    member b is transformed) gives every realization and its statistics bit for bit as the pass
    over one member computes them (a realization alone): one launch over all members here."""
    sts = [(2, 0.7, 80, 40, 10, 30, 15, 0)]
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 1)
    lx = npg[0].bit_length() - 1
    say(f"[c2r loop] grid {grid} periodic {npg} nreal {nreal} plan {rm}")
    old = eng.tuning_get(8, lx - 1)
    sp = eng.spectrum(sts, 0.3, 1, 1, 1, *npg)
    try:
        eng.tuning_set(8, lx - 1, rm)
        eng.profile(True)
        f, m, v, _ = eng.realize(sp, 69069, 1, 0, nreal, *grid, S.FP32)
        eng.profile(False)
        assert eng.profile_read()["real_x_c2r_fp32"][2] == 1  # one launch over every member
        for r in range(nreal):
            one, m1, v1, _ = eng.realize(sp, 69069, 1, r, 1, *grid, S.FP32)
            assert np.array_equal(f[r], one[0]), r
            assert m[r] == m1[0] and v[r] == v1[0], r
    finally:
        eng.tuning_set(8, lx - 1, old)
        sp.close()


@pytest.mark.parametrize("rm", [3, 5])
def test_c2r_output_alignment_does_not_change_values(eng, rm):
    """Realizations written to an output buffer that is only 4-byte aligned (4-byte stores of the
    output pairs) equal those written to an 8-byte aligned one (8-byte stores), bit for bit, with
    their statistics; 250x130x110 in 512x512x256, 8 realizations in one launch."""
    sts = [(2, 0.7, 80, 40, 10, 30, 15, 0)]
    nx, ny, nz, nreal = 250, 130, 110, 8
    npg = S.periodic_grid(sts, 1, 1, 1, nx, ny, nz, 1)
    lx = npg[0].bit_length() - 1
    ncell = nx * ny * nz
    say(f"[c2r align] plan {rm}")
    old = eng.tuning_get(8, lx - 1)
    sp = eng.spectrum(sts, 0.3, 1, 1, 1, *npg)
    out = DevBuf(nreal * ncell * 4 + 8)
    st = DevBuf(2 * nreal * 3 * 8)
    try:
        eng.tuning_set(8, lx - 1, rm)
        got = []
        for off in (0, 4):
            eng.realize_device(sp, 69069, 1, 0, nreal, nx, ny, nz, S.FP32, out.ptr + off, st.ptr + len(got) * nreal * 24)
            eng.synchronize()
            raw = out.get(np.uint8, nreal * ncell * 4 + 8)
            got.append(raw[off:off + nreal * ncell * 4].view(np.float32).copy())
        stats = st.get(np.float64, 2 * nreal * 3).reshape(2, nreal, 3)
    finally:
        eng.tuning_set(8, lx - 1, old)
        out.free()
        st.free()
        sp.close()
    assert np.array_equal(got[0], got[1])
    assert np.array_equal(stats[0], stats[1])
    assert np.all(np.isfinite(got[0])) and 0.5 < stats[0, :, 1].min()


def test_256cubed_full_size_fp32_vs_fp64(eng):
    """256^3 in a 512^3 periodic grid (exponential 0.7 + nugget 0.3, ranges 80/40/10, az 30, dip 15,
    igauss=1): FP32 vs FP64 from the same draws, sill reproduced by the sample variance."""
    sts = [(2, 0.7, 80, 40, 10, 30, 15, 0)]
    grid = S.periodic_grid(sts, 1, 1, 1, 256, 256, 256, 1)
    assert grid == (512, 512, 512)
    say("[c2] spectrum 512^3")
    sp = eng.spectrum(sts, 0.3, 1, 1, 1, *grid)
    d = sp.diag
    assert d["c0"] == pytest.approx(1.0) and d["spos"] + d["sneg"] == pytest.approx(1.0, abs=1e-9)
    assert d["sneg"] > -1e-10
    say("[c2] fp64 realization")
    f64, m64, v64, mi64 = eng.realize(sp, 69069, 1, 0, 1, 256, 256, 256, S.FP64)
    say("[c2] fp32 realizations")
    f32, m32, v32, _ = eng.realize(sp, 69069, 1, 0, 2, 256, 256, 256, S.FP32)
    sp.close()
    assert rel_rms(f32[0].astype(np.float64), f64[0]) < 1e-5
    assert abs(m64[0] - m32[0]) < 1e-5 and abs(v64[0] - v32[0]) < 1e-5
    assert mi64[0] < 1e-10
    assert 0.5 < v64[0] < 1.5 and 0.5 < v32[1] < 1.5
    assert not np.array_equal(f32[0], f32[1])


@pytest.mark.parametrize("rm", [3, 4, 5])
@pytest.mark.parametrize("grid", [(64, 48, 20), (100, 30, 1), (40, 1, 1), (24, 20, 12)])
def test_fp32_radix_variants(eng, rm, grid):
    """Every radix plan the autotuner may pick gives the FP64 field of the same draws."""
    say(f"[radix] rm {rm} grid {grid}")
    sts = [(2, 0.7, 8, 4, 2, 30, 15, 0)]
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 1)
    lx, ly, lz = (v.bit_length() - 1 for v in npg)
    for kind, L in ((6, lz), (7, ly), (8, max(lx - 1, 0))):
        eng.tuning_set(kind, L, rm)
    sp = eng.spectrum(sts, 0.3, 1, 1, 1, *npg)
    f64, m64, v64, _ = eng.realize(sp, 4711, 1, 0, 2, *grid, S.FP64)
    f32, m32, v32, _ = eng.realize(sp, 4711, 1, 0, 2, *grid, S.FP32)
    sp.close()
    for kind, L in ((6, lz), (7, ly), (8, max(lx - 1, 0))):
        eng.tuning_set(kind, L, 5)
    assert rel_rms(f32.astype(np.float64), f64) < 1e-5
    assert np.allclose(m32, m64, atol=1e-5) and np.allclose(v32, v64, rtol=1e-5)


def test_autotune_returns_a_plan(eng):
    sts = [(1, 1.0, 20, 10, 5, 45, 0, 0)]
    grid = S.periodic_grid(sts, 1, 1, 1, 64, 64, 16, 1)
    sp = eng.spectrum(sts, 0.0, 1, 1, 1, *grid)
    plan = eng.autotune(sp, 64, 64, 16, 0, reps=1)
    a, _, _, _ = eng.realize(sp, 69069, 0, 0, 1, 64, 64, 16, S.FP32)
    sp.close()
    assert set(plan.values()) <= {3, 4, 5}
    o = O.simulate(sts, 0.0, 1, 1, 1, 64, 64, 16, 1, 69069, 1, 0)
    assert rel_rms(a.astype(np.float64), o["res"]) < 1e-5


@pytest.mark.parametrize("grid", [(3000, 5, 1), (5, 6000, 1), (4, 3, 2500), (700, 600, 1)])
def test_long_axes(eng, grid):
    """Axes up to 8192 points (the longest the passes take): FP32 vs FP64 of the same draws."""
    say(f"[long] grid {grid}")
    sts = [(2, 1.0, 30, 20, 10, 20, 0, 0)]
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 0)
    sp = eng.spectrum(sts, 0.1, 1, 1, 1, *npg)
    f64, _, v64, _ = eng.realize(sp, 99, 1, 0, 1, *grid, S.FP64)
    f32, _, v32, _ = eng.realize(sp, 99, 1, 0, 1, *grid, S.FP32)
    sp.close()
    assert rel_rms(f32.astype(np.float64), f64) < 1e-5
    assert abs(v32[0] - v64[0]) < 1e-5 * max(1.0, v64[0])


def test_stream_pipeline_gives_identical_results(eng):
    """The two-stream FP32 pipeline (fill of batch k+1 beside the passes of batch k) changes no bit."""
    sts = [(2, 0.7, 20, 10, 5, 30, 15, 0)]
    grid = (100, 90, 40)
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 1)
    sp = eng.spectrum(sts, 0.3, 1, 1, 1, *npg)
    a, ma, va, _ = eng.realize(sp, 77, 1, 0, 40, *grid, S.FP32)
    eng.set_option(S.OPT_PIPELINE, 1)
    try:
        b, mb, vb, _ = eng.realize(sp, 77, 1, 0, 40, *grid, S.FP32)
    finally:
        eng.set_option(S.OPT_PIPELINE, 0)
    sp.close()
    assert np.array_equal(a, b) and np.array_equal(ma, mb) and np.array_equal(va, vb)


FP64_GRIDS = [(64, 64, 16, 1, 1), (40, 30, 1, 1, 0), (13, 7, 5, 0, 1), (1, 16, 8, 0, 1), (33, 1, 1, 1, 1),
              (256, 256, 256, 1, 1), (20, 12, 300, 1, 0), (6, 1400, 3, 1, 1), (700, 3, 2, 1, 1),
              (2000, 5, 1, 0, 0), (3, 1, 40, 1, 1), (2, 2, 2, 0, 1)]


@pytest.mark.parametrize("grid", FP64_GRIDS)
def test_fp64_compile_time_kernels_equal_generic(eng, grid):
    """The compile-time FP64 passes compute the generic pass kernel's values bit for bit."""
    nx, ny, nz, ipad, igauss = grid
    sts = [(2, 0.8, 9.0, 5.0, 3.0, 30.0, 10.0, 0.0), (1, 0.2, 25.0, 25.0, 12.0, 0.0, 0.0, 0.0)]
    npg = S.periodic_grid(sts, 1, 1, 1, nx, ny, nz, ipad)
    say(f"[fp64 kernels] grid {grid} periodic {npg}")
    sp = eng.spectrum(sts, 0.1, 1, 1, 1, *npg)
    nreal = 2 if nx * ny * nz > 1 << 22 else 3
    try:
        f1, m1, v1, i1 = eng.realize(sp, 9191, igauss, 5, nreal, nx, ny, nz, S.FP64)
        eng.set_option(S.OPT_FP64_GENERIC, 1)
        try:
            f2, m2, v2, i2 = eng.realize(sp, 9191, igauss, 5, nreal, nx, ny, nz, S.FP64)
        finally:
            eng.set_option(S.OPT_FP64_GENERIC, 0)
    finally:
        sp.close()
    assert np.array_equal(f1, f2), np.abs(f1 - f2).max()
    # statistics: the same sums over differently shaped blocks
    assert np.allclose(m1, m2, rtol=0, atol=1e-14) and np.allclose(v1, v2, rtol=1e-13)
    assert np.allclose(i1, i2, rtol=0, atol=1e-15)


@pytest.mark.parametrize("rm", [3, 4])
@pytest.mark.parametrize("grid", [(64, 48, 20), (60, 50, 50), (100, 30, 1), (40, 600, 20), (20, 12, 700)])
def test_tile_variants_give_identical_fields(eng, rm, grid):
    """Every tile variant the autotuner may pick (halved tile width up to 2^8, no LDS padding at
    2^10) computes the default tile's fields bit for bit: a column's transform does not depend on
    the tile.  Covers column-tiled fill output (z and y of one length), 2D grids (fill along y) and
    axes of 2^10."""
    sts = [(2, 0.7, 8, 4, 2, 30, 15, 0)]
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 1)
    lx, ly, lz = (v.bit_length() - 1 for v in npg)
    say(f"[tiles] rm {rm} grid {grid} periodic {npg}")
    sp = eng.spectrum(sts, 0.3, 1, 1, 1, *npg)
    try:
        out = {}
        for var in (0, 1):
            for kind, L in ((6, lz), (7, ly)):
                try:
                    eng.tuning_set_tile(kind, L, rm, var)
                except S.SS3DError:  # no variant at this length: default tile
                    eng.tuning_set_tile(kind, L, rm, 0)
            eng.tuning_set(8, max(lx - 1, 0), rm)
            out[var] = eng.realize(sp, 4711, 1, 0, 3, *grid, S.FP32)
            picked = [eng.tuning_get_tile(k, L) for k, L in ((6, lz), (7, ly))]
            say(f"  variant {var}: {picked}")
    finally:
        for kind, L in ((6, lz), (7, ly), (8, max(lx - 1, 0))):
            eng.tuning_set(kind, L, 5)
        sp.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
