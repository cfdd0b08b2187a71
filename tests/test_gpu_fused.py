"""The fused FP32 schedule (SS3D_OPT_FUSE, on by default): the z fill pass of batch k+1 and the y
pass of batch k run in one launch whose blocks interleave.  The launch performs the arithmetic of
the separate passes on the same tiles, so the fields and statistics equal those of separate
launches bit for bit.  The cases exercise the block-role map: more fill
blocks than y blocks and fewer (a short last batch), a tail of one role, lengths 2^5 .. 2^10, both
coefficient laws, a y axis twice the fill's, the instance that prunes the upper half of the last
radix pass (both passes keep at most half their axes) and the one that does not, a grid whose z
and y lengths differ otherwise (no fusion: separate launches either way), and the full-size
512^3 and 1024x1024x512 periodic grids (y twice the fill's length).  The fused pass is what runs:
pass kind 9 appears in the profile only with the option on.
"""
import sys

import numpy as np
import pytest

import specsim3d_amd as S
from hipmem import DevBuf

pytestmark = pytest.mark.gpu

# (structures, nugget, output grid, realizations, igauss, fused) -- periodic grids 128^3 (at most 16
# per batch: 14, 14, 12), 64^3 (at most 64: 35, 35), 128x64x64 (at most 32: 20, 20), 32^3 (at most
# 512: 300, 300), 256x256x64
CASES = [
    ([(2, 0.7, 20.0, 10.0, 5.0, 30.0, 15.0, 0.0)], 0.3, (64, 64, 64), 40, 1, True),
    ([(1, 1.0, 15.0, 15.0, 15.0, 0.0, 0.0, 0.0)], 0.0, (48, 40, 40), 70, 1, True),
    ([(3, 1.0, 6.0, 6.0, 6.0, 0.0, 0.0, 0.0)], 0.0, (60, 30, 30), 40, 0, True),
    ([(2, 1.0, 3.0, 3.0, 3.0, 0.0, 0.0, 0.0)], 0.0, (16, 16, 16), 600, 1, True),
    ([(1, 1.0, 30.0, 30.0, 30.0, 0.0, 0.0, 0.0)], 0.0, (100, 100, 20), 17, 1, False),
    # 128x256x256: z 90 and y 120 of 256, the pruned instance at length 2^8, igauss=0
    ([(1, 0.9, 40.0, 40.0, 40.0, 0.0, 0.0, 0.0)], 0.1, (50, 120, 90), 20, 0, True),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    with S.Engine(0) as e:
        yield e


@pytest.mark.parametrize("sts,nugget,grid,nreal,igauss,fused", CASES)
def test_fused_equals_separate_launches(eng, sts, nugget, grid, nreal, igauss, fused):
    nx, ny, nz = grid
    npg = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    say(f"[fused] grid {grid} periodic {npg} nreal {nreal} igauss {igauss}")
    sp = eng.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
    try:
        got = {}
        for opt in (1, 0):
            eng.set_option(S.OPT_FUSE, opt)
            eng.profile(True)
            got[opt] = eng.realize(sp, 777, igauss, 5, nreal, nx, ny, nz, precision=S.FP32)
            eng.profile(False)
            kinds = eng.profile_read()
            assert ("real_zy_fp32" in kinds) == (fused and opt == 1), kinds
    finally:
        eng.set_option(S.OPT_FUSE, 1)
        sp.close()
    (f1, m1, v1, i1), (f0, m0, v0, i0) = got[1], got[0]
    assert np.array_equal(f1, f0)
    assert np.array_equal(m1, m0) and np.array_equal(v1, v0) and np.array_equal(i1, i0)
    assert np.all(np.isfinite(f1)) and np.all(v1 > 0)


def test_fused_matches_fp64(eng):
    """The fused schedule's FP32 fields against the FP64 drop-in path: <= 1e-5 relative RMS."""
    sts, nugget, (nx, ny, nz), nreal, igauss, _ = CASES[0]
    npg = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    say(f"[fused] fp32 vs fp64 {npg}")
    sp = eng.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
    try:
        f32 = eng.realize(sp, 4242, igauss, 0, nreal, nx, ny, nz, precision=S.FP32)[0]
        f64 = eng.realize(sp, 4242, igauss, 0, nreal, nx, ny, nz, precision=S.FP64)[0]
    finally:
        sp.close()
    rel = np.sqrt(np.mean((f32 - f64) ** 2) / np.mean(f64 ** 2))
    say(f"relative RMS {rel:.2e}")
    assert rel < 1e-5


def test_no_fusion_across_radix_plans(eng):
    """A y pass whose radix plan differs from the fill's is not fused (the fused kernel runs one
    plan): the realizations are those of the separate launches either way."""
    sts, nugget, (nx, ny, nz), nreal, igauss, _ = CASES[0]
    npg = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    ly = int(np.log2(npg[1]))
    say(f"[fused] y plan 2^3 against fill plan 2^4, length 2^{ly}")
    old = {k: eng.tuning_get(k, ly) for k in (6, 7)}
    sp = eng.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
    try:
        eng.tuning_set(6, ly, 4)
        eng.tuning_set(7, ly, 3)
        got = {}
        for opt in (1, 0):
            eng.set_option(S.OPT_FUSE, opt)
            eng.profile(True)
            got[opt] = eng.realize(sp, 31, igauss, 0, nreal, nx, ny, nz, precision=S.FP32)[0]
            eng.profile(False)
            kinds = eng.profile_read()
            assert "real_zy_fp32" not in kinds, kinds
    finally:
        eng.set_option(S.OPT_FUSE, 1)
        for k, v in old.items():
            eng.tuning_set(k, ly, v)
        sp.close()
    assert np.array_equal(got[1], got[0])


def test_autotune_keeps_fusion(eng):
    """The autotuner gives the fill and the y pass of one length a single radix plan, so the tuned
    path runs fused."""
    sts, nugget, (nx, ny, nz), nreal, igauss, _ = CASES[0]
    npg = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    say(f"[fused] autotune {npg}")
    sp = eng.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
    try:
        plan = eng.autotune(sp, nx, ny, nz, igauss, reps=1)
        eng.profile(True)
        eng.realize(sp, 5, igauss, 0, nreal, nx, ny, nz, precision=S.FP32)
        eng.profile(False)
        kinds = eng.profile_read()
    finally:
        sp.close()
    say(f"plan {plan}")
    assert "real_zy_fp32" in kinds, (plan, kinds)


# fill and y pass of different lengths (y twice as long, y plan 3): (structures, grid, realizations,
# fill plan) -- periodic 128 x 256 x 128 and 64 x 1024 x 512
CASES_UP = [
    ([(1, 1.0, 20.0, 20.0, 20.0, 0.0, 0.0, 0.0)], (64, 200, 64), 19, 3),
    ([(2, 1.0, 10.0, 10.0, 10.0, 0.0, 0.0, 0.0)], (40, 600, 300), 20, 4),
    # both passes keep at most half their axes (z 60 of 128, y 120 of 256): the pruned instance
    ([(2, 1.0, 10.0, 40.0, 10.0, 0.0, 0.0, 0.0)], (48, 120, 60), 19, 3),
]


@pytest.mark.parametrize("sts,grid,nreal,rmz", CASES_UP)
def test_fused_longer_y_pass(eng, sts, grid, nreal, rmz):
    """A y pass twice as long as the fill (plan 3, tiles of the fill's thread count) runs fused
    and gives the fields of separate launches bit for bit."""
    nx, ny, nz = grid
    npg = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    lz, ly = int(np.log2(npg[2])), int(np.log2(npg[1]))
    assert ly == lz + 1
    say(f"[fused] fill 2^{lz} plan {rmz} beside y 2^{ly} plan 3, grid {grid}")
    old = {6: eng.tuning_get(6, lz), 7: eng.tuning_get(7, ly)}
    sp = eng.spectrum(sts, 0.0, 1.0, 1.0, 1.0, *npg)
    try:
        eng.tuning_set(6, lz, rmz)
        eng.tuning_set(7, ly, 3)
        got = {}
        for opt in (1, 0):
            eng.set_option(S.OPT_FUSE, opt)
            eng.profile(True)
            got[opt] = eng.realize(sp, 99, 1, 2, nreal, nx, ny, nz, precision=S.FP32)
            eng.profile(False)
            kinds = eng.profile_read()
            assert ("real_zy_fp32" in kinds) == (opt == 1), kinds
    finally:
        eng.set_option(S.OPT_FUSE, 1)
        eng.tuning_set(6, lz, old[6])
        eng.tuning_set(7, ly, old[7])
        sp.close()
    (f1, m1, v1, _), (f0, m0, v0, _) = got[1], got[0]
    assert np.array_equal(f1, f0)
    assert np.array_equal(m1, m0) and np.array_equal(v1, v0)


def full_size_runs(e, sts, nugget, grid, nreal, opts, seed, plans):
    """Realizations 0..nreal-1 in one device call (so that the schedule sees every batch) under each
    SS3D_OPT_FUSE value in `opts`, with the radix plans `plans` ({(kind, logn): rm}) in place:
    {opt: (field, mean, var, maximag)}, {opt: pass kinds run}, and the spectrum."""
    nx, ny, nz = grid
    ncell = nx * ny * nz
    npg = S.periodic_grid(sts, 1.0, 1.0, 1.0, nx, ny, nz, 1)
    sp = e.spectrum(sts, nugget, 1.0, 1.0, 1.0, *npg)
    old = {k: e.tuning_get(*k) for k in plans}
    got, kinds = {}, {}
    try:
        for k, rm in plans.items():
            e.tuning_set(*k, rm)
        with DevBuf(nreal * ncell * 4) as out, DevBuf(nreal * 3 * 8) as stats:
            for opt in opts:
                say(f"[fused] {grid} in {npg}, {nreal} realizations, plans {plans}, option {opt}")
                e.set_option(S.OPT_FUSE, opt)
                e.profile(True)
                e.realize_device(sp, seed, 1, 0, nreal, nx, ny, nz, S.FP32, out.ptr, stats.ptr)
                e.profile(False)
                kinds[opt] = e.profile_read()
                st = stats.get(np.float64, nreal * 3).reshape(nreal, 3)
                f = out.get(np.float32, nreal * ncell).reshape(nreal, nz, ny, nx)
                got[opt] = (f, st[:, 0].copy(), st[:, 1].copy(), st[:, 2].copy())
    except BaseException:
        sp.close()
        raise
    finally:
        e.set_option(S.OPT_FUSE, 1)
        for k, rm in old.items():
            e.tuning_set(*k, rm)
    return sp, got, kinds


def assert_same_as_separate(got):
    f0, m0, v0, i0 = got[0]
    for opt in sorted(got):
        f1, m1, v1, i1 = got[opt]
        assert np.array_equal(f1, f0), opt
        assert np.array_equal(m1, m0) and np.array_equal(v1, v0) and np.array_equal(i1, i0), opt


# (fill plan, x plan) at 512^3: the y pass takes the fill's plan
PLANS_512 = [(4, 3), (3, 5)]


@pytest.mark.parametrize("rmz,rmx", PLANS_512)
def test_fused_full_size(rmz, rmx):
    """The 256^3 output grid in a 512^3 periodic grid (exponential 0.7 + nugget 0.3, ranges 80/40/10,
    az 30, dip 15, igauss=1), 33 realizations in one call (batches 11, 11, 11): the fused
    launches give the fields and statistics of separate
    launches bit for bit, and realizations whose y pass (3), fill and y pass (20) or fill (32) ran
    fused match the FP64 path to 1e-5 relative RMS."""
    sts = [(2, 0.7, 80.0, 40.0, 10.0, 30.0, 15.0, 0.0)]
    plans = {(6, 9): rmz, (7, 9): rmz, (8, 8): rmx}
    with S.Engine(0) as e:
        sp, got, kinds = full_size_runs(e, sts, 0.3, (256, 256, 256), 33, (1, 0), 69069, plans)
        try:
            assert kinds[1].get("real_zy_fp32", (0, 0, 0))[2] == 2, kinds[1]
            assert "real_zy_fp32" not in kinds[0], kinds[0]
            say("[fused] full size: fp64 of realizations 3, 20 and 32")
            f64 = {r: e.realize(sp, 69069, 1, r, 1, 256, 256, 256, precision=S.FP64)[0][0] for r in (3, 20, 32)}
        finally:
            sp.close()
    assert_same_as_separate(got)
    for r, ref in f64.items():
        rel = np.sqrt(np.mean((got[1][0][r].astype(np.float64) - ref) ** 2) / np.mean(ref ** 2))
        say(f"realization {r}: relative RMS {rel:.2e}")
        assert rel < 1e-5
    v = got[1][2]
    assert 0.5 < v.min() and v.max() < 1.5


def test_fused_full_size_longer_y_axis():
    """512x512x256 in 1024x1024x512 (Gaussian sill 1, ranges 150/60/20, az 120, dip -20, igauss=1):
    the fused launch with a y pass twice the fill's length (2^10, plan 3, beside 2^9, plan 4); 30
    realizations (batches 10, 10, 10) give the fields and statistics of separate launches bit for bit."""
    sts = [(3, 1.0, 150.0, 60.0, 20.0, 120.0, -20.0, 0.0)]
    plans = {(6, 9): 4, (7, 10): 3, (8, 9): 4}
    with S.Engine(0) as e:
        sp, got, kinds = full_size_runs(e, sts, 0.0, (512, 512, 256), 30, (1, 0), 4711, plans)
        sp.close()
    assert kinds[1].get("real_zy_fp32", (0, 0, 0))[2] == 2, kinds[1]
    assert "real_zy_fp32" not in kinds[0], kinds[0]
    assert_same_as_separate(got)
    v = got[1][2]
    assert 0.3 < v.min() and v.max() < 2.0
