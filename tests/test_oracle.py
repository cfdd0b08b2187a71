"""Pins the CPU oracle (oracle/oracle.c) before anything is compared with it.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a wrong sign
or index, a transposed operand, an off-by-one in the draw order) fails at least one of them:
  - the reference's own FP64 realizations (tests/golden, written by the reference build);
  - textbook known answers (Park-Miller, DFT by brute force) and closed forms;
  - a brute-force walk of the fill loop for the draw-index closed form.
"""
import math

import numpy as np
import pytest

import golden_cases as G
import oracle_ffi as O

PM_M = 2147483647


@pytest.mark.parametrize("name", sorted(G.cases()))
def test_golden_fields_bit_exact(name):
    """FP64 realizations equal the reference build's, bit for bit (all models, igauss, ipad)."""
    c = G.cases()[name]
    r = O.simulate(*G.args(c))
    ref = G.field(name)
    assert tuple(r["np"]) == tuple(c["periodic_grid"])
    assert np.array_equal(r["res"], ref), f"max |diff| {np.abs(r['res'] - ref).max():g}"
    for (m, v), mo, vo in zip(G.stdout_stats(c), r["mean"], r["var"]):
        assert f"{m:9.4f}" == f"{mo:9.4f}" and f"{v:9.4f}" == f"{vo:9.4f}"


def test_golden_spectrum_diagnostics():
    """Stdout diagnostic lines of the reference (f90:1289-1292) from the oracle's spectrum."""
    for name, c in G.cases().items():
        r = O.simulate(*G.args(c))
        d = r["diag"]
        lines = {ln.split(":")[0].strip(): ln.split(":")[1].strip() for ln in c["stdout"] if ":" in ln}
        assert f"{d['c0']:10.5f}".strip() == lines["target variance C(0)"], name
        assert f"{d['spos']:10.5f}".strip() == lines["sum of positive spectrum"], name


def test_park_miller_known_answer():
    """Minimal standard generator: from state 1 the 10,000th state is 1043618065."""
    s = 1
    for _ in range(10000):
        s = (16807 * s) % PM_M
    assert s == 1043618065
    # the oracle's urand advances the same recurrence
    st = O.ctypes.c_int64(1)
    for _ in range(10000):
        O.lib().or_urand(O.ctypes.byref(st))
    assert st.value == 1043618065
    assert O.lib().or_pm_pow(16807, 10000) == 1043618065


@pytest.mark.parametrize("iseed", [0, 1, 69069, -4711, 2147483646, 2147483647, -2147483647])
def test_seed_and_burn_in(iseed):
    """f90:1302-1303: state = mod(|iseed|, 2147483646) + 1, then 10 draws are discarded."""
    s = abs(iseed) % 2147483646 + 1
    for _ in range(10):
        s = (16807 * s) % PM_M
    assert O.seed_state(iseed) == s


def test_gauss_pair_is_box_muller():
    st = O.ctypes.c_int64(12345)
    g1, g2 = O.ctypes.c_double(), O.ctypes.c_double()
    O.lib().or_gauss_pair(O.ctypes.byref(st), O.ctypes.byref(g1), O.ctypes.byref(g2))
    s1 = (16807 * 12345) % PM_M
    s2 = (16807 * s1) % PM_M
    u1, u2 = s1 / PM_M, s2 / PM_M
    rr = math.sqrt(-2.0 * math.log(u1))
    assert g1.value == rr * math.cos(2 * math.pi * u2) and g2.value == rr * math.sin(2 * math.pi * u2)


GRIDS = [(8, 8, 8), (16, 8, 4), (4, 16, 2), (2, 2, 2), (8, 1, 1), (1, 8, 1), (1, 1, 8), (1, 4, 8), (8, 4, 1),
         (2, 1, 4), (1, 1, 1), (32, 2, 16)]


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("igauss", [0, 1])
def test_draw_index_closed_form(grid, igauss):
    """Closed-form draw index == the index the serial walk f90:1305-1334 actually uses."""
    npx, npy, npz = grid
    amp = np.ones((npz, npy, npx))
    s0 = 987654
    c, s1, log = O.fill(amp, igauss, s0)
    assert (log >= 0).all(), "every point is written by the walk"
    d = O.lib().or_draws_per_realization(npx, npy, npz, igauss)
    assert O.lib().or_pm_pow(16807, d) * s0 % PM_M == s1, "draws per realization"
    for k in range(npz):
        for j in range(npy):
            for i in range(npx):
                assert O.draw_index(i, j, k, npx, npy, npz, igauss) == log[k, j, i], (i, j, k)


@pytest.mark.parametrize("grid", [(8, 8, 8), (16, 4, 2), (2, 2, 2), (8, 1, 1)])
def test_fill_is_hermitian(grid):
    """c(-k) = conj(c(k)) when amp is symmetric (each pair shares one draw); self-conjugate points are real."""
    npx, npy, npz = grid
    rng = np.random.default_rng(1)
    a = rng.random((npz, npy, npx))
    amp = 0.5 * (a + np.roll(np.flip(a, (0, 1, 2)), 1, (0, 1, 2)))
    for ig in (0, 1):
        c, _, _ = O.fill(amp, ig, 4242)
        partner = np.roll(np.flip(c, (0, 1, 2)), 1, (0, 1, 2))
        assert np.allclose(c, np.conj(partner), rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128])
@pytest.mark.parametrize("isign", [-1, 1])
def test_fft1d_is_unscaled_dft(n, isign):
    """fft1d(x) = sum_j x_j exp(isign 2 pi i jk/n), no scaling (brute-force DFT)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    dft = (x[None, :] * np.exp(isign * 2j * np.pi * np.outer(j, j) / n)).sum(axis=1)
    assert np.allclose(O.fft1d(x, isign), dft, rtol=0, atol=1e-12 * max(1, n))


def test_fft3d_axes_and_signs():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((4, 8, 16)) + 1j * rng.standard_normal((4, 8, 16))
    assert np.allclose(O.fft3d(a, -1), np.fft.fftn(a), atol=1e-11)
    assert np.allclose(O.fft3d(a, +1), np.fft.ifftn(a) * a.size, atol=1e-11)
    # a delta transforms to a constant
    d = np.zeros((4, 8, 16), complex)
    d[0, 0, 0] = 1.0
    assert np.array_equal(O.fft3d(d, -1), np.ones_like(d))


def test_twiddles_formula():
    for n in (2, 8, 64):
        for isign in (-1, 1):
            tw = O.twiddles(n, isign)
            k = np.arange(n // 2)
            assert np.array_equal(tw.real, np.cos(2.0 * np.pi * k / n))
            assert np.array_equal(tw.imag, isign * np.sin(2.0 * np.pi * k / n))


def test_setrot_major_axis():
    """tests/test_rotation.py:49-52 convention: az 45, dip +35.26 puts the major axis along (1,1,1)."""
    dip = math.degrees(math.atan(1 / math.sqrt(2)))
    rot, axes = O.setrot((1, 1.0, 20, 10, 5, 45, dip, 0))
    assert np.allclose(rot[0], np.ones(3) / math.sqrt(3), atol=1e-14)
    rot, _ = O.setrot((1, 1.0, 20, 10, 5, 45, -dip, 0))
    assert np.allclose(rot[0], np.array([1, 1, -1]) / math.sqrt(3), atol=1e-14)
    # rows 2 and 3 are scaled by ahmax/ahmin and ahmax/avert and orthogonal to the major axis
    rot, axes = O.setrot((1, 1.0, 20, 10, 5, 30, 15, 20))
    assert np.allclose(np.linalg.norm(rot, axis=1), [1, 2, 4])
    assert np.allclose(np.linalg.norm(axes, axis=1), [20, 10, 5])
    u = rot / np.linalg.norm(rot, axis=1)[:, None]
    assert np.allclose(u @ u.T, np.eye(3), atol=1e-14)
    # azimuth >= 270 uses 450 - az: same direction as az - 360
    assert np.allclose(O.setrot((1, 1, 9, 3, 2, 300, 25, -40))[0], O.setrot((1, 1, 9, 3, 2, -60, 25, -40))[0], atol=1e-14)


def test_covariance_closed_forms():
    az = 30.0
    u = np.array([math.sin(math.radians(az)), math.cos(math.radians(az)), 0.0])  # azimuth from North
    for itype, at_range in ((1, 0.0), (2, math.exp(-3.0)), (3, math.exp(-3.0))):
        st = [(itype, 0.8, 12, 6, 3, az, 0, 0)]
        assert O.covariance(st, 0.2, 0, 0, 0) == pytest.approx(1.0, abs=1e-15)
        h = 12 * u
        assert O.covariance(st, 0.2, *h) == pytest.approx(0.8 * at_range, abs=1e-14)
        h = 6 * np.array([u[1], -u[0], 0.0])  # minor horizontal axis, range 6
        assert O.covariance(st, 0.2, *h) == pytest.approx(0.8 * at_range, abs=1e-14)
        assert O.covariance(st, 0.2, 0, 0, 3) == pytest.approx(0.8 * at_range, abs=1e-14)
    # spherical: 1 - 1.5 r + 0.5 r^3 inside the range, 0 beyond
    st = [(1, 1.0, 10, 10, 10, 0, 0, 0)]
    assert O.covariance(st, 0.0, 5, 0, 0) == pytest.approx(1 - 0.75 + 0.0625, abs=1e-15)
    assert O.covariance(st, 0.0, 10.5, 0, 0) == 0.0
    # nested structures add
    st2 = [(2, 0.5, 8, 8, 8, 0, 0, 0), (3, 0.3, 4, 4, 4, 0, 0, 0)]
    v = O.covariance(st2, 0.1, 2, 0, 0)
    assert v == pytest.approx(0.5 * math.exp(-0.75) + 0.3 * math.exp(-0.75), abs=1e-15)


@pytest.mark.parametrize("itype", [1, 2])
def test_spectrum_invariants(itype):
    sts = [(itype, 0.7, 10, 6, 4, 30, 15, 0)]
    amp, S, d = O.spectrum(sts, 0.3, 1, 1, 1, 32, 32, 16)
    assert d["c0"] == pytest.approx(1.0)
    assert d["spos"] + d["sneg"] == pytest.approx(d["c0"], abs=1e-12)  # sum S = C(0)
    assert d["sneg"] > -1e-12  # no negative mass for spherical / exponential
    Sre = S.real
    part = np.roll(np.flip(Sre, (0, 1, 2)), 1, (0, 1, 2))
    assert np.allclose(Sre, part, atol=1e-16)
    assert np.array_equal(amp, np.sqrt(np.maximum(Sre, 0.0)) * (Sre >= 0))


def test_periodic_grid_rule():
    sts = [(1, 1.0, 20, 10, 5, 45, 0, 0)]
    assert O.periodic_grid(sts, 1, 1, 1, 64, 64, 16, 1)[0] == (128, 128, 32)
    assert O.periodic_grid(sts, 1, 1, 1, 64, 64, 16, 0)[0] == (64, 64, 16)
    assert O.periodic_grid(sts, 1, 1, 1, 100, 60, 1, 0)[0] == (128, 64, 1)
    assert O.periodic_grid([(2, 0.7, 80, 40, 10, 30, 15, 0)], 1, 1, 1, 256, 256, 256, 1)[0] == (512, 512, 512)
    g, rc = O.periodic_grid([(3, 1.0, 150, 60, 20, 120, -20, 0)], 1, 1, 1, 512, 512, 256, 1)
    assert g == (1024, 1024, 512) and rc == 1  # beyond the reference's 2^27 limit


def test_fixed_amplitudes_give_deterministic_variance():
    """Legacy grid (ipad=0, grid = periodic grid), igauss=0: every realization has the same sample
    variance (Parseval; tests/test_gauss_amplitudes.py:56)."""
    r = O.simulate([(2, 1.0, 5, 5, 5, 0, 0, 0)], 0.0, 1, 1, 1, 16, 16, 16, 0, 4711, 6, 0)
    assert np.std(r["var"]) < 1e-12
    r = O.simulate([(2, 1.0, 5, 5, 5, 0, 0, 0)], 0.0, 1, 1, 1, 16, 16, 16, 0, 4711, 6, 1)
    assert np.std(r["var"]) > 1e-3


def _numpy_vario(f, lx, ly, lz):
    """Brute force over shifted slices (f: nz, ny, nx): 0.5 mean (f(x+l) - f(x))^2."""
    nz, ny, nx = f.shape
    if abs(lx) >= nx or abs(ly) >= ny or abs(lz) >= nz:
        return 0.0
    a = f[max(0, -lz):nz - max(0, lz), max(0, -ly):ny - max(0, ly), max(0, -lx):nx - max(0, lx)]
    b = f[max(0, lz):nz - max(0, -lz), max(0, ly):ny - max(0, -ly), max(0, lx):nx - max(0, -lx)]
    return 0.5 * float(np.mean((b - a) ** 2))


@pytest.mark.parametrize("lag", [(1, 0, 0), (0, 3, 0), (0, 0, 2), (2, -1, 1), (-3, 2, -1), (9, 0, 0), (0, 0, 5)])
def test_vario_lag_brute_force(lag):
    """vario_lag (f90:405-431) equals the shifted-slice mean, and is even in the lag."""
    rng = np.random.default_rng(7)
    f = rng.standard_normal((5, 6, 9))
    g = O.vario_lag(f, *lag)
    assert g == pytest.approx(_numpy_vario(f, *lag), rel=1e-13, abs=0.0)
    assert O.vario_lag(f, *(-v for v in lag)) == pytest.approx(g, rel=1e-13)
    nz, ny, nx = f.shape
    exp_np = (nx - abs(lag[0])) * (ny - abs(lag[1])) * (nz - abs(lag[2])) if all(
        abs(v) < n for v, n in zip(lag, (nx, ny, nz))) else 0
    assert O.npair_count(nx, ny, nz, *lag) == exp_np


def test_vario_known_answers():
    """A linear ramp along x has gamma(l) = l^2/2 exactly; a constant field has 0; no pairs gives 0."""
    f = np.tile(np.arange(7.0), (3, 4, 1))
    for l in range(1, 7):
        assert O.vario_lag(f, l, 0, 0) == 0.5 * l * l
    assert O.vario_lag(np.full((3, 4, 7), 2.5), 1, 1, 1) == 0.0
    assert O.vario_lag(f, 7, 0, 0) == 0.0 and O.npair_count(7, 4, 3, 0, 4, 0) == 0


@pytest.mark.parametrize("name", sorted(G.vario_cases()))
def test_vario_matches_reference_table(name):
    """The oracle's variograms of the reference's FP64 realizations reproduce the reference's
    --vario table (es14.6): pairs, lag distance, model, per-realization values, mean and sd."""
    c = G.vario_cases()[name]
    f = G.vario_field(name)
    head, rows = G.vario_table(name)
    assert head[:10] == ["dir", "sx", "sy", "sz", "lag", "h", "npairs", "gamma_model", "gamma_mean", "gamma_sd"]
    nx, ny, nz = c["grid"]
    dx, dy, dz = c["spacing"]
    c0 = c["nugget"] + sum(s[1] for s in c["structures"])
    expected = []
    for d, (sx, sy, sz, nlag) in enumerate(c["dirs"], 1):
        for il in range(1, nlag + 1):
            lx, ly, lz = il * sx, il * sy, il * sz
            npairs = O.npair_count(nx, ny, nz, lx, ly, lz)
            if npairs == 0:
                continue
            g = np.array([O.vario_lag(f[r], lx, ly, lz) for r in range(c["nreal"])])
            h = math.sqrt((lx * dx) ** 2 + (ly * dy) ** 2 + (lz * dz) ** 2)
            model = c0 - O.covariance(c["structures"], c["nugget"], lx * dx, ly * dy, lz * dz)
            sd = float(np.sqrt(np.sum((g - g.mean()) ** 2) / (len(g) - 1))) if len(g) > 1 else 0.0
            expected.append((d, sx, sy, sz, il, h, npairs, model, g.mean(), sd, g))
    assert len(rows) == len(expected)
    for row, (d, sx, sy, sz, il, h, npairs, model, gm, sd, g) in zip(rows, expected):
        assert [int(v) for v in row[:5]] == [d, sx, sy, sz, il] and int(row[6]) == npairs
        got = [float(v) for v in row[5:6] + row[7:]]
        want = [h, model, gm, sd] + (list(g) if c["iper"] else [])
        assert len(got) == len(want)
        for a, b in zip(got, want):
            assert a == pytest.approx(b, rel=6e-7, abs=1e-300), (name, row)
