"""The C ABI without a GPU: the library loads, exports what include/specsim3d.h declares, and its
host-only entry points (grid sizing, stream bookkeeping, argument checks) agree with the oracle."""
import os
import re

import numpy as np
import pytest

import oracle_ffi as O
import specsim3d_amd as S

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "specsim3d.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|double)\s+(ss3d_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 14
    L = S.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(S.SIGNATURES), "Python binding and header disagree"


def test_version():
    assert S.lib().ss3d_version() >= 10000


def random_models(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        nst = int(rng.integers(1, 4))
        sts = [(int(rng.integers(1, 4)), float(rng.uniform(0.1, 1)), *sorted(rng.uniform(2, 60, 3))[::-1],
                float(rng.uniform(0, 360)), float(rng.uniform(-60, 60)), float(rng.uniform(-90, 90)))
               for _ in range(nst)]
        grid = tuple(int(v) for v in rng.integers(1, 200, 3))
        d = tuple(float(v) for v in rng.uniform(0.3, 3, 3))
        yield sts, d, grid, int(rng.integers(0, 2))


def test_periodic_grid_matches_oracle():
    for sts, d, grid, ipad in random_models(300, 5):
        ref, _ = O.periodic_grid(sts, *d, *grid, ipad)
        assert S.periodic_grid(sts, *d, *grid, ipad) == ref, (sts, d, grid, ipad)


@pytest.mark.parametrize("grid", [(128, 128, 32), (512, 512, 512), (64, 32, 1), (1, 16, 8), (32, 1, 1)])
@pytest.mark.parametrize("igauss", [0, 1])
def test_stream_bookkeeping_matches_oracle(grid, igauss):
    assert S.draws_per_realization(*grid, igauss) == O.lib().or_draws_per_realization(*grid, igauss)
    d = O.lib().or_draws_per_realization(*grid, igauss)
    for iseed in (69069, -5, 0):
        for r in (0, 1, 7, 1000):
            want = O.seed_state(iseed) * O.lib().or_pm_pow(16807, r * d) % 2147483647
            assert S.realization_state(iseed, *grid, igauss, r) == want


def test_argument_errors_are_reported():
    with pytest.raises(S.SS3DError) as e:
        S.periodic_grid([], 1, 1, 1, 8, 8, 8, 1)
    assert e.value.code == 1 and "nstruct" in str(e.value)
    with pytest.raises(S.SS3DError) as e:
        S.periodic_grid([(1, 1, 0, 1, 1, 0, 0, 0)], 1, 1, 1, 8, 8, 8, 1)
    assert "ranges must be positive" in str(e.value)
    with pytest.raises(S.SS3DError):
        S.periodic_grid([(4, 1, 1, 1, 1, 0, 0, 0)], 1, 1, 1, 8, 8, 8, 1)
    with pytest.raises(S.SS3DError):
        S.periodic_grid([(1, 1, 1, 1, 1, 0, 0, 0)], 1, 1, 1, 8, 8, 8, 2)


def test_host_covariance_and_pairs_equal_oracle():
    """ss3d_covariance / ss3d_npairs are host functions: no GPU needed."""
    import oracle_ffi as O
    rng = np.random.default_rng(11)
    for _ in range(200):
        nst = int(rng.integers(1, 4))
        sts = [(int(rng.integers(1, 4)), float(rng.uniform(0.1, 1)), *map(float, rng.uniform(1, 50, 3)),
                *map(float, rng.uniform(-180, 360, 3))) for _ in range(nst)]
        h = [float(v) for v in rng.uniform(-40, 40, 3)] if rng.random() < 0.9 else [0.0, 0.0, 0.0]
        nug = float(rng.uniform(0, 0.5))
        assert S.covariance(sts, nug, *h) == O.covariance(sts, nug, *h)
        dims = [int(v) for v in rng.integers(1, 20, 3)]
        lag = [int(v) for v in rng.integers(-20, 20, 3)]
        assert S.npairs(*dims, *lag) == O.npair_count(*dims, *lag)
