"""Request loop (continuous batching): every result is the same alone and in the mix, the
spectral-density cache is hit by repeated models, FP64 spot checks stay within the FP32
tolerance, and a request the engine cannot take fails alone.  A mix that holds requests of the
benchmark's size (256^3 grid, 512^3 periodic) is checked against the reference: the FP64 request
against the reference program's own realizations (tests/golden/c2, 1e-10), the FP32 one within
1e-5 relative RMS of them, and a small FP64 request against the oracle and its golden fixture."""
import sys

import numpy as np
import pytest

import golden_cases as G
import oracle_ffi as O
import specsim3d_amd as S
from test_gpu_configs import check_against_fixture, fixture, rel_max, rel_rms

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


MODELS = [
    [(2, 0.7, 12, 6, 3, 30, 15, 0)],
    [(1, 1.0, 10, 8, 4, 120, -10, 5)],
    [(3, 0.9, 6, 5, 3, 60, 0, 0), (2, 0.1, 20, 20, 20, 0, 0, 0)],
]
# (model index, grid, nreal, igauss, precision, seed)
JOBS = [(0, (48, 40, 20), 5, 1, S.FP32, 11), (1, (64, 30, 1), 7, 0, S.FP32, 12), (0, (48, 40, 20), 3, 0, S.FP32, 13),
        (2, (32, 32, 16), 4, 1, S.FP32, 14), (1, (20, 70, 12), 2, 1, S.FP64, 15), (0, (48, 40, 20), 6, 1, S.FP32, 16),
        (2, (40, 1, 1), 3, 0, S.FP32, 17), (1, (64, 30, 1), 9, 1, S.FP32, 18)]


def alone(eng, job):
    m, grid, nreal, igauss, prec, seed = job
    sts = MODELS[m]
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 1)
    sp = eng.spectrum(sts, 0.1, 1, 1, 1, *npg)
    f, mean, var, _ = eng.realize(sp, seed, igauss, 0, nreal, *grid, prec)
    sp.close()
    return f, mean, var


def test_results_identical_alone_and_in_the_mix():
    say("[server] alone")
    with S.Engine(0) as eng:
        ref = [alone(eng, j) for j in JOBS]
        srv = S.Server(eng, nslots=3)
        outs = []
        try:
            say("[server] mix")
            for i, (m, grid, nreal, igauss, prec, seed) in enumerate(JOBS):
                out = np.empty((nreal, grid[2], grid[1], grid[0]), dtype=np.float32 if prec == S.FP32 else np.float64)
                outs.append(out)
                req = S.Server.request(i, MODELS[m], 0.1, 1, 1, 1, *grid, 1, seed, nreal, igauss, prec,
                                       spot_check=(i % 3 == 0 and prec == S.FP32), host_out=out)
                srv.submit(req, keep=out)
            done = {}
            while len(done) < len(JOBS):
                for c in srv.poll(timeout=1.0):
                    done[c["id"]] = c
        finally:
            srv.close()
    for i, job in enumerate(JOBS):
        c = done[i]
        assert c["status"] == 0, c
        f, mean, var = ref[i]
        assert np.array_equal(outs[i], f), f"request {i} differs from its result alone"
        assert c["mean0"] == mean[0] and c["var0"] == var[0]
        assert c["t_submit"] <= c["t_admit"] <= c["t_done"]
        if i % 3 == 0 and job[4] == S.FP32:
            assert 0 <= c["spot_rel_rms"] < 1e-5, c
    # the second request of a (model, grid) finds the cached spectral density
    assert done[2]["cache_hit"] == 1 and done[5]["cache_hit"] == 1 and done[7]["cache_hit"] == 1
    assert done[0]["cache_hit"] == 0 and done[1]["cache_hit"] == 0


def test_failing_request_does_not_stop_others():
    with S.Engine(0) as eng:
        with S.Server(eng, nslots=2) as srv:
            bad = S.Server.request(1, MODELS[0], 0.0, 1, 1, 1, 20000, 2, 2, 0, 5, 1, 0)  # axis 32768 > 8192
            good = S.Server.request(2, MODELS[0], 0.0, 1, 1, 1, 32, 32, 8, 1, 5, 2, 0)
            srv.submit(bad)
            srv.submit(good)
            done = {}
            while len(done) < 2:
                for c in srv.poll(timeout=1.0):
                    done[c["id"]] = c
    assert done[1]["status"] == 2  # SS3D_ETOOBIG
    assert done[2]["status"] == 0


def test_benchmark_sized_requests_in_the_mix_match_the_reference():
    meta, fx = fixture("c2")
    big = [tuple(v) for v in meta["structures"]]
    nxb, nyb, nzb = meta["grid"]
    small = G.cases()["exp_nug_gauss"]
    # (structures, nugget, grid, seed, nreal, igauss, precision, spot check)
    jobs = [(big, meta["nugget"], (nxb, nyb, nzb), meta["iseed"], 2, meta["igauss"], S.FP64, False),
            (big, meta["nugget"], (nxb, nyb, nzb), meta["iseed"], 2, meta["igauss"], S.FP32, True),
            (small["structures"], small["nugget"], tuple(small["grid"]), small["iseed"], small["nreal"],
             small["igauss"], S.FP64, False),
            (big, meta["nugget"], (nxb, nyb, nzb), 777, 8, 1, S.FP32, False),
            ([(2, 0.7, 90, 35, 12, 10, -5, 0)], 0.3, (nxb, nyb, nzb), 778, 4, 1, S.FP32, False)]
    say("[server] benchmark-sized mix: alone")
    with S.Engine(0) as eng:
        ref = []
        for sts, nug, grid, seed, nreal, ig, prec, _ in jobs:
            npg = S.periodic_grid(sts, 1, 1, 1, *grid, 1)
            sp = eng.spectrum(sts, nug, 1, 1, 1, *npg)
            ref.append(eng.realize(sp, seed, ig, 0, nreal, *grid, prec))
            sp.close()
        say("[server] benchmark-sized mix: in the mix")
        srv = S.Server(eng, nslots=4)
        outs, done = [], {}
        try:
            for i, (sts, nug, grid, seed, nreal, ig, prec, spot) in enumerate(jobs):
                out = np.empty((nreal, grid[2], grid[1], grid[0]), dtype=np.float32 if prec == S.FP32 else np.float64)
                outs.append(out)
                srv.submit(S.Server.request(i, sts, nug, 1, 1, 1, *grid, 1, seed, nreal, ig, prec, spot_check=spot,
                                            host_out=out), keep=out)
            while len(done) < len(jobs):
                for c in srv.poll(timeout=1.0):
                    done[c["id"]] = c
        finally:
            srv.close()
    for i in range(len(jobs)):
        assert done[i]["status"] == 0, done[i]
        assert np.array_equal(outs[i], ref[i][0]), f"request {i} differs from its result alone"
        assert done[i]["mean0"] == ref[i][1][0] and done[i]["var0"] == ref[i][2][0]
    check_against_fixture(outs[0], fx, 1e-10, rel_max)   # FP64 256^3: the reference's realizations
    check_against_fixture(outs[1], fx, 1e-5, rel_rms)    # FP32 256^3: within the FP32 bar of them
    assert 0 <= done[1]["spot_rel_rms"] < 1e-5
    o = O.simulate(*G.args(small))
    assert rel_max(outs[2], o["res"]) < 1e-10 and rel_max(outs[2], G.field("exp_nug_gauss")) < 1e-10
    assert done[3]["cache_hit"] == 1  # same model and grid as requests 0 and 1
