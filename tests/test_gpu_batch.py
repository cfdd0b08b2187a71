"""Uneven-grid batch replayed as one HIP graph: every replay (with other device work between
replays) gives the same fields and statistics as the eager enqueue, as each job simulated
alone, and, for a small job, as the oracle within the FP32 tolerance."""
import sys

import numpy as np
import pytest

import oracle_ffi as O
import specsim3d_amd as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


MODELS = [
    [(1, 1.0, 10, 8, 4, 120, -10, 5)],
    [(2, 0.8, 14, 9, 5, 30, 15, 0)],
    [(3, 0.9, 6, 5, 3, 60, 0, 0), (2, 0.1, 20, 20, 20, 0, 0, 0)],
]
# (model, nugget, grid, nreal, igauss, seed): uneven grids, repeated (model, grid) pairs, 1-D and 2-D
JOBS = [(0, 0.1, (32, 32, 16), 3, 1, 101), (1, 0.0, (64, 16, 1), 5, 0, 102), (2, 0.2, (20, 30, 12), 2, 1, 103),
        (0, 0.1, (32, 32, 16), 4, 0, 104), (1, 0.3, (48, 40, 24), 1, 1, 105), (2, 0.0, (40, 1, 1), 6, 1, 106),
        (0, 0.05, (16, 64, 8), 2, 1, 107), (1, 0.1, (24, 24, 24), 3, 1, 108)]


def requests():
    return [S.Server.request(i, MODELS[m], nug, 1, 1, 1, *grid, 1, seed, nreal, ig, S.FP32)
            for i, (m, nug, grid, nreal, ig, seed) in enumerate(JOBS)]


def alone(eng, job):
    m, nug, grid, nreal, ig, seed = job
    sts = MODELS[m]
    npg = S.periodic_grid(sts, 1, 1, 1, *grid, 1)
    sp = eng.spectrum(sts, nug, 1, 1, 1, *npg)
    f, mean, var, _ = eng.realize(sp, seed, ig, 0, nreal, *grid, S.FP32)
    sp.close()
    return f, mean, var


@pytest.mark.parametrize("nlanes", [1, 4, 8])
def test_graph_replays_match_eager_and_alone(nlanes):
    say(f"[batch] nlanes={nlanes} create")
    with S.Engine(0) as eng:
        with S.Batch(eng, requests(), nlanes=nlanes) as b:
            assert b.cells == sum(g[0] * g[1] * g[2] * n for _, _, g, n, _, _ in JOBS)
            eng.synchronize()
            first = [b.output(i) for i in range(len(JOBS))]
            ref = [alone(eng, j) for j in JOBS]
            for i in range(len(JOBS)):
                f, st = first[i]
                assert np.array_equal(f, ref[i][0]), f"job {i}: batch differs from the job alone"
                assert np.array_equal(st[:, 0], ref[i][1]) and np.array_equal(st[:, 1], ref[i][2])
            for rep in range(3):
                say(f"[batch] nlanes={nlanes} replay {rep}")
                b.run(graph=True)
                eng.synchronize()
                for i in range(len(JOBS)):
                    f, st = b.output(i)
                    assert np.array_equal(f, first[i][0]), f"replay {rep} job {i} differs"
                    assert np.array_equal(st, first[i][1])
                # other device work between replays: one job alone and the eager batch
                alone(eng, JOBS[rep])
                b.run(graph=False)
                eng.synchronize()
                for i in range(len(JOBS)):
                    assert np.array_equal(b.output(i)[0], first[i][0]), f"eager after replay {rep} job {i} differs"


@pytest.mark.parametrize("job", [1, 4])
def test_batch_job_matches_oracle(job):
    m, nug, grid, nreal, ig, seed = JOBS[job]
    say(f"[batch] oracle job {job}")
    with S.Engine(0) as eng:
        with S.Batch(eng, requests(), nlanes=2) as b:
            b.run(graph=True)
            eng.synchronize()
            f, _ = b.output(job)
    ref = O.simulate(MODELS[m], nug, 1, 1, 1, *grid, 1, seed, nreal, ig)["res"]
    rms = np.sqrt(np.mean((f - ref) ** 2) / np.mean(ref ** 2))
    assert rms < 1e-5, rms  # FP32 path tolerance (relative RMS)


def test_batch_rejects_fp64_and_bad_lanes():
    reqs = requests()
    reqs[1].precision = S.FP64
    with S.Engine(0) as eng:
        with pytest.raises(S.SS3DError):
            S.Batch(eng, reqs)
        with pytest.raises(S.SS3DError):
            S.Batch(eng, requests(), nlanes=0)
