"""Batched solves of uneven solution intervals: every interval gives exactly
what its own single solve gives under the same schedule, and matches the CPU
oracle to the parity tolerances; the autotuner returns a valid schedule."""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from dp3_amd import warmstart as WS
from oracle import pyoracle as O
from problems import phase_ref_relerr, small_problem

pytestmark = pytest.mark.gpu

SHAPES = [  # stations, directions, channels, blocks, solint, sols per direction
    (10, 3, 16, 2, 4, None),
    (14, 2, 12, 3, 2, None),
    (8, 5, 8, 1, 6, (1, 2, 3, 1, 2)),
    (12, 4, 24, 4, 3, None),
    (9, 6, 10, 2, 5, None),
    (16, 3, 20, 5, 2, (2, 1, 1)),
]


@pytest.fixture(scope="module")
def intervals():
    out = []
    for i, (st, nd, nch, ncb, solint, nsd) in enumerate(SHAPES):
        w = S.Workload(f"b{i}", st, nd, 1, 1, nch, ncb, solint, 40, 1e-6,
                       n_solutions_per_direction=nsd, seed=100 + i)
        _, _, _, gains, P = small_problem(w)
        out.append((w, P))
    return out


def make(P, w, **kw):
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 max_iterations=w.max_iterations, accuracy=w.tolerance, **kw)
    dp = H.DeviceProblem.from_problem(s, P)
    dp.upload(P)
    return s, dp


@pytest.mark.parametrize("waves,rows", [(2, 1), (4, 2), (1, 2)])
def test_batch_equals_single_solves(intervals, waves, rows):
    print(f"case: batch of {len(intervals)} waves={waves} rows={rows}", flush=True)
    singles, pairs = [], []
    for w, P in intervals:
        s, dp = make(P, w)
        s.set_schedule(waves_per_tile=waves, rows_per_trip=rows, use_graph=0)
        sol = P.unit_solutions()
        r = s.solve_problem(dp, sol)
        singles.append((sol, r))
        pairs.append((s, dp))
    b = H.Batch([p[0] for p in pairs], [p[1] for p in pairs])
    b.set_schedule(waves_per_tile=waves, rows_per_trip=rows, use_graph=1, poll_interval=4)
    sols = [P.unit_solutions() for _, P in intervals]
    res = b.solve(sols)
    for i, ((w, P), (ref, r1), sol, r2) in enumerate(zip(intervals, singles, sols, res)):
        assert r1.iterations == r2.iterations, i
        np.testing.assert_array_equal(sol, ref)
    # and against the CPU oracle
    for i, ((w, P), sol, r) in enumerate(zip(intervals, sols, res)):
        ref = P.unit_solutions()
        rr = O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance)
        assert r.iterations == rr.iterations, (i, r.iterations, rr.iterations)
        assert phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub) <= 1e-4
    # a second batched solve reuses the captured graph
    sols2 = [P.unit_solutions() for _, P in intervals]
    b.solve(sols2)
    for a, c in zip(sols, sols2):
        np.testing.assert_array_equal(a, c)


def test_autotune(intervals):
    print("case: autotune", flush=True)
    pairs = [make(P, w) for w, P in intervals]
    b = H.Batch([p[0] for p in pairs], [p[1] for p in pairs])
    waves, rows, ms = b.autotune()
    assert 1 <= waves <= 4 and rows in (1, 2) and ms > 0
    sols = [P.unit_solutions() for _, P in intervals]
    res = b.solve(sols)
    for (w, P), sol, r in zip(intervals, sols, res):
        ref = P.unit_solutions()
        rr = O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance)
        assert r.iterations == rr.iterations
        assert phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub) <= 1e-4
    s, dp = pairs[0]
    w2, r2, ms2 = H.autotune(s, dp)
    assert 1 <= w2 <= 4 and r2 in (1, 2) and ms2 > 0


def test_gram_batch_equals_single_gram_solves():
    """Batched Gram iterations (cross products per interval, one apply launch
    per direction count) give the same results as single Gram solves."""
    print("case: gram batch", flush=True)
    ws = [S.Workload(f"g{i}", 10 + 3 * i, 2 + i, 1, 1, 8 * (1 + i % 2), 2, 2 + i, 40, 1e-5,
                     seed=i) for i in range(4)]
    seqs = [WS.IntervalSequence(w, 1, seed=50 + i, solver_kwargs=dict(iteration=H.ITERATION_GRAM))
            for i, w in enumerate(ws)]
    singles = []
    for q in seqs:
        q.solver.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=0, poll_interval=8)
        g = q.unit()
        r = q.solver.solve_problem(q.problems[0], g)
        singles.append((g, r.iterations))
    b = H.Batch([q.solver for q in seqs], [q.problems[0] for q in seqs])
    b.set_schedule(2, 1, 1, 8)
    got = [q.unit() for q in seqs]
    res = b.solve(got)
    for (g, its), x, r in zip(singles, got, res):
        assert r.iterations == its
        np.testing.assert_array_equal(x, g)
