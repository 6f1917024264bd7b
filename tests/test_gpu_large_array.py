"""Arrays beyond what one LDS stage of a block's antenna-solve partials holds
(about 13,000 baselines, 160 stations): the antenna solve stages the
partials in chunks of baselines and walks each antenna's baseline list with
a cursor, in the same order as a single stage. 200 stations (19,900
baselines) against the CPU oracle, streaming and Gram iterations, and a batch
against single solves."""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from oracle import pyoracle as O
from problems import phase_ref_relerr, small_problem

pytestmark = pytest.mark.gpu

GAIN_TOL = 1e-4
CHI2_TOL = 1e-3


@pytest.fixture(scope="module")
def large():
    w = S.Workload("large200", 200, 3, 1, 1, 4, 2, 2, 200, 1e-5, seed=7)
    return small_problem(w)


@pytest.mark.parametrize("iteration", [H.ITERATION_STREAM, H.ITERATION_GRAM])
def test_two_hundred_stations_against_oracle(large, iteration):
    w, geom, sky, gains, P = large
    print(f"case: 200 stations, iteration {iteration}", flush=True)
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 max_iterations=w.max_iterations, accuracy=w.tolerance, iteration=iteration)
    sol = P.unit_solutions()
    r = s.solve(P, sol)
    s.close()
    ref = P.unit_solutions()
    rr = O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance, n_threads=8)
    e = phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub)
    c_gpu, c_ref = O.chi2(P, sol), O.chi2(P, ref)
    print(f"iterations gpu {r.iterations} oracle {rr.iterations}, gain error {e:.2e}, "
          f"chi2 {c_gpu:.6g} / {c_ref:.6g}")
    assert r.iterations == rr.iterations
    assert e <= GAIN_TOL
    assert abs(c_gpu - c_ref) <= CHI2_TOL * c_ref


def test_two_hundred_stations_batch_equals_single(large):
    w, geom, sky, gains, P = large
    P2 = small_problem(S.Workload("large200b", 200, 3, 1, 1, 4, 2, 2, 200, 1e-5, seed=8))[4]
    print("case: 200 stations, batch of 2", flush=True)
    problems = [P, P2]
    singles = []
    for X in problems:
        s = H.Solver(X.n_antennas, X.n_solutions_per_direction, X.n_channel_blocks,
                     max_iterations=w.max_iterations, accuracy=w.tolerance)
        s.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=0)
        g = X.unit_solutions()
        r = s.solve(X, g)
        singles.append((g, r.iterations))
        s.close()
    solvers = [H.Solver(X.n_antennas, X.n_solutions_per_direction, X.n_channel_blocks,
                        max_iterations=w.max_iterations, accuracy=w.tolerance) for X in problems]
    dps = [H.DeviceProblem.from_problem(s, X) for s, X in zip(solvers, problems)]
    for dp, X in zip(dps, problems):
        dp.upload(X)
    b = H.Batch(solvers, dps)
    b.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=1, poll_interval=4)
    got = [X.unit_solutions() for X in problems]
    res = b.solve(got)
    for (g, its), x, r in zip(singles, got, res):
        assert r.iterations == its
        np.testing.assert_array_equal(x, g)
    b.close()
    for x in dps + solvers:
        x.close()
