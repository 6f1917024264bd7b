"""Antenna solve schedules: every split of an antenna's baseline list over 1..16
threads and every block size computes the same iteration (FP64 sums in
another order), so the solves match the oracle under any schedule, the
autotuners pick a valid one, a batch under one schedule equals single solves
under it, and schedules that do not cover every (antenna, solution) are
rejected. The reference solves each antenna serially
(IterativeDiagonalSolver.cc:130-222)."""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from dp3_amd import warmstart as WS
from oracle import pyoracle as O
from problems import phase_ref_relerr, small_problem, solvertester_problem

pytestmark = pytest.mark.gpu

GAIN_TOL = 1e-4
# one schedule against another: FP64 sums in another order
ORDER_TOL = 1e-9


@pytest.fixture(scope="module")
def small():
    return small_problem()


@pytest.mark.parametrize("iteration", [H.ITERATION_STREAM, H.ITERATION_GRAM])
def test_schedules_match_oracle(small, iteration):
    w, geom, sky, gains, P = small
    print(f"case: antenna solve schedules, iteration {iteration}", flush=True)
    kw = dict(max_iterations=w.max_iterations, accuracy=w.tolerance)
    ref = P.unit_solutions()
    rr = O.solve(P, ref, **kw)
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 n_correlations=P.n_correlations, iteration=iteration, **kw)
    s.set_gram_fusion(False)
    base = P.unit_solutions()
    s.solve(P, base)
    default = s.antenna_solve_schedule()
    items = P.n_antennas * max(P.n_solutions_per_direction)
    for split in (1, 4, 16):
        need = max(64, -(-items * split // 64) * 64)
        for threads in (0, need, min(1024, 2 * need)):
            s.set_antenna_solve_schedule(split, threads)
            assert s.antenna_solve_schedule()[0] == split
            sol = P.unit_solutions()
            r = s.solve(P, sol)
            e = phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub)
            e_base = phase_ref_relerr(sol, base, P.n_antennas, P.n_sub)
            print(f"split {split} threads {threads}: {r.iterations} iterations "
                  f"(oracle {rr.iterations}), vs oracle {e:.2e}, vs default {e_base:.2e}")
            assert r.iterations == rr.iterations
            assert e <= GAIN_TOL
            assert e_base <= ORDER_TOL
    s.set_antenna_solve_schedule(0, 0)
    assert s.antenna_solve_schedule() == default
    s.close()


def test_autotune_picks_a_valid_schedule_for_several_solutions_per_direction():
    """SolverTester's dd(1, 2, 5) intervals: up to five solutions per
    direction, so a split multiplies 5 x n_ant items."""
    print("case: autotuned antenna solve schedule, dd(1,2,5)", flush=True)
    P, inp = solvertester_problem((1, 2, 5), n_corr=2)
    kw = dict(max_iterations=100, accuracy=1e-8, step_size=0.2)
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks, **kw)
    dp = H.DeviceProblem.from_problem(s, P)
    dp.upload(P)
    H.autotune(s, dp)
    split, threads = s.antenna_solve_schedule()
    items = P.n_antennas * max(P.n_solutions_per_direction)
    print(f"autotuned: split {split}, {threads} threads for {items} items")
    assert split in (1, 2, 4, 8, 16) and threads % 64 == 0 and items * split <= threads <= 1024
    sol = P.unit_solutions()
    r = s.solve_problem(dp, sol)
    ref = P.unit_solutions()
    rr = O.solve(P, ref, **kw)
    # accuracy 1e-8 lies below the noise floor of the reference's FP32
    # per-visibility sums on this noiseless fixture: the oracle stops at the
    # limit, the device (FP64 antenna sums) may stop earlier, never later
    print(f"iterations {r.iterations} (oracle {rr.iterations})")
    assert r.iterations <= rr.iterations
    assert phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub) <= GAIN_TOL
    s.set_antenna_solve_schedule(0, 0)
    base = P.unit_solutions()
    rb = s.solve_problem(dp, base)
    assert rb.iterations == r.iterations
    assert phase_ref_relerr(sol, base, P.n_antennas, P.n_sub) <= ORDER_TOL
    dp.close()
    s.close()


def test_batch_schedule_equals_single_and_rejects_uncovered_blocks():
    print("case: batched antenna solve schedule", flush=True)
    ws = [S.Workload(f"sb{i}", 10 + 4 * i, 2 + i, 1, 1, 8, 2, 2 + i, 40, 1e-5, seed=i)
          for i in range(3)]
    seqs = [WS.IntervalSequence(w, 1, seed=80 + i) for i, w in enumerate(ws)]
    b = H.Batch([q.solver for q in seqs], [q.problems[0] for q in seqs])
    b.autotune()
    split, threads = b.antenna_solve_schedule()
    items = max(w.n_stations * max(w.sols_per_dir()) for w in ws)
    print(f"batch autotuned: split {split}, {threads} threads, largest interval {items} items")
    assert items * split <= threads <= 1024 and threads % 64 == 0
    b.set_schedule(2, 1, 0, 8)
    b.set_antenna_solve_schedule(4, 0)
    got = [q.unit() for q in seqs]
    res = b.solve(got)
    for q, g, r in zip(seqs, got, res):
        q.solver.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=0, poll_interval=8)
        q.solver.set_antenna_solve_schedule(4, 0)
        x = q.unit()
        rs = q.solver.solve_problem(q.problems[0], x)
        assert rs.iterations == r.iterations
        np.testing.assert_array_equal(x, g)
    # 16 threads per item need more than one wave per 64 items: a 64-thread
    # block does not cover the largest interval and is refused; the batch
    # keeps its previous schedule
    with pytest.raises(H.HipError):
        b.set_antenna_solve_schedule(16, 64)
    assert b.antenna_solve_schedule()[0] == 4
    for bad in ((3, 0), (2, 100), (128, 0)):
        with pytest.raises(H.HipError):
            b.set_antenna_solve_schedule(*bad)
    again = [q.unit() for q in seqs]
    b.solve(again)
    for g, x in zip(got, again):
        np.testing.assert_array_equal(x, g)
    b.close()
