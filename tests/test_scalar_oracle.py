"""The oracle's scalar solver (IterativeScalarSolver) against the
reference's own cases: tSolvers.cc:155-200 (iterative_scalar_dd_intervals
with full-Jones data, iterative_duo_scalar_dd_intervals with Duo data) on the
SolverTester fixture with SetScalarSolutions, checked with CheckScalarResults
at 1e-3. CPU only."""
import time

import numpy as np
import pytest

from oracle import pyoracle as O
from problems import check_scalar_results, solvertester_problem


@pytest.mark.parametrize("n_corr", [2, 4])
def test_scalar_solvertester_dd_intervals(n_corr):
    P, inp = solvertester_problem((1, 2, 5), n_corr=n_corr, scalar=True)
    sol = np.ones((P.n_channel_blocks, P.n_antennas * P.n_sub), np.complex128)
    t0 = time.perf_counter()
    r = O.solve(P, sol, max_iterations=100, accuracy=1e-8, step_size=0.2, scalar=True,
                n_threads=8)
    print(f"scalar n_corr={n_corr}: {r.iterations} iterations, {time.perf_counter() - t0:.1f} s")
    assert r.iterations <= 101
    check_scalar_results(sol, inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-3)


def test_scalar_step_magnitude_and_repair():
    """One solution polarization: the step magnitude averages |Δg / g| over
    the solutions (SolverBase.cc:244-253), so a solve from the converged
    solutions stops at once; and a NaN start is repaired to the mean |g| of
    the finite solutions (SolverBase.cc:308-333), after which the solve still
    recovers the fixture's gains."""
    P, inp = solvertester_problem((1, 1, 1), n_corr=2, scalar=True)
    sol = np.ones((P.n_channel_blocks, P.n_antennas * P.n_sub), np.complex128)
    O.solve(P, sol, max_iterations=100, accuracy=1e-8, step_size=0.2, scalar=True, n_threads=8)
    again = sol.copy()
    r = O.solve(P, again, max_iterations=100, accuracy=1e-6, step_size=0.2, scalar=True,
                n_threads=8)
    assert r.iterations == 1 and r.has_converged
    assert r.step_magnitudes[0] <= 1e-6
    start = np.ones_like(sol)
    start[:, 7] = complex(np.nan, 0.0)
    O.solve(P, start, max_iterations=100, accuracy=1e-8, step_size=0.2, scalar=True,
            n_threads=8)
    check_scalar_results(start, inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-3)
