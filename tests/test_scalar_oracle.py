"""The oracle's scalar solver (IterativeScalarSolver) against the
reference's own cases: tSolvers.cc:155-200 (iterative_scalar_dd_intervals
with full-Jones data, iterative_uni_scalar_dd_intervals with single-
correlation data, iterative_duo_scalar_dd_intervals with Duo data) on the
SolverTester fixture with SetScalarSolutions, checked with CheckScalarResults
at 1e-3. CPU only."""
import time

import numpy as np
import pytest

from oracle import pyoracle as O
from problems import check_scalar_results, solvertester_problem


@pytest.mark.parametrize("n_corr", [1, 2, 4])
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


def test_single_correlation_packing_and_duo_equivalence():
    """Single-correlation data is 0.5 (XX + YY) of data and model
    (ToSpecificMatrix, SolveData.cc:29-31); a scalar solve of it equals, bit
    for bit, the scalar solve of Duo data holding that value in both
    correlations (both per-correlation terms are equal, so the float sums
    double exactly and the ratio is unchanged)."""
    P1, _ = solvertester_problem((1, 2, 5), n_corr=1, scalar=True)
    P4, _ = solvertester_problem((1, 2, 5), n_corr=4, scalar=True)
    want = (0.5 * (P4.data[:, 0].astype(np.complex64) + P4.data[:, 3])).astype(np.complex64)
    np.testing.assert_array_equal(P1.data[:, 0], want)
    wm = (0.5 * (P4.model[:, 0].astype(np.complex64) + P4.model[:, 3])).astype(np.complex64)
    np.testing.assert_array_equal(P1.model[:, 0], wm)
    import copy
    P2 = copy.copy(P1)
    P2.n_correlations = 2
    P2.data = np.ascontiguousarray(np.repeat(P1.data, 2, axis=1))
    P2.model = np.ascontiguousarray(np.repeat(P1.model, 2, axis=1))
    kw = dict(max_iterations=40, accuracy=1e-8, step_size=0.2, scalar=True, n_threads=8)
    a = np.ones((P1.n_channel_blocks, P1.n_antennas * P1.n_sub), np.complex128)
    b = a.copy()
    ra = O.solve(P1, a, **kw)
    rb = O.solve(P2, b, **kw)
    assert ra.iterations == rb.iterations
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ra.step_magnitudes, rb.step_magnitudes)
