"""The oracle's full-Jones solver (IterativeFullJonesSolver.cc:18-224) against
the reference's own cases and against what the mathematics fixes. CPU only.

* iterative_full_jones_dd_intervals (tSolvers.cc:409-450): SolverTester data
  made with diagonal gains and direction-dependent solution intervals; the
  diagonals of the 2x2 solutions pass CheckDualResults at 1e-2 %.
* Leakage-only data (SetLeakageSolutions, FillDdIntervalData(true)): the
  solver recovers J = [[1, l_0], [l_1, 1]] up to the one phase per
  (block, solution) that the data cannot fix.
* Data and models with zero off-diagonals: every iterate stays diagonal and
  equals the diagonal solver's iterate, and the step magnitude of the 4-pol
  convergence rule is half the 2-pol one (the same |Δ/g| terms over four
  entries instead of two, SolverBase.cc:254-283).
"""
import numpy as np

from oracle import pyoracle as O
from problems import (check_dual_results, matrix_ref_relerr, small_problem, solvertester_problem,
                      unit_matrices)

SETTINGS = dict(max_iterations=100, accuracy=1e-8, step_size=0.2, n_threads=8)


def diagonals(sol, n_cb):
    m = sol.reshape(n_cb, -1, 4)
    return np.ascontiguousarray(m[:, :, [0, 3]].reshape(n_cb, -1))


def test_iterative_full_jones_dd_intervals():
    P, inp = solvertester_problem((1, 2, 5), n_corr=4)
    sol = unit_matrices(P.n_channel_blocks, P.n_antennas, P.n_sub)
    r = O.solve(P, sol, full_jones=True, **SETTINGS)
    assert r.iterations <= SETTINGS["max_iterations"] + 1
    check_dual_results(diagonals(sol, P.n_channel_blocks), inp, P.n_antennas, P.n_sub,
                       P.n_channel_blocks, 1e-2)
    # the data has no leakage: the off-diagonal terms end near 0 (at the level
    # the dual check allows, 1e-4 relative)
    m = sol.reshape(P.n_channel_blocks, -1, 4)
    assert np.max(np.abs(m[:, :, 1:3])) < 1e-3 * np.max(np.abs(m[:, :, 0]))


def test_leakage_recovery():
    P, inp = solvertester_problem((1, 2, 5), n_corr=4, leakage=True)
    sol = unit_matrices(P.n_channel_blocks, P.n_antennas, P.n_sub)
    O.solve(P, sol, full_jones=True, **SETTINGS)
    leak = inp.reshape(P.n_antennas, P.n_sub, 2)
    truth = np.zeros((P.n_channel_blocks, P.n_antennas, P.n_sub, 4), np.complex128)
    truth[..., 0] = 1.0
    truth[..., 3] = 1.0
    truth[..., 1] = leak[..., 0]
    truth[..., 2] = leak[..., 1]
    err = matrix_ref_relerr(sol, truth.reshape(P.n_channel_blocks, -1), P.n_antennas, P.n_sub)
    # 100 iterations leave 5e-4 (the step magnitude is still 8e-7 and falling:
    # a slow mode, 3e-4 after 300); a transposed or conjugated operand gives
    # errors of order the leakage (0.1-0.5)
    assert err < 1e-3, err
    # the noiseless model is reproduced
    assert O.chi2(P, sol, full_jones=True) < 1e-9 * np.sum(np.abs(P.data) ** 2)


def test_diagonal_data_follows_the_diagonal_solver():
    w, geom, sky, gains, P = small_problem(n_corr=4)
    n_it = 12
    kw = dict(max_iterations=n_it, accuracy=0.0, detect_stalling=False, n_threads=8)
    full = unit_matrices(P.n_channel_blocks, P.n_antennas, P.n_sub)
    rf = O.solve(P, full, full_jones=True, **kw)
    diag = P.unit_solutions()
    rd = O.solve(P, diag, **kw)
    assert rf.iterations == rd.iterations == n_it + 1
    m = full.reshape(P.n_channel_blocks, -1, 4)
    assert np.all(m[:, :, 1:3] == 0)
    d = diagonals(full, P.n_channel_blocks)
    np.testing.assert_allclose(d, diag, rtol=1e-5, atol=0)
    np.testing.assert_allclose(rf.step_magnitudes, rd.step_magnitudes / 2, rtol=1e-4)
    np.testing.assert_allclose(O.chi2(P, full, full_jones=True), O.chi2(P, diag), rtol=1e-6)


def test_singular_denominator_gives_nan_then_repair():
    """An antenna with no data has a zero denominator: its solution becomes NaN
    (IterativeFullJonesSolver.cc:178-185), is repaired to the block's mean
    |J_00| in all four entries before the next iteration (SolverBase.cc:
    373-405), and ends as NaN again; the other antennas are unaffected."""
    w, geom, sky, gains, P = small_problem(n_corr=4)
    dead = 3
    rows = (P.ant1 == dead) | (P.ant2 == dead)
    P.data[rows] = 0
    n_dir = P.n_directions
    model = P.model.copy()
    for cb in range(P.n_channel_blocks):
        v0, v1 = P.vis_offsets[cb], P.vis_offsets[cb + 1]
        blk = model[v0 * n_dir:v1 * n_dir].reshape(n_dir, v1 - v0, 4)
        blk[:, rows[v0:v1]] = 0
    P.model = model
    sol = unit_matrices(P.n_channel_blocks, P.n_antennas, P.n_sub)
    O.solve(P, sol, full_jones=True, max_iterations=5, accuracy=0.0, detect_stalling=False)
    m = sol.reshape(P.n_channel_blocks, P.n_antennas, P.n_sub, 4)
    assert np.all(np.isnan(m[:, dead]))
    others = np.delete(m, dead, axis=1)
    assert np.all(np.isfinite(others))
