"""The scalar solver on the GPU (n_solution_polarizations = 1,
IterativeScalarSolver.cc:30-226) against the oracle's scalar solve and the
reference's own cases (tSolvers.cc:155-200: the SolverTester fixture with
SetScalarSolutions, CheckScalarResults at 1e-3, Duo and full-Jones data).
Tolerances as for the diagonal solver: gains within 1e-4 relative after
referencing to antenna 0, chi^2 within 0.1 %, equal iteration counts on noisy
data (on the noiseless fixture the GPU may stop earlier, never later)."""
import numpy as np
import pytest

from dp3_amd import hip as H
from oracle import pyoracle as O
from problems import check_scalar_results, small_problem, solvertester_problem

pytestmark = pytest.mark.gpu

GAIN_TOL = 1e-4
CHI2_TOL = 1e-3


def scalar_units(P):
    return np.ones((P.n_channel_blocks, P.n_antennas * P.n_sub), np.complex128)


def scalar_relerr(a, b, n_ant, n_sub):
    """Largest |a - b| / |b| after referencing both to antenna 0 per (block,
    sub-solution)."""
    a = a.reshape(a.shape[0], n_ant, n_sub)
    b = b.reshape(b.shape[0], n_ant, n_sub)
    ra = a * np.conj(a[:, :1]) / np.abs(a[:, :1])
    rb = b * np.conj(b[:, :1]) / np.abs(b[:, :1])
    return float(np.max(np.abs(ra - rb) / np.abs(rb)))


def as_diagonal(sol):
    """Scalar solutions as equal-polarization pairs (for chi^2)."""
    return np.ascontiguousarray(np.repeat(sol[..., None], 2, axis=-1).reshape(sol.shape[0], -1))


def gpu_scalar(P, **kw):
    return H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                    n_correlations=P.n_correlations, n_solution_polarizations=1, **kw)


@pytest.mark.parametrize("n_corr", [1, 2, 4])
def test_scalar_solvertester_dd_intervals(n_corr):
    """n_corr 1: single-correlation data 0.5 (XX + YY), the reference's
    iterative_uni_scalar_dd_intervals (tSolvers.cc:171-185); kept on the
    device as Duo pairs of equal values."""
    print(f"case: scalar solvertester n_corr={n_corr}", flush=True)
    P, inp = solvertester_problem((1, 2, 5), n_corr=n_corr, scalar=True)
    s = gpu_scalar(P, max_iterations=100, accuracy=1e-8, step_size=0.2)
    sol = scalar_units(P)
    r = s.solve(P, sol)
    check_scalar_results(sol, inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-3)
    ref = scalar_units(P)
    rr = O.solve(P, ref, max_iterations=100, accuracy=1e-8, step_size=0.2, scalar=True,
                 n_threads=8)
    err = scalar_relerr(sol, ref, P.n_antennas, P.n_sub)
    print(f"iterations gpu {r.iterations} oracle {rr.iterations}, gain err {err:.2e}")
    assert r.iterations <= rr.iterations
    assert err <= GAIN_TOL
    s.close()


@pytest.fixture(scope="module")
def small():
    return small_problem()


@pytest.mark.parametrize("iteration", [H.ITERATION_STREAM, H.ITERATION_GRAM])
def test_scalar_matches_oracle(small, iteration):
    """A scalar solve of data made with diagonal gains (one gain per antenna
    fitted to both polarizations)."""
    w, geom, sky, gains, P = small
    print(f"case: scalar small iteration={iteration}", flush=True)
    s = gpu_scalar(P, max_iterations=w.max_iterations, accuracy=w.tolerance, iteration=iteration)
    sol = scalar_units(P)
    r = s.solve(P, sol)
    ref = scalar_units(P)
    rr = O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance, scalar=True)
    err = scalar_relerr(sol, ref, P.n_antennas, P.n_sub)
    c_gpu, c_ref = O.chi2(P, as_diagonal(sol)), O.chi2(P, as_diagonal(ref))
    print(f"iterations gpu {r.iterations} oracle {rr.iterations}, gain err {err:.2e}, "
          f"chi2 {c_gpu:.6g} / {c_ref:.6g}")
    assert r.iterations == rr.iterations
    assert err <= GAIN_TOL
    assert abs(c_gpu - c_ref) <= CHI2_TOL * c_ref
    s.close()


def test_scalar_with_smoothness(small):
    w, geom, sky, gains, P = small
    print("case: scalar smoothness", flush=True)
    f = 120e6 + (np.arange(P.n_channel_blocks) + 0.5) * 30e6 / P.n_channel_blocks
    rng = np.random.default_rng(8)
    c = O.Smoothness(f, 12e6, 0.0, -1.0, True, weights=rng.uniform(0.5, 2, (P.n_antennas,
                                                                            P.n_channel_blocks)),
                     antenna_factors=rng.uniform(0.6, 1.6, P.n_antennas))
    s = gpu_scalar(P, max_iterations=400, accuracy=w.tolerance)
    s.set_smoothness(c.frequencies, c.bandwidth_hz, 0.0, -1.0, True, c.antenna_factors)
    s.set_constraint_weights(c.weights)
    sol = scalar_units(P)
    r = s.solve(P, sol)
    ref = scalar_units(P)
    rr = O.solve(P, ref, max_iterations=400, accuracy=w.tolerance, scalar=True, smoothness=c)
    err = scalar_relerr(sol, ref, P.n_antennas, P.n_sub)
    print(f"iterations gpu {r.iterations} oracle {rr.iterations}, gain err {err:.2e}")
    assert r.iterations == rr.iterations
    assert err <= GAIN_TOL
    s.close()


def test_scalar_batch_equals_single_solves(small):
    w, geom, sky, gains, P = small
    print("case: scalar batch", flush=True)
    s = gpu_scalar(P, max_iterations=w.max_iterations, accuracy=w.tolerance)
    s.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=0)
    single = scalar_units(P)
    r1 = s.solve(P, single)
    s.close()
    solvers = [gpu_scalar(P, max_iterations=w.max_iterations, accuracy=w.tolerance)
               for _ in range(2)]
    problems = [H.DeviceProblem.from_problem(x, P) for x in solvers]
    for dp in problems:
        dp.upload(P)
    b = H.Batch(solvers, problems)
    b.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=1, poll_interval=4)
    sols = [scalar_units(P) for _ in solvers]
    res = b.solve(sols)
    for got, r in zip(sols, res):
        assert r.iterations == r1.iterations
        np.testing.assert_array_equal(got, single)
    b.close()
    for dp in problems:
        dp.close()
    for x in solvers:
        x.close()


def test_scalar_rejects_bad_polarizations(small):
    w, geom, sky, gains, P = small
    with pytest.raises(H.HipError, match="n_solution_polarizations"):
        H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 n_solution_polarizations=3)


def test_scalar_single_correlation_chi2_and_device_problem():
    """A single-correlation problem uploaded to the device: the solve equals
    the drop-in solve, and chi^2 counts each visibility once (as the oracle's
    one-correlation chi^2)."""
    P, _ = solvertester_problem((1, 2, 5), n_corr=1, scalar=True)
    print("case: scalar single correlation device problem", flush=True)
    s = gpu_scalar(P, max_iterations=20, accuracy=0.0, detect_stalling=False)
    a = scalar_units(P)
    s.solve(P, a)
    dp = H.DeviceProblem.from_problem(s, P)
    dp.upload(P)
    b = scalar_units(P)
    s.solve_problem(dp, b)
    np.testing.assert_array_equal(a, b)
    got = H.chi2(s, dp, b)
    want = O.chi2(P, as_diagonal(b))
    print(f"chi2 gpu {got:.10g} oracle {want:.10g}")
    assert abs(got - want) <= 1e-9 * want
    with pytest.raises(H.HipError, match="scalar"):
        H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 n_correlations=1, n_solution_polarizations=2)
    dp.close()
    s.close()


def test_scalar_phase_only(small):
    """The scalarphase mode: the scalar solver with SetPhaseOnly(true)
    (SolverFactory.cc:297-303), a unit-modulus step along the circle
    (SolverBase.cc:128-155)."""
    w, geom, sky, gains, P = small
    print("case: scalar phase only", flush=True)
    s = gpu_scalar(P, max_iterations=30, accuracy=w.tolerance, phase_only=True)
    sol = scalar_units(P)
    r = s.solve(P, sol)
    s.close()
    ref = scalar_units(P)
    rr = O.solve(P, ref, max_iterations=30, accuracy=w.tolerance, phase_only=True, scalar=True)
    err = scalar_relerr(sol, ref, P.n_antennas, P.n_sub)
    print(f"iterations gpu {r.iterations} oracle {rr.iterations}, gain err {err:.2e}")
    assert r.iterations == rr.iterations
    np.testing.assert_allclose(np.abs(sol), 1.0, atol=1e-12)
    assert err <= GAIN_TOL


def test_scalar_single_correlation_gram_iteration():
    """Single-correlation data through the Gram iteration (its device planes
    are Duo pairs): the same iteration count and solutions as the streaming
    iteration's reference parity (tSolvers.cc:171-185)."""
    P, inp = solvertester_problem((1, 2, 5), n_corr=1, scalar=True)
    print("case: scalar single correlation gram", flush=True)
    s = gpu_scalar(P, max_iterations=100, accuracy=1e-8, step_size=0.2,
                   iteration=H.ITERATION_GRAM)
    sol = scalar_units(P)
    r = s.solve(P, sol)
    s.close()
    check_scalar_results(sol, inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-3)
    ref = scalar_units(P)
    rr = O.solve(P, ref, max_iterations=100, accuracy=1e-8, step_size=0.2, scalar=True,
                 n_threads=8)
    err = scalar_relerr(sol, ref, P.n_antennas, P.n_sub)
    print(f"iterations gpu {r.iterations} oracle {rr.iterations}, gain err {err:.2e}")
    assert r.iterations <= rr.iterations
    assert err <= GAIN_TOL
