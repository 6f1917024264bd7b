"""The full-Jones solver on the GPU (n_solution_polarizations = 4,
IterativeFullJonesSolver.cc:18-224) against the oracle's full-Jones solve and
the reference's own case (tSolvers.cc:409-450: SolverTester dd-interval data,
the diagonals of the 2x2 solutions through CheckDualResults at 1e-2 %).
Tolerances as for the diagonal solver: 2x2 gains within 1e-4 relative
(Frobenius, per matrix) after removing antenna 0's J_00 phase, chi^2 within
0.1 %, equal iteration counts on noisy data (on noiseless fixtures the GPU may
stop earlier, never later)."""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from oracle import pyoracle as O
from problems import (check_dual_results, matrix_ref_relerr, small_problem, solvertester_problem,
                      unit_matrices)

pytestmark = pytest.mark.gpu

GAIN_TOL = 1e-4
CHI2_TOL = 1e-3


def gpu_full_jones(P, **kw):
    return H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                    n_correlations=4, n_solution_polarizations=4, **kw)


def units(P):
    return unit_matrices(P.n_channel_blocks, P.n_antennas, P.n_sub)


def compare(P, noiseless=False, oracle_kw=None, **kw):
    s = gpu_full_jones(P, **kw)
    sol = units(P)
    r = s.solve(P, sol)
    s.close()
    okw = dict(max_iterations=kw.get("max_iterations", 100), accuracy=kw.get("accuracy", 1e-5),
               n_threads=8)
    if "step_size" in kw:
        okw["step_size"] = kw["step_size"]
    if "detect_stalling" in kw:
        okw["detect_stalling"] = kw["detect_stalling"]
    okw.update(oracle_kw or {})
    ref = units(P)
    rr = O.solve(P, ref, full_jones=True, **okw)
    err = matrix_ref_relerr(sol, ref, P.n_antennas, P.n_sub)
    c_gpu, c_ref = O.chi2(P, sol, full_jones=True), O.chi2(P, ref, full_jones=True)
    print(f"iterations gpu {r.iterations} oracle {rr.iterations}, gain err {err:.2e}, "
          f"chi2 {c_gpu:.6g} / {c_ref:.6g}")
    if noiseless:
        assert r.iterations <= rr.iterations
    else:
        assert r.iterations == rr.iterations
        assert abs(c_gpu - c_ref) <= CHI2_TOL * c_ref
    assert err <= GAIN_TOL, err
    return sol, ref


def test_full_jones_solvertester_dd_intervals():
    print("case: full-Jones solvertester dd intervals", flush=True)
    P, inp = solvertester_problem((1, 2, 5), n_corr=4)
    sol, _ = compare(P, noiseless=True, max_iterations=100, accuracy=1e-8, step_size=0.2)
    m = sol.reshape(P.n_channel_blocks, -1, 4)
    diag = np.ascontiguousarray(m[:, :, [0, 3]].reshape(P.n_channel_blocks, -1))
    check_dual_results(diag, inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-2)


def test_full_jones_leakage_fixture():
    """Polarized models and 2x2 gains with off-diagonal leakage terms: the
    full 2x2 algebra (both antennas' sides, the Hermitian denominators)."""
    print("case: full-Jones leakage", flush=True)
    P, _ = solvertester_problem((1, 2, 5), n_corr=4, leakage=True)
    compare(P, noiseless=True, max_iterations=100, accuracy=1e-8, step_size=0.2)


@pytest.fixture(scope="module")
def small4():
    return small_problem(n_corr=4)


@pytest.mark.parametrize("storage", [H.MODEL_C64, H.MODEL_F16])
def test_full_jones_matches_oracle(small4, storage):
    w, geom, sky, gains, P = small4
    print(f"case: full-Jones small storage={storage}", flush=True)
    if storage == H.MODEL_C64:
        compare(P, max_iterations=w.max_iterations, accuracy=w.tolerance)
        return
    # FP16 model planes against the FP64 oracle at a fixed iteration count
    # (11 bits per model value relative to its direction's scale)
    s = gpu_full_jones(P, max_iterations=30, accuracy=0.0, detect_stalling=False,
                       model_storage=storage)
    sol = units(P)
    r = s.solve(P, sol)
    s.close()
    ref = units(P)
    rr = O.solve(P, ref, full_jones=True, max_iterations=30, accuracy=0.0, detect_stalling=False,
                 fp64=True, n_threads=8)
    err = matrix_ref_relerr(sol, ref, P.n_antennas, P.n_sub)
    c_gpu, c_ref = O.chi2(P, sol, full_jones=True), O.chi2(P, ref, full_jones=True)
    print(f"FP16: gain err {err:.2e}, chi2 {c_gpu:.6g} / {c_ref:.6g}")
    assert r.iterations == rr.iterations == 31
    assert err <= 2e-3
    assert abs(c_gpu - c_ref) <= CHI2_TOL * c_ref


def test_full_jones_direction_groups():
    """Ten directions: the correlate accumulates them in two groups of at most
    eight (each group forms the residual over all ten)."""
    import dataclasses
    w = dataclasses.replace(S.SMALL, name="24st-10dir", n_directions=10,
                            n_solutions_per_direction=(1, 2, 1, 4, 1, 2, 1, 1, 2, 1), seed=5)
    w_, geom, sky, gains, P = small_problem(w, n_corr=4)
    print("case: full-Jones 10 directions", flush=True)
    compare(P, max_iterations=w.max_iterations, accuracy=w.tolerance)


def test_full_jones_with_smoothness(small4):
    w, geom, sky, gains, P = small4
    print("case: full-Jones smoothness", flush=True)
    f = 120e6 + (np.arange(P.n_channel_blocks) + 0.5) * 30e6 / P.n_channel_blocks
    rng = np.random.default_rng(9)
    c = O.Smoothness(f, 12e6, 0.0, -1.0, True,
                     weights=rng.uniform(0.5, 2, (P.n_antennas, P.n_channel_blocks)),
                     antenna_factors=rng.uniform(0.6, 1.6, P.n_antennas))
    s = gpu_full_jones(P, max_iterations=400, accuracy=w.tolerance)
    s.set_smoothness(c.frequencies, c.bandwidth_hz, 0.0, -1.0, True, c.antenna_factors)
    s.set_constraint_weights(c.weights)
    sol = units(P)
    r = s.solve(P, sol)
    s.close()
    ref = units(P)
    rr = O.solve(P, ref, full_jones=True, max_iterations=400, accuracy=w.tolerance, smoothness=c,
                 n_threads=8)
    err = matrix_ref_relerr(sol, ref, P.n_antennas, P.n_sub)
    print(f"iterations gpu {r.iterations} oracle {rr.iterations}, gain err {err:.2e}")
    assert r.iterations == rr.iterations
    assert err <= GAIN_TOL


def test_full_jones_dead_antenna(small4):
    """An antenna without data: NaN solutions (singular denominators), the
    4-pol repair between iterations, other antennas as the oracle's."""
    w, geom, sky, gains, P0 = small4
    print("case: full-Jones dead antenna", flush=True)
    import copy
    P = copy.deepcopy(P0)
    dead = 3
    rows = (P.ant1 == dead) | (P.ant2 == dead)
    P.data = P.data.copy()
    P.data[rows] = 0
    model = P.model.copy()
    n_dir = P.n_directions
    for cb in range(P.n_channel_blocks):
        v0, v1 = P.vis_offsets[cb], P.vis_offsets[cb + 1]
        model[v0 * n_dir:v1 * n_dir].reshape(n_dir, v1 - v0, 4)[:, rows[v0:v1]] = 0
    P.model = model
    kw = dict(max_iterations=8, accuracy=0.0, detect_stalling=False)
    s = gpu_full_jones(P, **kw)
    sol = units(P)
    s.solve(P, sol)
    s.close()
    ref = units(P)
    O.solve(P, ref, full_jones=True, n_threads=8, **kw)
    m = sol.reshape(P.n_channel_blocks, P.n_antennas, P.n_sub, 4)
    mr = ref.reshape(P.n_channel_blocks, P.n_antennas, P.n_sub, 4)
    assert np.all(np.isnan(m[:, dead])) and np.all(np.isnan(mr[:, dead]))
    a = np.delete(m, dead, axis=1).reshape(P.n_channel_blocks, -1)
    b = np.delete(mr, dead, axis=1).reshape(P.n_channel_blocks, -1)
    err = matrix_ref_relerr(a, b, P.n_antennas - 1, P.n_sub)
    print(f"gain err {err:.2e}")
    assert err <= GAIN_TOL


def test_full_jones_chi2(small4):
    w, geom, sky, gains, P = small4
    s = gpu_full_jones(P, max_iterations=10, accuracy=0.0)
    sol = units(P)
    s.solve(P, sol)
    dp = H.DeviceProblem.from_problem(s, P)
    dp.upload(P)
    got = H.chi2(s, dp, sol)
    want = O.chi2(P, sol, full_jones=True)
    print(f"chi2 gpu {got:.10g} oracle {want:.10g}")
    assert abs(got - want) <= 1e-9 * want
    dp.close()
    s.close()


def test_full_jones_rejections(small4):
    w, geom, sky, gains, P = small4
    with pytest.raises(H.HipError, match="Full data"):
        H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 n_correlations=2, n_solution_polarizations=4)
    s = gpu_full_jones(P)
    dp = H.DeviceProblem.from_problem(s, P)
    dp.upload(P)
    diag = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                    n_correlations=4)
    dq = H.DeviceProblem.from_problem(diag, P)
    dq.upload(P)
    with pytest.raises(H.HipError, match="full-Jones or none"):
        H.Batch([s, diag], [dp, dq])
    for x in (dq, diag, dp, s):
        x.close()


@pytest.mark.parametrize("storage", [H.MODEL_C64, H.MODEL_F16])
def test_full_jones_batch_equals_single_solves(small4, storage):
    """Batched full-Jones solves (one launch per kernel per iteration over
    every interval, model products once per solve) equal single solves under
    the same schedule: the same iteration counts, and solutions to FP64
    rounding (the batched and single kernels are compiled separately, and
    their FP64 contractions may differ in the last bit)."""
    w, geom, sky, gains, P = small4
    print(f"case: full-Jones batch storage={storage}", flush=True)
    P2 = small_problem(n_corr=4, seed=11)[4]
    problems = [P, P2]
    kw = dict(max_iterations=w.max_iterations, accuracy=w.tolerance, model_storage=storage)
    singles = []
    for X in problems:
        s = gpu_full_jones(X, **kw)
        s.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=0)
        sol = units(X)
        r = s.solve(X, sol)
        singles.append((sol, r.iterations))
        s.close()
    solvers = [gpu_full_jones(X, **kw) for X in problems]
    dps = [H.DeviceProblem.from_problem(s, X) for s, X in zip(solvers, problems)]
    for dp, X in zip(dps, problems):
        dp.upload(X)
    b = H.Batch(solvers, dps)
    b.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=1, poll_interval=4)
    sols = [units(X) for X in problems]
    res = b.solve(sols)
    for (sol, it), got, r in zip(singles, sols, res):
        print(f"iterations batch {r.iterations} single {it}")
        assert r.iterations == it
        np.testing.assert_allclose(got, sol, rtol=1e-13, atol=1e-15)
    b.close()
    for x in dps + solvers:
        x.close()
