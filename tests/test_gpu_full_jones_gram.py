"""The full-Jones solver's Gram iteration on the GPU: per-baseline cross
products of the data and model planes made once per solve (C_d, G_e,d),
turned into the correlate's X and Y sums by the gains every iteration
(IterativeFullJonesSolver.cc:99-188 restated on the cross products). Against
the oracle's full-Jones solve at the tolerances of the streaming iteration
(2x2 gains within 1e-4 relative after removing antenna 0's J_00 phase, chi^2
within 0.1 %, equal iteration counts on noisy data), and against the
streaming iteration on the same problem (the same iteration, FP32 sums in
another order)."""
import numpy as np
import pytest

from dp3_amd import hip as H
from oracle import pyoracle as O
from problems import check_dual_results, matrix_ref_relerr, small_problem, solvertester_problem
from test_gpu_full_jones import GAIN_TOL, compare, gpu_full_jones, units

pytestmark = pytest.mark.gpu

# Gram against streaming: the same iteration with FP32 sums in another order
ORDER_TOL = 1e-5


@pytest.fixture(scope="module")
def small4():
    return small_problem(n_corr=4)


def test_fj_gram_solvertester_dd_intervals_and_leakage():
    """The reference's SolverTester dd-interval case (three directions with
    1, 2 and 5 sub-solutions: several time segments per block) and the
    leakage fixture (off-diagonal gain terms, polarized models)."""
    print("case: full-Jones Gram solvertester", flush=True)
    P, inp = solvertester_problem((1, 2, 5), n_corr=4)
    sol, _ = compare(P, noiseless=True, max_iterations=100, accuracy=1e-8, step_size=0.2,
                     iteration=H.ITERATION_GRAM)
    m = sol.reshape(P.n_channel_blocks, -1, 4)
    diag = np.ascontiguousarray(m[:, :, [0, 3]].reshape(P.n_channel_blocks, -1))
    check_dual_results(diag, inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-2)
    print("case: full-Jones Gram leakage", flush=True)
    P, _ = solvertester_problem((1, 2, 5), n_corr=4, leakage=True)
    compare(P, noiseless=True, max_iterations=100, accuracy=1e-8, step_size=0.2,
            iteration=H.ITERATION_GRAM)


@pytest.mark.parametrize("storage", [H.MODEL_C64, H.MODEL_F16])
def test_fj_gram_matches_oracle_and_streaming(small4, storage):
    w, geom, sky, gains, P = small4
    print(f"case: full-Jones Gram small storage={storage}", flush=True)
    if storage == H.MODEL_C64:
        gram, _ = compare(P, max_iterations=w.max_iterations, accuracy=w.tolerance,
                          iteration=H.ITERATION_GRAM)
    else:
        s = gpu_full_jones(P, max_iterations=30, accuracy=0.0, detect_stalling=False,
                           model_storage=storage, iteration=H.ITERATION_GRAM)
        gram = units(P)
        r = s.solve(P, gram)
        s.close()
        ref = units(P)
        rr = O.solve(P, ref, full_jones=True, max_iterations=30, accuracy=0.0,
                     detect_stalling=False, fp64=True, n_threads=8)
        err = matrix_ref_relerr(gram, ref, P.n_antennas, P.n_sub)
        print(f"FP16 Gram: gain err {err:.2e} against the FP64 oracle")
        assert r.iterations == rr.iterations == 31
        assert err <= 2e-3
    kw = dict(max_iterations=w.max_iterations if storage == H.MODEL_C64 else 30,
              accuracy=w.tolerance if storage == H.MODEL_C64 else 0.0,
              detect_stalling=storage == H.MODEL_C64, model_storage=storage)
    s = gpu_full_jones(P, **kw)
    stream = units(P)
    s.solve(P, stream)
    s.close()
    e = matrix_ref_relerr(gram, stream, P.n_antennas, P.n_sub)
    print(f"Gram vs streaming: {e:.2e}")
    assert e <= ORDER_TOL


def test_fj_gram_with_smoothness(small4):
    w, geom, sky, gains, P = small4
    print("case: full-Jones Gram smoothness", flush=True)
    f = 120e6 + (np.arange(P.n_channel_blocks) + 0.5) * 30e6 / P.n_channel_blocks
    c = O.Smoothness(f, 12e6, 0.0, -1.0, True)
    s = gpu_full_jones(P, max_iterations=400, accuracy=w.tolerance, iteration=H.ITERATION_GRAM)
    s.set_smoothness(c.frequencies, c.bandwidth_hz, 0.0, -1.0, True)
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


@pytest.mark.parametrize("storage", [H.MODEL_C64, H.MODEL_F16])
def test_fj_gram_batch_equals_single_solves(small4, storage):
    """Batched full-Jones Gram solves (cross products once per solve for every
    interval, one apply launch per iteration) equal single Gram solves."""
    w, geom, sky, gains, P = small4
    print(f"case: full-Jones Gram batch storage={storage}", flush=True)
    P2 = small_problem(n_corr=4, seed=11)[4]
    problems = [P, P2]
    kw = dict(max_iterations=w.max_iterations, accuracy=w.tolerance, model_storage=storage,
              iteration=H.ITERATION_GRAM)
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


def test_fj_gram_rejects_more_than_eight_directions(small4):
    w, geom, sky, gains, P = small4
    with pytest.raises(H.HipError, match="at most 8 directions"):
        H.Solver(P.n_antennas, (1,) * 9, P.n_channel_blocks, n_correlations=4,
                 n_solution_polarizations=4, iteration=H.ITERATION_GRAM)
