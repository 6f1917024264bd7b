"""GPU parity of the HIP iterative diagonal solver against the CPU oracle.

Tolerances: gains within 1e-4 relative after referencing to antenna 0, final
chi^2 within 0.1 %, identical iteration counts (complex64 model storage).
FP16 model storage is compared with the FP64 oracle at looser, stated bounds.
"""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from oracle import pyoracle as O
from problems import (check_dual_results, phase_ref_relerr, small_problem,
                      solvertester_problem)

pytestmark = pytest.mark.gpu

GAIN_TOL = 1e-4
CHI2_TOL = 1e-3


def gpu_solver(P, **kw):
    return H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                    n_correlations=P.n_correlations, **kw)


@pytest.fixture(scope="module")
def small():
    return small_problem()


def test_small_workload_matches_oracle(small):
    w, geom, sky, gains, P = small
    print(f"case: {w.name} duo c64", flush=True)
    ref = P.unit_solutions()
    r_ref = O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance)
    sol = P.unit_solutions()
    s = gpu_solver(P, max_iterations=w.max_iterations, accuracy=w.tolerance)
    r = s.solve(P, sol)
    assert r.iterations == r_ref.iterations
    err = phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub)
    assert err <= GAIN_TOL, err
    c_gpu, c_ref = O.chi2(P, sol), O.chi2(P, ref)
    assert abs(c_gpu - c_ref) <= CHI2_TOL * c_ref
    # error against the FP64 oracle, reported
    ref64 = P.unit_solutions()
    O.solve(P, ref64, max_iterations=w.max_iterations, accuracy=w.tolerance, fp64=True)
    print(f"gain err vs oracle {err:.2e}, vs fp64 oracle "
          f"{phase_ref_relerr(sol, ref64, P.n_antennas, P.n_sub):.2e}, chi2 {c_gpu:.6g} / {c_ref:.6g}")


@pytest.mark.parametrize("n_corr", [2, 4])
def test_solvertester_dd_intervals(n_corr):
    """tSolvers.cc:280-310 (iterative_[duo_]diagonal_dd_intervals): recovery of
    the fixture's gains with CheckDualResults(1e-2), and parity with the oracle."""
    print(f"case: solvertester dd(1,2,5) n_corr={n_corr}", flush=True)
    P, inp = solvertester_problem((1, 2, 5), n_corr=n_corr)
    sol = P.unit_solutions()
    s = gpu_solver(P, max_iterations=100, accuracy=1e-8, step_size=0.2)
    r = s.solve(P, sol)
    assert r.iterations <= 101
    check_dual_results(sol, inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-2)
    ref = P.unit_solutions()
    r_ref = O.solve(P, ref, max_iterations=100, accuracy=1e-8, step_size=0.2, n_threads=4)
    # Accuracy 1e-8 lies below the noise floor of the reference's per-visibility
    # FP32 numerator sums (its step magnitude levels off near 8e-8 on this
    # noiseless fixture, so it stops at max+1). The factored per-baseline sums
    # with an FP64 antenna reduction get below it, so the GPU may stop earlier,
    # never later.
    print(f"iterations gpu {r.iterations} oracle {r_ref.iterations}")
    assert r.iterations <= r_ref.iterations
    assert phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub) <= GAIN_TOL


def test_device_problem_matches_dropin(small):
    w, geom, sky, gains, P = small
    print(f"case: {w.name} device problem", flush=True)
    s = gpu_solver(P, max_iterations=w.max_iterations, accuracy=w.tolerance)
    a = P.unit_solutions()
    ra = s.solve(P, a)
    dp = H.DeviceProblem.from_problem(s, P)
    dp.upload(P)
    b = P.unit_solutions()
    rb = s.solve_problem(dp, b)
    assert ra.iterations == rb.iterations
    np.testing.assert_array_equal(a, b)
    # the device layout holds the data plane transposed to baseline-fastest
    plane = dp.download_plane(0, 0)
    n_ch = int(dp.channels_per_block[0])
    ref = P.data[: int(P.vis_offsets[1])].reshape(P.n_times, P.n_baselines, n_ch, 2)
    got = plane.reshape(P.n_times, n_ch, dp.bl_stride, 2)[:, :, : P.n_baselines]
    np.testing.assert_array_equal(got, ref.transpose(0, 2, 1, 3))
    assert np.all(plane.reshape(P.n_times, n_ch, dp.bl_stride, 2)[:, :, P.n_baselines:] == 0)


def test_iteration_count_semantics(small):
    """min_iterations honoured (tSolvers.cc:512-529); max+1 when unconverged
    (SolverBase.cc:197-200)."""
    w, geom, sky, gains, P = small
    print("case: iteration semantics", flush=True)
    s = gpu_solver(P, max_iterations=50, min_iterations=10, accuracy=1e8)
    sol = P.unit_solutions()
    assert s.solve(P, sol).iterations == 10
    s2 = gpu_solver(P, max_iterations=7, accuracy=1e-12)
    sol = P.unit_solutions()
    r = s2.solve(P, sol)
    assert r.iterations == 8 and not r.has_converged


def test_zero_denominator_and_nan_repair(small):
    """Antenna 3 fully flagged in block 1 (zero data and model, as
    AssignAndWeight leaves it): den == 0 gives NaN (IterativeDiagonalSolver.cc
    :212-214), repaired each iteration by MakeSolutionsFinite2Pol."""
    w, geom, sky, gains, P = small
    print("case: zero denominator", flush=True)
    import copy
    Q = copy.deepcopy(P)
    v0, v1 = int(Q.vis_offsets[1]), int(Q.vis_offsets[2])
    sel = (Q.ant1[v0:v1] == 3) | (Q.ant2[v0:v1] == 3)
    Q.data[v0:v1][sel] = 0
    m = Q.model[v0 * Q.n_directions:v1 * Q.n_directions].reshape(Q.n_directions, v1 - v0, 2)
    m[:, sel] = 0
    ref = Q.unit_solutions()
    r_ref = O.solve(Q, ref, max_iterations=20, accuracy=w.tolerance)
    sol = Q.unit_solutions()
    r = gpu_solver(Q, max_iterations=20, accuracy=w.tolerance).solve(Q, sol)
    assert r.iterations == r_ref.iterations
    # the flagged antenna's final solutions are NaN in both (not repaired after the last step)
    s = sol.reshape(Q.n_channel_blocks, Q.n_antennas, Q.n_sub, 2)
    e = ref.reshape(Q.n_channel_blocks, Q.n_antennas, Q.n_sub, 2)
    assert np.all(np.isnan(s[1, 3])) and np.all(np.isnan(e[1, 3]))
    ok = np.ones(s.shape[:2], bool)
    ok[1, 3] = False
    assert np.max(np.abs(s[ok] - e[ok]) / np.abs(e[ok])) <= GAIN_TOL


def test_phase_only(small):
    w, geom, sky, gains, P = small
    print("case: phase only", flush=True)
    ref = P.unit_solutions()
    r_ref = O.solve(P, ref, max_iterations=30, accuracy=w.tolerance, phase_only=True)
    sol = P.unit_solutions()
    r = gpu_solver(P, max_iterations=30, accuracy=w.tolerance, phase_only=True).solve(P, sol)
    assert r.iterations == r_ref.iterations
    np.testing.assert_allclose(np.abs(sol), 1.0, atol=1e-12)
    assert phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub) <= GAIN_TOL


def test_fp16_model_storage(small):
    """FP16 model planes with a per (block, direction) scale and FP32
    accumulation, against the FP64 oracle: gains within 2e-3 relative, chi^2
    within 0.1 %."""
    w, geom, sky, gains, P = small
    print("case: fp16 model", flush=True)
    ref64 = P.unit_solutions()
    O.solve(P, ref64, max_iterations=w.max_iterations, accuracy=w.tolerance, fp64=True)
    sol = P.unit_solutions()
    s = gpu_solver(P, max_iterations=w.max_iterations, accuracy=w.tolerance,
                   model_storage=H.MODEL_F16)
    s.solve(P, sol)
    err = phase_ref_relerr(sol, ref64, P.n_antennas, P.n_sub)
    c16, c64 = O.chi2(P, sol), O.chi2(P, ref64)
    print(f"fp16 model: gain err vs fp64 oracle {err:.2e}, chi2 rel {abs(c16 - c64) / c64:.2e}")
    assert err <= 2e-3
    assert abs(c16 - c64) <= CHI2_TOL * c64


@pytest.mark.parametrize("waves,rows,graph,poll", [(1, 1, 0, 5), (2, 2, 0, 3), (3, 1, 1, 4),
                                                    (4, 2, 1, 8), (4, 1, 1, 1)])
def test_schedule_does_not_change_results_beyond_rounding(small, waves, rows, graph, poll):
    w, geom, sky, gains, P = small
    print(f"case: waves={waves} rows={rows} graph={graph} poll={poll}", flush=True)
    s = gpu_solver(P, max_iterations=w.max_iterations, accuracy=w.tolerance)
    s.set_schedule(waves_per_tile=waves, rows_per_trip=rows, use_graph=graph, poll_interval=poll)
    sol = P.unit_solutions()
    r = s.solve(P, sol)
    ref = P.unit_solutions()
    rr = O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance)
    assert r.iterations == rr.iterations
    assert phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub) <= GAIN_TOL
    # a second solve on the same solver (graph reuse) gives the same result
    sol2 = P.unit_solutions()
    s.solve(P, sol2)
    np.testing.assert_array_equal(sol, sol2)


@pytest.mark.parametrize("storage", [H.MODEL_C64, H.MODEL_F16])
def test_gram_iteration_matches_oracle(small, storage):
    """The Gram iteration (cross products once per solve, applied to the gains
    each iteration) against the oracle: same tolerances as the streaming
    iteration (FP16 model: against the FP64 oracle)."""
    w, geom, sky, gains, P = small
    print(f"case: gram iteration storage={storage}", flush=True)
    sol = P.unit_solutions()
    r = gpu_solver(P, max_iterations=w.max_iterations, accuracy=w.tolerance,
                   model_storage=storage, iteration=H.ITERATION_GRAM).solve(P, sol)
    if storage == H.MODEL_C64:
        ref = P.unit_solutions()
        rr = O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance)
        assert r.iterations == rr.iterations
        tol = GAIN_TOL
    else:
        ref = P.unit_solutions()
        O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance, fp64=True)
        tol = 2e-3
    err = phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub)
    c, cr = O.chi2(P, sol), O.chi2(P, ref)
    print(f"gram: iterations {r.iterations}, gain err {err:.2e}, chi2 {c:.6g} / {cr:.6g}")
    assert err <= tol
    assert abs(c - cr) <= CHI2_TOL * cr
    # streaming and Gram iterations agree
    s2 = P.unit_solutions()
    gpu_solver(P, max_iterations=w.max_iterations, accuracy=w.tolerance,
               model_storage=storage).solve(P, s2)
    assert phase_ref_relerr(sol, s2, P.n_antennas, P.n_sub) <= tol


def test_gram_iteration_dd_intervals():
    """Direction-dependent solution intervals (segments with several solution
    indices) under the Gram iteration: SolverTester recovery and oracle parity."""
    print("case: gram iteration solvertester dd(1,2,5)", flush=True)
    P, inp = solvertester_problem((1, 2, 5), n_corr=2)
    sol = P.unit_solutions()
    s = gpu_solver(P, max_iterations=100, accuracy=1e-8, step_size=0.2,
                   iteration=H.ITERATION_GRAM)
    r = s.solve(P, sol)
    check_dual_results(sol, inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-2)
    ref = P.unit_solutions()
    r_ref = O.solve(P, ref, max_iterations=100, accuracy=1e-8, step_size=0.2, n_threads=4)
    print(f"iterations gram {r.iterations} oracle {r_ref.iterations}")
    assert r.iterations <= r_ref.iterations
    assert phase_ref_relerr(sol, ref, P.n_antennas, P.n_sub) <= GAIN_TOL


def test_gram_iteration_rejects_unsupported_shapes(small):
    w, geom, sky, gains, P = small
    with pytest.raises(H.HipError):
        H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 n_correlations=4, iteration=H.ITERATION_GRAM)
    with pytest.raises(H.HipError):
        H.Solver(P.n_antennas, [1] * 9, P.n_channel_blocks, iteration=H.ITERATION_GRAM)
