"""The committed golden fixtures (tests/golden/, written by make_golden.py
with the CPU oracle) against the oracle as built now: the oracle still
produces them, and the fixtures satisfy the properties that pin them."""
import numpy as np
import pytest

from oracle import pyoracle as O
from problems import (check_dual_results, golden, golden_components, golden_problem,
                      phase_ref_relerr, solvertester_problem)


@pytest.mark.parametrize("name", ["golden_small_duo.npz", "golden_small_full.npz"])
def test_oracle_reproduces_small_fixture(name):
    P, g = golden_problem(name)
    sol = P.unit_solutions()
    r = O.solve(P, sol, max_iterations=int(g["max_iterations"]), accuracy=float(g["accuracy"]),
                step_size=float(g["step_size"]), n_threads=1)
    assert r.iterations == int(g["iterations"])
    np.testing.assert_allclose(sol, g["final"], rtol=1e-9, atol=0)
    assert abs(O.chi2(P, sol) - float(g["chi2"])) <= 1e-9 * float(g["chi2"])
    # the reference precision mix and the FP64 restatement agree
    assert phase_ref_relerr(g["final"], g["final_fp64"], P.n_antennas, P.n_sub) <= 1e-4
    assert abs(float(g["chi2"]) - float(g["chi2_fp64"])) <= 1e-3 * float(g["chi2_fp64"])


def test_small_fixture_is_a_fixed_point_near_the_truth():
    """The converged fixture (tolerance 1e-5 on the damped step, so about
    5e-5 on the undamped one) is stationary under one undamped iteration,
    and its chi^2 lies below that of the true gains (the noise is fitted)
    and far below that of the unit start."""
    P, g = golden_problem("golden_small_duo.npz")
    nxt = O.iterate_once(P, g["final_fp64"], fp64=True)
    assert phase_ref_relerr(nxt, g["final_fp64"], P.n_antennas, P.n_sub) <= 2e-4
    chi_truth = O.chi2(P, g["true_gains"].astype(np.complex128))
    assert float(g["chi2_fp64"]) < chi_truth
    assert chi_truth < 1e-3 * O.chi2(P, P.unit_solutions())


def test_oracle_reproduces_solvertester_fixture():
    g = golden("golden_solvertester.npz")
    P, inp = solvertester_problem(tuple(int(x) for x in g["n_solutions_per_direction"]))
    np.testing.assert_array_equal(inp, g["input_solutions"])
    sol = P.unit_solutions()
    r = O.solve(P, sol, max_iterations=int(g["max_iterations"]), accuracy=float(g["accuracy"]),
                step_size=float(g["step_size"]), n_threads=1)
    assert r.iterations == int(g["iterations"])
    np.testing.assert_allclose(sol, g["final"], rtol=1e-9, atol=0)
    check_dual_results(g["final"], inp, P.n_antennas, P.n_sub, P.n_channel_blocks, 1e-2)


def test_oracle_reproduces_predict_fixture():
    g = golden("golden_predict.npz")
    comps = golden_components(g)
    for t in range(g["uvw"].shape[0]):
        v = O.simulate(tuple(g["ref"]), g["uvw"][t], g["baselines"], g["freq"], g["chan_width"],
                       comps, stokes_i_only=False, correct_freq_smearing=True)
        np.testing.assert_allclose(v, g["visibilities"][t], rtol=0, atol=1e-12)
