"""Seeded test problems shared by the oracle and GPU tests.

Model visibilities here come from the CPU oracle's predict (oracle/), which
is test infrastructure; the GPU tests feed the same inputs to the HIP path.
"""
import os

import numpy as np

from dp3_amd import synth as S
from oracle import pyoracle as O


def oracle_components(sky, d):
    return [O.make_component(sky.ra[i], sky.dec[i], sky.flux[i],
                             gaussian=bool(sky.is_gaussian[i]), major=sky.major[i],
                             minor=sky.minor[i], pa=sky.pa[i])
            for i in np.where(sky.direction == d)[0]]


def oracle_model(w, geom, sky):
    """[n_dir, n_t, n_bl, n_ch] complex128 Stokes-I model visibilities."""
    n_t = geom.uvw.shape[0]
    out = np.zeros((w.n_directions, n_t, len(geom.baselines), len(geom.freq)), np.complex128)
    for d in range(w.n_directions):
        comps = oracle_components(sky, d)
        for t in range(n_t):
            out[d, t] = O.simulate(geom.ref, geom.uvw[t], geom.baselines, geom.freq,
                                   geom.chan_width, comps)[:, :, 0]
    return out


def small_problem(w=S.SMALL, n_corr=2, noise_sigma=None, seed=None):
    """Returns (workload, geometry, sky, true gains, SolveProblem)."""
    rng = np.random.default_rng(w.seed if seed is None else seed)
    geom = S.make_geometry(w, rng)
    sky = S.make_sky(w, geom, rng)
    model = oracle_model(w, geom, sky)
    gains = S.true_gains(w, rng)
    P = S.assemble(w, geom, model, gains, rng, noise_sigma=noise_sigma, n_corr=n_corr)
    return w, geom, sky, gains, P


def solvertester_problem(n_sol_per_dir=(1, 2, 5), n_corr=2, scalar=False, leakage=False):
    """SolverTester::FillDdIntervalData in SolveData layout + input solutions
    (SetScalarSolutions when `scalar`, SetLeakageSolutions with leakage-only
    data when `leakage`, else SetDiagonalSolutions)."""
    n_ant, n_dir, n_ch, n_t, n_cb = O.solvertester_shape()
    data, model, inp = O.solvertester_fill(n_sol_per_dir, scalar=scalar, leakage=leakage)
    a1 = [a for a in range(n_ant) for b in range(a + 1, n_ant)]
    a2 = [b for a in range(n_ant) for b in range(a + 1, n_ant)]
    pk = O.pack_solve_data(data, model, a1, a2, n_ant, n_cb, n_sol_per_dir, n_corr=n_corr)
    P = S.problem_from_solvertester(data, model, n_sol_per_dir, n_ant, n_cb, pk)
    P.channel_begin = S.channel_begin(n_ch, n_cb)
    return P, inp


def check_scalar_results(solutions, inp, n_ant, n_sub, n_cb, tolerance):
    """SolverTester::CheckScalarResults (SolverTester.cc:382-407): |g|^2 within
    `tolerance` percent, and the antenna-0-referenced product within
    `tolerance` absolute in real and imaginary part. solutions [n_cb,
    n_ant*n_sub] (one polarization), inp [n_ant*n_sub*2]."""
    sol = np.asarray(solutions).reshape(n_cb, n_ant, n_sub)
    x = np.asarray(inp).reshape(n_ant, n_sub, 2)[:, :, 0]
    for cb in range(n_cb):
        s0, i0 = sol[cb, 0], x[0]
        for ant in range(n_ant):
            s, i = sol[cb, ant], x[ant]
            a, b = np.abs(s) ** 2, np.abs(i) ** 2
            d = np.abs(a - b)
            assert np.all(d <= tolerance / 100 * a) and np.all(d <= tolerance / 100 * b), (cb, ant)
            rs, ri = s * np.conj(s0), i * np.conj(i0)
            assert np.all(np.abs(rs.real - ri.real) < tolerance), (cb, ant)
            assert np.all(np.abs(rs.imag - ri.imag) < tolerance), (cb, ant)


def check_dual_results(solutions, inp, n_ant, n_sub, n_cb, tolerance_percent):
    """SolverTester::CheckDualResults (SolverTester.cc:409-439): |g|^2 and the
    real part of g * conj(g_ant0) within `tolerance_percent` (BOOST_CHECK_CLOSE:
    the difference relative to each of the two values, in percent). Returns
    the worst relative error."""
    s = solutions.reshape(n_cb, n_ant, n_sub, 2)
    i = inp.reshape(n_ant, n_sub, 2)
    worst = 0.0
    for a, b in [(np.abs(s) ** 2, np.broadcast_to(np.abs(i) ** 2, s.shape)),
                 ((s * np.conj(s[:, 0:1])).real,
                  np.broadcast_to((i * np.conj(i[0:1])).real, s.shape))]:
        rel = np.abs(a - b) / np.minimum(np.abs(a), np.abs(b))
        worst = max(worst, float(rel.max()))
    assert worst * 100.0 <= tolerance_percent, f"worst relative error {worst:.3e}"
    return worst


def phase_ref_relerr(sol_a, sol_b, n_ant, n_sub):
    """Max |a_ref - b_ref| / |b_ref| after referencing both to antenna 0."""
    ra = S.phase_referenced(sol_a, n_ant, n_sub)
    rb = S.phase_referenced(sol_b, n_ant, n_sub)
    return float(np.max(np.abs(ra - rb) / np.abs(rb)))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    """Arrays of a committed fixture in tests/golden/ (numpy, no pickles)."""
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_problem(name):
    """(SolveProblem, fixture arrays) of golden_small_{duo,full}.npz."""
    g = golden(name)
    P = S.SolveProblem(int(g["n_antennas"]), int(g["n_directions"]), int(g["n_channel_blocks"]),
                       int(g["n_correlations"]), g["n_solutions_per_direction"].astype(np.int64),
                       g["vis_offsets"], g["data"], g["model"], g["ant1"], g["ant2"],
                       g["sol_map"], int(g["n_times"]), int(g["n_baselines"]),
                       g["channel_begin"])
    return P, g


def golden_components(g):
    """Oracle components of golden_predict.npz."""
    out = []
    for i in range(len(g["ra"])):
        terms = list(g["spectral_terms"][i][: int(g["n_spectral_terms"][i])])
        out.append(O.make_component(
            g["ra"][i], g["dec"][i], g["I"][i], g["Q"][i], g["U"][i], g["V"][i],
            gaussian=bool(g["gaussian"][i]), major=g["major"][i], minor=g["minor"][i],
            pa=g["pa"][i], pa_absolute=bool(g["pa_absolute"][i]), spectral_terms=terms or None,
            ref_freq=g["ref_freq"][i], log_si=bool(g["log_si"][i]),
            rm=tuple(g["rm"][i]) if g["has_rm"][i] else None))
    return out


def unit_matrices(n_cb, n_ant, n_sub):
    """Identity 2x2 solutions [n_cb, n_ant*n_sub*4] (the reference's full-Jones
    tests start from unit matrices, tSolvers.cc:374-381)."""
    s = np.zeros((n_cb, n_ant * n_sub, 4), np.complex128)
    s[:, :, 0] = 1.0
    s[:, :, 3] = 1.0
    return s.reshape(n_cb, -1)


def matrix_ref_relerr(sol_a, sol_b, n_ant, n_sub):
    """Largest |A - B| / |B| (Frobenius, per matrix) after removing the phase of
    antenna 0's J_00 per (block, sub-solution) from both: full-Jones
    solutions are unique up to one phase per (block, direction solution)."""
    def ref(s):
        m = np.asarray(s).reshape(s.shape[0], n_ant, n_sub, 4)
        p = m[:, :1, :, :1]
        return m * (np.conj(p) / np.abs(p))
    ra, rb = ref(sol_a), ref(sol_b)
    return float(np.max(np.linalg.norm(ra - rb, axis=-1) / np.linalg.norm(rb, axis=-1)))
