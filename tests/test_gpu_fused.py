"""The fused Gram iteration (apply and antenna solve in one kernel per channel
block, baselines visited in a ring walk of the antennas) against the oracle
and against the unfused apply + antenna solve kernels.

The fused kernel forms each baseline's two SolveDirection terms in FP64 from
its FP32 U and sums them per antenna in FP64 over the walk steps, its waves
and its row groups; the unfused antenna solve adds per-baseline FP32 partials
in FP64. Both follow the reference's rules (IterativeDiagonalSolver.cc:
130-222), so they agree with the oracle to the parity bars and with each
other to summation order.
Baseline orientation and order are free in SolveData: reversed baselines
(data and model conjugated) and reordered or missing ones must give the same
solutions."""
import dataclasses

import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from dp3_amd import warmstart as WS
from oracle import pyoracle as O
from problems import phase_ref_relerr, small_problem

pytestmark = pytest.mark.gpu

GAIN_TOL = 1e-4
CHI2_TOL = 1e-3
# fused against unfused: the same iteration with a different summation order
ORDER_TOL = 1e-5


def solver(P, fused=True, **kw):
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 n_correlations=P.n_correlations, iteration=H.ITERATION_GRAM, **kw)
    s.set_gram_fusion(fused)
    return s


def variant(P, reverse=(), drop=()):
    """P with the baselines in `reverse` turned around (antennas swapped, data
    and models conjugated) and those in `drop` removed (SolveData order:
    time -> baseline -> channel per block)."""
    n_bl, n_t, n_dir = P.n_baselines, P.n_times, P.n_directions
    keep = np.array([b not in set(drop) for b in range(n_bl)])
    rev = np.array([b in set(reverse) for b in range(n_bl)])
    datas, models, a1s, a2s, sms, offs = [], [], [], [], [], [0]
    for cb in range(P.n_channel_blocks):
        v0, v1 = int(P.vis_offsets[cb]), int(P.vis_offsets[cb + 1])
        nc = (v1 - v0) // (n_t * n_bl)
        shape = (n_t, n_bl, nc)
        d = P.data[v0:v1].reshape(shape + (-1,)).copy()
        m = P.model[v0 * n_dir:v1 * n_dir].reshape((n_dir,) + shape + (-1,)).copy()
        a1 = P.ant1[v0:v1].reshape(shape).copy()
        a2 = P.ant2[v0:v1].reshape(shape).copy()
        sm = P.sol_map[v0 * n_dir:v1 * n_dir].reshape((n_dir,) + shape)
        d[:, rev] = np.conj(d[:, rev])
        m[:, :, rev] = np.conj(m[:, :, rev])
        a1[:, rev], a2[:, rev] = a2[:, rev].copy(), a1[:, rev].copy()
        datas.append(d[:, keep].reshape(-1, d.shape[-1]))
        models.append(m[:, :, keep].reshape(-1, m.shape[-1]))
        a1s.append(a1[:, keep].ravel())
        a2s.append(a2[:, keep].ravel())
        sms.append(np.ascontiguousarray(sm[:, :, keep]).ravel())
        offs.append(offs[-1] + n_t * int(keep.sum()) * nc)
    return S.SolveProblem(P.n_antennas, n_dir, P.n_channel_blocks, P.n_correlations,
                          P.n_solutions_per_direction, np.array(offs, np.int64),
                          np.concatenate(datas), np.concatenate(models), np.concatenate(a1s),
                          np.concatenate(a2s), np.concatenate(sms), n_t, int(keep.sum()),
                          P.channel_begin)


@pytest.fixture(scope="module")
def small():
    return small_problem()


@pytest.mark.parametrize("storage", [H.MODEL_C64, H.MODEL_F16])
def test_fused_matches_oracle_and_unfused(small, storage):
    w, geom, sky, gains, P = small
    print(f"case: fused gram storage={storage}", flush=True)
    kw = dict(max_iterations=w.max_iterations, accuracy=w.tolerance, model_storage=storage)
    fused = P.unit_solutions()
    rf = solver(P, True, **kw).solve(P, fused)
    plain = P.unit_solutions()
    ru = solver(P, False, **kw).solve(P, plain)
    ref = P.unit_solutions()
    rr = O.solve(P, ref, max_iterations=w.max_iterations, accuracy=w.tolerance,
                 fp64=storage == H.MODEL_F16)
    e_ref = phase_ref_relerr(fused, ref, P.n_antennas, P.n_sub)
    e_un = phase_ref_relerr(fused, plain, P.n_antennas, P.n_sub)
    print(f"iterations fused {rf.iterations} unfused {ru.iterations} oracle {rr.iterations}; "
          f"fused vs oracle {e_ref:.2e}, vs unfused {e_un:.2e}")
    assert rf.iterations == ru.iterations
    assert e_un <= ORDER_TOL
    if storage == H.MODEL_C64:
        assert rf.iterations == rr.iterations
        assert e_ref <= GAIN_TOL
    else:
        assert e_ref <= 2e-3
    assert abs(O.chi2(P, fused) - O.chi2(P, ref)) <= CHI2_TOL * O.chi2(P, ref)


def test_fused_reversed_reordered_and_missing_baselines(small):
    """Reversing a third of the baselines (conjugated data and models) leaves
    the solution unchanged; dropping some baselines leaves pairs without a
    baseline in the walk, and the oracle on the same data agrees."""
    w, geom, sky, gains, P = small
    print("case: fused gram reversed / missing baselines", flush=True)
    rng = np.random.default_rng(3)
    rev = rng.choice(P.n_baselines, P.n_baselines // 3, replace=False)
    Q = variant(P, reverse=rev)
    kw = dict(max_iterations=w.max_iterations, accuracy=w.tolerance)
    a = P.unit_solutions()
    solver(P, **kw).solve(P, a)
    b = Q.unit_solutions()
    solver(Q, **kw).solve(Q, b)
    e = phase_ref_relerr(b, a, P.n_antennas, P.n_sub)
    print(f"reversed baselines: solutions differ by {e:.2e}")
    assert e <= ORDER_TOL
    R = variant(P, drop=rng.choice(P.n_baselines, 7, replace=False))
    c = R.unit_solutions()
    r = solver(R, **kw).solve(R, c)
    ref = R.unit_solutions()
    rr = O.solve(R, ref, **kw)
    e = phase_ref_relerr(c, ref, R.n_antennas, R.n_sub)
    print(f"7 baselines missing: iterations {r.iterations} / oracle {rr.iterations}, "
          f"gain error {e:.2e}")
    assert r.iterations == rr.iterations
    assert e <= GAIN_TOL


@pytest.mark.parametrize("n_st", [23, 64, 2])
def test_fused_antenna_counts(n_st):
    """Odd antenna count (the last walk step covers all lanes), the full 64
    lanes, and the smallest problem (one baseline)."""
    wl = dataclasses.replace(S.SMALL, name=f"fused{n_st}", n_stations=n_st)
    w, geom, sky, gains, P = small_problem(wl)
    print(f"case: fused gram {n_st} stations", flush=True)
    kw = dict(max_iterations=30, accuracy=w.tolerance)
    a = P.unit_solutions()
    r = solver(P, **kw).solve(P, a)
    ref = P.unit_solutions()
    rr = O.solve(P, ref, **kw)
    e = phase_ref_relerr(a, ref, P.n_antennas, P.n_sub)
    print(f"iterations {r.iterations} / oracle {rr.iterations}, gain error {e:.2e}")
    assert r.iterations == rr.iterations
    assert e <= GAIN_TOL


def test_fused_phase_only_and_scalar(small):
    w, geom, sky, gains, P = small
    print("case: fused gram phase-only and scalar", flush=True)
    kw = dict(max_iterations=30, accuracy=w.tolerance)
    a = P.unit_solutions()
    solver(P, phase_only=True, **kw).solve(P, a)
    b = P.unit_solutions()
    solver(P, False, phase_only=True, **kw).solve(P, b)
    ref = P.unit_solutions()
    O.solve(P, ref, phase_only=True, **kw)
    assert phase_ref_relerr(a, ref, P.n_antennas, P.n_sub) <= GAIN_TOL
    assert phase_ref_relerr(a, b, P.n_antennas, P.n_sub) <= ORDER_TOL
    # scalar solver: one gain per antenna and direction for both polarizations
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks,
                 iteration=H.ITERATION_GRAM, n_solution_polarizations=1, **kw)
    c = np.ones((P.n_channel_blocks, P.n_antennas * P.n_sub), np.complex128)
    s.solve(P, c)
    s.set_gram_fusion(False)
    d = np.ones_like(c)
    s.solve(P, d)
    e = np.max(np.abs(c - d) / np.abs(d))
    print(f"scalar fused vs unfused {e:.2e}")
    assert e <= ORDER_TOL


def test_fused_batch_equals_single_and_warm_start():
    """Batched fused Gram solves equal single fused solves bit for bit, and
    the batched warm start's one-step denominators and start come through the
    fused kernel."""
    print("case: fused gram batch", flush=True)
    ws = [S.Workload(f"fb{i}", 10 + 5 * i, 2 + i, 1, 1, 8 * (1 + i % 2), 2, 2 + i, 40, 1e-5,
                     seed=i) for i in range(4)]
    seqs = [WS.IntervalSequence(w, 2, seed=60 + i,
                                solver_kwargs=dict(iteration=H.ITERATION_GRAM))
            for i, w in enumerate(ws)]
    singles = []
    for q in seqs:
        # the cross products' FP32 summation order follows the waves per tile
        q.solver.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=0, poll_interval=8)
        g = q.unit()
        r = q.solver.solve_problem(q.problems[0], g)
        singles.append((g, r.iterations))
    b = H.Batch([q.solver for q in seqs], [q.problems[0] for q in seqs])
    b.set_schedule(2, 1, 1, 8)
    got = [q.unit() for q in seqs]
    res = b.solve(got)
    for (g, its), x, r in zip(singles, got, res):
        assert r.iterations == its
        np.testing.assert_array_equal(x, g)
    # warm start from the first interval's solutions on the second interval
    ws_net = WS.WarmStart(hidden=32, seed=1, stretch=2)
    b.rebind([q.problems[1] for q in seqs])
    warm = [g.copy() for g in got]
    b.solve_warm(ws_net.net, warm, stretch=2)
    for q, g0, x in zip(seqs, got, warm):
        host = ws_net.initial(q.solver, q.problems[1], g0)
        ref = host.copy()
        q.solver.solve_problem(q.problems[1], ref)
        assert phase_ref_relerr(x, ref, q.w.n_stations, q.w.n_sub) <= ORDER_TOL
    b.close()


@pytest.mark.parametrize("storage,phase_only,graph",
                         [(H.MODEL_C64, False, 1), (H.MODEL_F16, False, 0),
                          (H.MODEL_F16, True, 1)])
def test_merged_finishing_step_equals_finishing_kernel(small, storage, phase_only, graph):
    """The finishing step inside the next iteration's kernel (fusion 1) gives
    the solutions, iteration count and step magnitudes of the separate
    finishing kernel (fusion 2) bit for bit, in single and batched solves."""
    w, geom, sky, gains, P = small
    print(f"case: merged finishing step storage={storage} phase_only={phase_only} "
          f"graph={graph}", flush=True)
    kw = dict(max_iterations=w.max_iterations, accuracy=w.tolerance, model_storage=storage,
              phase_only=phase_only)
    out = {}
    for mode in (1, 2):
        s = solver(P, mode, **kw)
        s.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=graph, poll_interval=8)
        g = P.unit_solutions()
        r = s.solve(P, g)
        out[mode] = (g, r.iterations)
    assert out[1][1] == out[2][1]
    np.testing.assert_array_equal(out[1][0], out[2][0])
    print(f"single: {out[1][1]} iterations in both forms")


def test_merged_finishing_step_batch_and_row_groups():
    """Batches of intervals of different shapes: the merged form equals the
    separate finishing kernel bit for bit at every row-group count (one to
    eight, the sums of groups beyond two added in a second pass)."""
    print("case: merged finishing step, batch", flush=True)
    ws = [S.Workload(f"mb{i}", 12 + 7 * i, 3 + i, 1, 1, 8, 2, 2 + 2 * i, 60, 1e-5, seed=20 + i)
          for i in range(3)]
    seqs = [WS.IntervalSequence(w, 1, seed=80 + i,
                                solver_kwargs=dict(iteration=H.ITERATION_GRAM))
            for i, w in enumerate(ws)]
    for groups in (1, 3, 8):
        res = {}
        for mode in (1, 2):
            for q in seqs:
                q.solver.set_gram_fusion(mode)
                q.solver.set_gram_solve_groups(groups)
            b = H.Batch([q.solver for q in seqs], [q.problems[0] for q in seqs])
            b.set_schedule(2, 1, 1, 8)
            b.set_gram_solve_groups(groups)
            got = [q.unit() for q in seqs]
            rs = b.solve(got)
            res[mode] = (got, [r.iterations for r in rs])
            b.close()
        assert res[1][1] == res[2][1]
        for x, y in zip(res[1][0], res[2][0]):
            np.testing.assert_array_equal(x, y)
        print(f"groups {groups}: iterations {res[1][1]}")
    for q in seqs:
        q.solver.set_gram_fusion(1)


@pytest.mark.parametrize("kw", [
    dict(max_iterations=7, accuracy=0.0, detect_stalling=False),      # the iteration limit
    dict(max_iterations=40, min_iterations=25, accuracy=1e-2),         # converged before min
    dict(max_iterations=300, accuracy=1e-12, detect_stalling=True),    # stall detection
    dict(max_iterations=1, accuracy=0.0, detect_stalling=False),       # a single iteration
], ids=["limit", "min_iterations", "stall", "one"])
def test_merged_decision_stopping_rules(small, kw):
    """The decision inside the merged kernel (fusion 1) against the decision
    kernel (fusion 2) under each stopping rule: the same iteration count,
    convergence flag, step magnitude and solutions, bit for bit."""
    w, geom, sky, gains, P = small
    print(f"case: merged decision {kw}", flush=True)
    out = {}
    for mode in (1, 2):
        s = solver(P, mode, **kw)
        s.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=1, poll_interval=4)
        g = P.unit_solutions()
        r = s.solve(P, g)
        out[mode] = (g, r)
        s.close()
    (g1, r1), (g2, r2) = out[1], out[2]
    print(f"iterations {r1.iterations} / {r2.iterations}, converged {r1.has_converged}, "
          f"last step {r1.last_step_magnitude:.3e}")
    assert (r1.iterations, r1.has_converged, r1.n_step_magnitudes) == \
        (r2.iterations, r2.has_converged, r2.n_step_magnitudes)
    assert r1.last_step_magnitude == r2.last_step_magnitude
    np.testing.assert_array_equal(g1, g2)
    if kw["max_iterations"] in (1, 7) and kw["accuracy"] == 0.0:
        assert r1.iterations == kw["max_iterations"] + 1  # max+1 when not converged
    if "min_iterations" in kw:
        assert r1.iterations >= kw["min_iterations"]
