"""Learned warm start for consecutive solution intervals.

DP3 starts each interval from unit gains or from the previous interval's
solutions (propagatesolutions, steps/DDECal.cc:487-499). This module adds a
third option: a small network, trained on the solver's own outputs, maps per
(channel block, antenna, solution)
  * the undamped one-step update of the previous solutions on the current
    interval (the data/model correlation summary, ddh_one_step), relative to
    the previous solutions, with its denominator as a confidence, and
  * the previous interval's drift
to a multiplicative correction of the previous solutions. For the full-Jones
solver (2x2 solutions) the same network reads the diagonal entries J_00,
J_11 and scales row p of the matrix by the correction of polarization p.
The plain solver
then runs to its stopping criterion from that start, so the network only
decides where the iteration begins, never when it ends.

Everything numerical runs on the GPU through libddecal_hip (features, network
forward/backward in bf16 MFMA GEMMs, Adam, application of the correction).
"""
from __future__ import annotations

import numpy as np

from . import hip as H
from . import synth as S


def sibling(solver: H.Solver, max_iterations, accuracy, detect_stalling=True, iteration=None):
    """A solver with the same shapes and step as `solver` but its own stopping
    criterion (device problems of one are valid for the other)."""
    c = solver.config
    return H.Solver(solver.n_antennas, solver._nsd, solver.n_channel_blocks,
                    n_correlations=solver.n_correlations, max_iterations=max_iterations,
                    accuracy=accuracy, step_size=c.step_size, detect_stalling=detect_stalling,
                    phase_only=bool(c.phase_only), model_storage=solver.model_storage,
                    iteration=solver.iteration if iteration is None else iteration,
                    n_solution_polarizations=solver.n_solution_polarizations)


def is_full_jones(solver: H.Solver):
    return solver.n_solution_polarizations == 4


def is_scalar(solver: H.Solver):
    return solver.n_solution_polarizations == 1


def as_pairs(solutions, scalar):
    """Scalar solutions [n_cb, n] as the pairs [n_cb, 2n] that hold each gain
    twice (the layout the device solves them in); other layouts unchanged."""
    g = np.asarray(solutions, np.complex128)
    return np.repeat(g, 2, axis=-1) if scalar else g


def diagonal(solutions, full_jones):
    """[n, 2] polarization values: a pair, or a 2x2 matrix's J_00, J_11."""
    g = np.asarray(solutions, np.complex128)
    return g.reshape(-1, 4)[:, ::3] if full_jones else g.reshape(-1, 2)


class WarmStart:
    """Features: the undamped one-step update of the previous solutions g0 on
    the current interval and two iterates of the damped solver (after
    `stretch` and 2 * `stretch` iterations: ga, gb), all relative to g0, plus
    the update's denominators. Output: log(g_start / gb) per polarization,
    the correction left after the two stretches (the start keeps their
    progress)."""

    N_IN = H.N_WARM_FEATURES
    N_OUT = H.N_WARM_OUTPUTS

    def __init__(self, hidden=256, seed=0, stretch=4):
        self.net = H.MLP(self.N_IN, hidden, self.N_OUT, seed)
        self.stretch = stretch
        self.trained = False
        self._aux = {}

    def _aux_solver(self, solver):
        key = id(solver)
        if key not in self._aux:
            self._aux[key] = (solver, sibling(solver, self.stretch, 0.0, False))
        return self._aux[key][1]

    def features(self, solver, dproblem, g0, return_stretch=False):
        """Features [n_pairs, 16] of the current interval (and gb, the
        iterate the start is formed from, with return_stretch)."""
        nxt, den = H.one_step(solver, dproblem, g0)
        aux = self._aux_solver(solver)
        ga = np.array(g0, np.complex128, copy=True)
        aux.solve_problem(dproblem, ga)
        gb = ga.copy()
        aux.solve_problem(dproblem, gb)
        sc = is_scalar(solver)
        x = H.warmstart_features(*[as_pairs(g, sc) for g in (g0, nxt, ga, gb)], den,
                                 solver.n_channel_blocks, solver.n_antennas, solver.n_sub,
                                 is_full_jones(solver))
        return (x, gb) if return_stretch else x

    @staticmethod
    def targets(base, final, full_jones=False, scalar=False):
        """log(final / base) per polarization as (log-magnitude, phase) pairs
        (base: the second stretch's iterate gb; for 2x2 solutions the ratio of
        the diagonal entries; a scalar gain's ratio for both)."""
        p = diagonal(as_pairs(base, scalar), full_jones)
        f = diagonal(as_pairs(final, scalar), full_jones)
        with np.errstate(all="ignore"):
            r = f / p
            out = np.stack([np.log(np.abs(r[:, 0])), np.angle(r[:, 0]),
                            np.log(np.abs(r[:, 1])), np.angle(r[:, 1])], axis=1)
        out[~np.isfinite(out)] = 0.0
        return out.astype(np.float32)

    def fit(self, x, t, epochs=10, batch=4096, lr=1e-3, seed=0):
        loss = self.net.train(x, t, epochs=epochs, batch=batch, lr=lr, seed=seed)
        self.trained = True
        return loss

    def initial(self, solver, dproblem, g0):
        x, gb = self.features(solver, dproblem, g0, return_stretch=True)
        pred = self.net.forward(x)
        if is_scalar(solver):
            # one gain: both outputs' mean (in float), applied to the pair
            sym = np.float32(0.5) * (pred[:, 0:2] + pred[:, 2:4])  # (log-magnitude, phase)
            pred = np.ascontiguousarray(np.concatenate([sym, sym], axis=1), np.float32)
            start = H.warmstart_apply(as_pairs(gb, True), pred)
            return np.ascontiguousarray(start.reshape(-1, 2)[:, 0].reshape(np.asarray(g0).shape))
        return H.warmstart_apply(gb, pred, is_full_jones(solver)).reshape(np.asarray(g0).shape)


def drifting_gains(w: S.Workload, rng, n_intervals, sigma_phase=0.05, sigma_amp=0.005):
    """Gains per interval: amplitude and phase random walks from the usual
    initial draw (amplitude 1 +- 0.1, phase uniform +-pi)."""
    shape = (w.n_channel_blocks, w.n_stations, w.n_sub, 2)
    amp = rng.uniform(0.9, 1.1, shape)
    ph = rng.uniform(-np.pi, np.pi, shape)
    out = []
    for _ in range(n_intervals):
        out.append((amp * np.exp(1j * ph)).reshape(w.n_channel_blocks, -1))
        amp = amp + rng.normal(0.0, sigma_amp, shape)
        ph = ph + rng.normal(0.0, sigma_phase, shape)
    return out


class IntervalSequence:
    """Consecutive solution intervals of one observation, built on the GPU on
    first use: same stations and sky, UVW advancing with earth rotation,
    drifting gains. `problems[i]` builds interval i (predict + data synthesis
    on the device) and keeps it resident until `release(i)`; the result does
    not depend on the order in which intervals are built."""

    def __init__(self, w: S.Workload, n_intervals, seed=0, storage=H.MODEL_C64,
                 sigma_phase=0.05, sigma_amp=0.005, solver_kwargs=None):
        rng = np.random.default_rng(seed)
        self.w = w
        self.seed = seed
        self.n_intervals = n_intervals
        self.geom0 = S.make_geometry(w, rng)
        self.sky = S.make_sky(w, self.geom0, rng)
        self.gains = drifting_gains(w, rng, n_intervals, sigma_phase, sigma_amp)
        kw = dict(max_iterations=w.max_iterations, accuracy=w.tolerance, step_size=w.step_size,
                  model_storage=storage)
        kw.update(solver_kwargs or {})
        self.solver = H.Solver(w.n_stations, w.sols_per_dir(), w.n_channel_blocks, **kw)
        self._ch = np.diff(S.channel_begin(w.n_channels, w.n_channel_blocks)).astype(np.int32)
        n_t = w.solint
        start = np.concatenate([[0], np.cumsum(w.sols_per_dir())[:-1]])
        self.sol_index = np.zeros((w.n_directions, n_t), np.uint32)
        for d in range(w.n_directions):
            self.sol_index[d] = (np.arange(n_t) * w.sols_per_dir()[d]) // n_t + start[d]
        self._comps = H.components_from_sky(self.sky)
        self.problems = _LazyProblems(self)

    def build(self, i, into=None):
        """Interval i as a device problem; `into` (a problem of this sequence)
        is refilled instead of allocating a new one."""
        w = self.w
        geom = S.make_geometry(w, None, start_step=i * w.solint, station_enu=self.geom0.station_enu)
        dp = into
        if dp is None:
            dp = H.DeviceProblem(self.solver, w.solint, geom.baselines[:, 0], geom.baselines[:, 1],
                                 self._ch, self.sol_index)
        dp.predict(geom.ref, geom.uvw, geom.freq, geom.chan_width, self._comps)
        dp.synthesize_data(self.gains[i], self.sol_index, w.noise_sigma,
                           seed=self.seed * 1000 + i)
        return dp

    def release(self, i):
        self.problems.release(i)

    def transient(self, i):
        """Interval i in one scratch problem reused by every call (no device
        allocation after the first), valid until the next call."""
        self._scratch = self.build(i, into=getattr(self, "_scratch", None))
        return self._scratch

    def unit(self):
        """Unit gains: ones, or identity matrices for 2x2 solutions."""
        n = self.w.n_stations * self.w.n_sub
        if is_scalar(self.solver):
            return np.ones((self.w.n_channel_blocks, n), np.complex128)
        if is_full_jones(self.solver):
            eye = np.tile(np.array([1, 0, 0, 1], np.complex128), n)
            return np.tile(eye, (self.w.n_channel_blocks, 1))
        return np.ones((self.w.n_channel_blocks, n * 2), np.complex128)


class _LazyProblems:
    def __init__(self, seq):
        self._seq = seq
        self._built = {}

    def __len__(self):
        return self._seq.n_intervals

    def __getitem__(self, i):
        if not 0 <= i < self._seq.n_intervals:
            raise IndexError(i)
        if i not in self._built:
            self._built[i] = self._seq.build(i)
        return self._built[i]

    def release(self, i):
        dp = self._built.pop(i, None)
        if dp is not None:
            dp.close()


def training_set(seq: IntervalSequence, ws: WarmStart, first, last, converge_iterations=600,
                 transient=False):
    """Solves intervals [first, last) to convergence (tolerance 1e-9, at most
    `converge_iterations`) from the previous interval's solutions and collects
    (features, targets) for the learned start. Returns x, t and the last
    converged solutions. With `transient`, the intervals are built one after
    another into one scratch problem instead of staying resident."""
    long = sibling(seq.solver, converge_iterations, 1e-9, False)
    prev = None
    xs, ts = [], []
    for i in range(first, last):
        dp = seq.transient(i) if transient else seq.problems[i]
        if prev is None:
            sol = seq.unit()
            long.solve_problem(dp, sol)
            prev = sol
            continue
        x, gb = ws.features(seq.solver, dp, prev, return_stretch=True)
        xs.append(x)
        sol = prev.copy()
        long.solve_problem(dp, sol)
        ts.append(ws.targets(gb, sol, is_full_jones(seq.solver), is_scalar(seq.solver)))
        prev = sol
    return np.concatenate(xs), np.concatenate(ts), prev


def training_sets_batched(seqs, ws: WarmStart, first, last, converge_iterations=1000,
                          use_graph=1):
    """training_set for several interval sequences of one shape at once: each
    step builds interval i of every sequence into its scratch problem, takes
    the features from the previous converged solutions, and solves all of
    them to convergence (tolerance 1e-9) in one batch with the Gram
    iteration (its iterations replayed as HIP graphs unless use_graph is 0).
    Returns per sequence (x, t, last converged solutions)."""
    dps = [q.transient(first) for q in seqs]
    longs = [sibling(q.solver, converge_iterations, 1e-9, False, iteration=H.ITERATION_GRAM)
             for q in seqs]
    batch = H.Batch(longs, dps)
    batch.set_schedule(use_graph=use_graph)
    prev = [q.unit() for q in seqs]
    batch.solve(prev)  # the first interval of each sequence from unit gains
    xs = [[] for _ in seqs]
    ts = [[] for _ in seqs]
    for i in range(first + 1, last):
        dps = [q.transient(i) for q in seqs]  # same scratch problems, refilled
        gbs = []
        for k, q in enumerate(seqs):
            x, gb = ws.features(q.solver, dps[k], prev[k], return_stretch=True)
            xs[k].append(x)
            gbs.append(gb)
        sol = [p.copy() for p in prev]
        batch.solve(sol)
        for k in range(len(seqs)):
            ts[k].append(ws.targets(gbs[k], sol[k], is_full_jones(seqs[k].solver),
                                    is_scalar(seqs[k].solver)))
        prev = sol
    batch.close()
    for s in longs:
        s.close()
    return [(np.concatenate(x), np.concatenate(t), p) for x, t, p in zip(xs, ts, prev)]
