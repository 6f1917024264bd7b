#!/usr/bin/env python3
"""DDECal solves/s on one MI355X for the LOFAR-scale solution interval:
62 stations (1891 baselines), 8 directions x 20 point+Gaussian components,
240 channels in 24 channel blocks, solution interval 10, step 0.2,
at most 100 iterations, tolerance 1e-6. Inputs are resident in HBM.

A step is one batch of solves (one per interval sequence); `value` is the
whole-job throughput in solves/s of the optimized form:

  * plain          reference settings: unit gains, complex64 model planes,
                   stopping criterion max 100 iterations / tolerance 1e-6.
                   On this workload no start reaches the tolerance within 100
                   iterations, and from unit gains the chi^2 is then far from
                   converged, so it is reported with its chi^2 excess.
  * plain_equal    unit gains, complex64, iteration budget raised to the
                   smallest count (from a ladder) whose chi^2 is within 0.1 %
                   of the converged chi^2 of the interval.
  * optimized      batched (one launch per kernel over all sequences'
                   intervals), FP16 model planes with per-direction scale,
                   learned warm start from the previous interval's solutions,
                   iteration budget calibrated the same way to 0.1 % of the
                   converged chi^2. Feature passes (one undamped step and
                   2 x 4 damped iterations) and the network run inside the
                   timed region.

Synthetic data: seeded station layout, HIP-predicted sky, gains drifting by a
phase random walk (0.05 rad / interval) and an amplitude random walk (0.005),
thermal noise 0.01 Jy per part, data synthesized on the device.

`roofline` is for the dominant kernel of the optimized form (the batched
correlate over FP16 planes): algorithmic bytes per launch
= sum over intervals of n_vis * (16 + n_dir * 8), over its mean duration from
HIP events on the batch stream around each launch that did work.
`roofline_plain` is the same for the complex64 correlate of the plain form.
`cpu_baseline` times the CPU oracle (a restatement of DP3's
IterativeDiagonalSolver, threads over channel blocks) on a bounded sample of
the plain form and scales it to solves/s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from dp3_amd import hip as H  # noqa: E402
from dp3_amd import synth as S  # noqa: E402
from dp3_amd import warmstart as WS  # noqa: E402

METRIC = "DDECal solves/s (62 st, 8 dir, 240 ch): plain vs batched+warm-start, same χ²"
HBM_PEAK_GBS = 8000.0
CHI2_TOL = 1e-3


def algorithmic_bytes(w, model_bytes):
    n_vis = w.n_baselines * w.n_channels * w.solint
    return n_vis * (16 + w.n_directions * model_bytes)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def converged_chi2(solver, dp, start, iterations=1500):
    s = WS.sibling(solver, iterations, 1e-10, False)
    g = np.array(start, copy=True)
    s.solve_problem(dp, g)
    return H.chi2(solver, dp, g), g


def min_iterations(solver, dp, start, target, ladder):
    """Smallest ladder entry whose chi^2 from `start` is within CHI2_TOL of target."""
    for k in ladder:
        s = WS.sibling(solver, k, 1e-6, False)
        g = np.array(start, copy=True)
        s.solve_problem(dp, g)
        if H.chi2(solver, dp, g) <= target * (1 + CHI2_TOL):
            return k
    return ladder[-1]


def cpu_baseline(dp, w, iterations_per_solve, sample_iterations, threads):
    """Oracle (reference loop structure) on `sample_iterations` iterations."""
    from oracle import pyoracle as O
    cbe = S.channel_begin(w.n_channels, w.n_channel_blocks)
    n_t, n_bl = dp.n_times, dp.n_baselines
    datas, models, a1s, a2s, sms, offs = [], [], [], [], [], [0]
    bl = np.array([(a, b) for a in range(w.n_stations) for b in range(a + 1, w.n_stations)])
    for cb in range(w.n_channel_blocks):
        nc = int(cbe[cb + 1] - cbe[cb])
        d = dp.download_plane(cb, 0).reshape(n_t, nc, dp.bl_stride, 2)[:, :, :n_bl]
        datas.append(np.ascontiguousarray(d.transpose(0, 2, 1, 3)).reshape(-1, 2))
        for di in range(w.n_directions):
            m = dp.download_plane(cb, di + 1).reshape(n_t, nc, dp.bl_stride, 2)[:, :, :n_bl]
            models.append(np.ascontiguousarray(m.transpose(0, 2, 1, 3)).reshape(-1, 2))
        nv = n_t * n_bl * nc
        offs.append(offs[-1] + nv)
        sms.append(np.repeat(np.arange(w.n_directions, dtype=np.uint32), nv))
        a1s.append(np.repeat(np.tile(bl[:, 0], n_t), nc).astype(np.uint32))
        a2s.append(np.repeat(np.tile(bl[:, 1], n_t), nc).astype(np.uint32))
    P = S.SolveProblem(w.n_stations, w.n_directions, w.n_channel_blocks, 2,
                       w.sols_per_dir(), np.array(offs, np.int64), np.concatenate(datas),
                       np.concatenate(models), np.concatenate(a1s), np.concatenate(a2s),
                       np.concatenate(sms), n_t, n_bl, cbe)
    sol = P.unit_solutions()
    t0 = time.perf_counter()
    O.solve(P, sol, max_iterations=sample_iterations, accuracy=0.0, detect_stalling=False,
            n_threads=threads)
    per_iter = (time.perf_counter() - t0) / sample_iterations
    return {
        "value": 1.0 / (per_iter * iterations_per_solve),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{sample_iterations} of {iterations_per_solve} iterations of one "
                   f"{w.name} interval ({per_iter:.3f} s/iteration, {threads} threads over "
                   f"channel blocks), scaled to a plain-form solve"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="timed batch steps (intervals per sequence)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sequences", type=int, default=4, help="intervals per batch")
    ap.add_argument("--train", type=int, default=6, help="training intervals per sequence")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-iterations", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="also time a plain copy and a plain read of 1 GiB")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(solution intervals shard across GPUs without communication; "
              "the one-GPU number is the per-GPU throughput)", file=sys.stderr)
        sys.exit(2)

    w = S.LOFAR
    B, T, TR, WU = args.sequences, args.steps, args.train, args.warmup
    n_total = TR + WU + T
    eval0 = TR + WU
    t_setup = time.perf_counter()
    seqs16 = [WS.IntervalSequence(w, n_total, seed=args.seed + 11 * b, storage=H.MODEL_F16)
              for b in range(B)]
    seqs64 = [WS.IntervalSequence(w, n_total, seed=args.seed + 11 * b, storage=H.MODEL_C64,
                                  build_from=TR)
              for b in range(B)]
    log(f"built {B} x {n_total} intervals in {time.perf_counter() - t_setup:.1f} s")

    # ---- learned warm start: trained on the solver's own converged outputs
    ws = WS.WarmStart(hidden=256, seed=args.seed)
    xs, ts, prevs = [], [], []
    for q in seqs16:
        x, t, prev = WS.training_set(q, ws, 0, TR)
        xs.append(x)
        ts.append(t)
        prevs.append(prev)
    x, t = np.concatenate(xs), np.concatenate(ts)
    loss = ws.fit(x, t, epochs=10, batch=4096, lr=1e-3, seed=args.seed)
    log(f"warm start: {len(x)} samples, loss {loss:.3e}, target variance {t.var():.3e}")

    # ---- converged chi^2 of every later interval, following the chain
    chain = list(prevs)
    conv = {}
    for i in range(TR, n_total):
        for b, q in enumerate(seqs16):
            c, g = converged_chi2(q.solver, q.problems[i], chain[b])
            conv[(b, i)] = c
            chain[b] = g
    # ---- iteration budgets reaching 0.1 % of the converged chi^2 (interval TR)
    q0, q64 = seqs16[0], seqs64[0]
    start_l = ws.initial(q0.solver, q0.problems[TR], prevs[0])
    n_opt = min_iterations(q0.solver, q0.problems[TR], start_l, conv[(0, TR)],
                           list(range(60, 401, 5)))
    n_plain = min_iterations(q64.solver, q64.problems[TR], q64.unit(), conv[(0, TR)],
                             list(range(100, 801, 10)))
    log(f"iteration budgets: optimized {n_opt}, plain_equal {n_plain}")

    # ---- optimized form: batched + FP16 + learned warm start
    solvers_opt = [WS.sibling(q.solver, n_opt, 1e-6) for q in seqs16]
    prev = list(prevs)
    chi_opt, tuned = [], None
    t_opt = 0.0
    n_solves = 0
    batch = None
    for i in range(TR, n_total):
        timed = i >= eval0
        t0 = time.perf_counter()
        inits = [ws.initial(q.solver, q.problems[i], prev[b]) for b, q in enumerate(seqs16)]
        batch = H.Batch(solvers_opt, [q.problems[i] for q in seqs16])
        if tuned is None:
            tuned = batch.autotune()
            log(f"autotuned batch schedule: waves {tuned[0]}, rows {tuned[1]}, "
                f"correlate {tuned[2]:.4f} ms")
        batch.set_schedule(tuned[0], tuned[1], 1, 8)
        res = batch.solve(inits)
        dt = time.perf_counter() - t0
        if timed:
            t_opt += dt
            n_solves += B
            for b, q in enumerate(seqs16):
                chi_opt.append(H.chi2(q.solver, q.problems[i], inits[b]) / conv[(b, i)] - 1)
        prev = inits
        its_opt = [r.iterations for r in res]
    value = n_solves / t_opt
    # roofline of the batched FP16 correlate: one more profiled batch
    batch.set_profiling(True)
    batch.set_schedule(tuned[0], tuned[1], 0, 8)
    batch.solve([np.array(p, copy=True) for p in prev])
    corr_ms, corr_n = batch.last_timing()
    avg_ms = corr_ms / max(corr_n, 1)
    bytes_opt = B * algorithmic_bytes(w, 8)
    achieved = bytes_opt / (avg_ms * 1e-3) / 1e9

    # ---- plain forms (complex64, unit gains, one interval at a time)
    plain = WS.sibling(q64.solver, w.max_iterations, w.tolerance)
    plain_eq = WS.sibling(q64.solver, n_plain, 1e-6, False)
    for s in (plain, plain_eq):
        s.set_schedule(waves_per_tile=2, rows_per_trip=1, use_graph=1, poll_interval=8)
    s_warm = q64.unit()
    plain.solve_problem(seqs64[0].problems[TR], s_warm)
    results = {}
    for name, s in (("plain", plain), ("plain_equal", plain_eq)):
        t_all, chis, its = 0.0, [], []
        for i in range(eval0, n_total):
            for b, q in enumerate(seqs64):
                g = q.unit()
                t0 = time.perf_counter()
                r = s.solve_problem(q.problems[i], g)
                t_all += time.perf_counter() - t0
                chis.append(H.chi2(s, q.problems[i], g) / conv[(b, i)] - 1)
                its.append(r.iterations)
        results[name] = (B * T / t_all, chis, its)
    # roofline of the complex64 correlate (plain form)
    plain.set_profiling(True)
    g = q64.unit()
    plain.solve_problem(seqs64[0].problems[eval0], g)
    pms, pn, _ = plain.last_timing()
    plain.set_profiling(False)
    p_avg = pms / max(pn, 1)
    p_achieved = algorithmic_bytes(w, 16) / (p_avg * 1e-3) / 1e9

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": T,
        "warmup": WU,
        "ms_per_step": round(t_opt / T * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic: seeded station layout, HIP-predicted sky, drifting gains, noise "
                 "0.01 Jy; data synthesized on the device"),
        "config": {
            "workload": (f"{w.name}: 62 stations/1891 baselines, 8 directions x 20 "
                         "point+Gaussian components, 240 channels in 24 blocks, solint 10, "
                         "step 0.2, max 100 iterations, tol 1e-6"),
            "form": ("optimized: batched x%d interval sequences, FP16 model planes, learned warm "
                     "start (3-layer 256-hidden bf16 MLP), iteration budget %d to chi^2 within "
                     "0.1%% of converged" % (B, n_opt)),
            "intervals_per_step": B,
            "optimized_iterations": its_opt,
            "optimized_chi2_excess_max": float(np.max(chi_opt)),
            "plain_solves_per_s": round(results["plain"][0], 3),
            "plain_chi2_excess_median": float(np.median(results["plain"][1])),
            "plain_iterations": int(np.median(results["plain"][2])),
            "plain_equal_chi2_solves_per_s": round(results["plain_equal"][0], 3),
            "plain_equal_chi2_iterations": n_plain,
            "plain_equal_chi2_excess_max": float(np.max(results["plain_equal"][1])),
            "speedup_vs_plain_equal_chi2": round(value / results["plain_equal"][0], 3),
            "speedup_vs_plain": round(value / results["plain"][0], 3),
            "warm_start_train_samples": int(len(x)),
            "autotuned_schedule": {"waves_per_tile": tuned[0], "rows_per_trip": tuned[1]},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "batched correlate, FP16 model planes",
            "algorithmic_bytes_per_launch": bytes_opt,
            "mean_launch_ms": round(avg_ms, 5),
        },
        "roofline_plain": {
            "bound": "hbm",
            "achieved": round(p_achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(p_achieved / HBM_PEAK_GBS, 4),
            "kernel": "correlate, complex64 model planes",
            "algorithmic_bytes_per_launch": algorithmic_bytes(w, 16),
            "mean_launch_ms": round(p_avg, 5),
        },
    }
    if args.calibrate:
        copy_gbs, read_gbs = H.calibrate_bandwidth(1 << 30, 10)
        line["calibration"] = {"copy_GBps": round(copy_gbs, 1), "read_GBps": round(read_gbs, 1),
                               "bytes": 1 << 30}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(seqs64[0].problems[eval0], w, w.max_iterations + 1,
                                            args.cpu_iterations, args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
