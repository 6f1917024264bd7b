#!/usr/bin/env python3
"""DDECal solves/s on one MI355X: the iterative diagonal solver at LOFAR scale
(62 stations / 1891 baselines, 8 directions x 20 point+Gaussian components,
240 channels in 24 channel blocks, solution interval 10, 100 iterations,
tolerance 1e-6), with inputs resident in HBM.

A step is one solve: one solution interval iterated from unit gains to the
stopping criterion (max iterations, convergence or stall). The synthetic
interval is built on the device: HIP predict of the sky model into the solver's
model planes, data = sum_d G_a M_d G_b^H + noise (sigma 0.01 Jy per part).

Prints one JSON line. `roofline` is for the correlate kernel (the streaming
pass over data and model planes): algorithmic bytes per launch
= sum_cb n_vis * (16 + n_dir * 16) divided by its mean launch duration from HIP
events recorded on the solver's stream around every launch that did work.
`cpu_baseline` times the CPU oracle (a restatement of DP3's
IterativeDiagonalSolver, threads over channel blocks) on a bounded sample of
the same interval and scales to solves/s with the GPU's iteration count.
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

METRIC = "DDECal solves/s (62 st, 8 dir, 240 ch): plain vs batched+warm-start, same χ²"
HBM_PEAK_GBS = 8000.0


def build_interval(w, seed, storage=H.MODEL_C64, solver_kwargs=None):
    rng = np.random.default_rng(seed)
    geom = S.make_geometry(w, rng)
    sky = S.make_sky(w, geom, rng)
    gains = S.true_gains(w, rng)
    kw = dict(max_iterations=w.max_iterations, accuracy=w.tolerance, step_size=w.step_size,
              model_storage=storage)
    kw.update(solver_kwargs or {})
    solver = H.Solver(w.n_stations, w.sols_per_dir(), w.n_channel_blocks, **kw)
    ch = np.diff(S.channel_begin(w.n_channels, w.n_channel_blocks)).astype(np.int32)
    n_t = geom.uvw.shape[0]
    sol_index = np.tile(np.arange(w.n_directions, dtype=np.uint32)[:, None], (1, n_t))
    dp = H.DeviceProblem(solver, n_t, geom.baselines[:, 0], geom.baselines[:, 1], ch, sol_index)
    dp.predict(geom.ref, geom.uvw, geom.freq, geom.chan_width, H.components_from_sky(sky))
    dp.synthesize_data(gains, sol_index, w.noise_sigma, seed=seed + 1)
    return solver, dp, geom, sky, gains, sol_index


def algorithmic_bytes(w, model_bytes=16):
    n_vis = w.n_baselines * w.n_channels * w.solint
    return n_vis * (16 + w.n_directions * model_bytes)


def cpu_baseline(dp, w, iterations_per_solve, sample_iterations, threads):
    """Oracle (reference loop structure) on `sample_iterations` iterations."""
    from oracle import pyoracle as O
    cbe = S.channel_begin(w.n_channels, w.n_channel_blocks)
    n_t, n_bl = dp.n_times, dp.n_baselines
    datas, models, a1s, a2s, sms, offs = [], [], [], [], [], [0]
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
    bl = np.array([(a, b) for a in range(w.n_stations) for b in range(a + 1, w.n_stations)])
    for cb in range(w.n_channel_blocks):
        nc = int(cbe[cb + 1] - cbe[cb])
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
    dt = time.perf_counter() - t0
    per_iter = dt / sample_iterations
    return {
        "value": 1.0 / (per_iter * iterations_per_solve),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{sample_iterations} of {iterations_per_solve} iterations of one "
                   f"{w.name} interval ({per_iter:.3f} s/iteration, {threads} threads over "
                   f"channel blocks), scaled to a full solve"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--waves", type=int, default=2)
    ap.add_argument("--poll", type=int, default=8)
    ap.add_argument("--rows", type=int, default=1)
    ap.add_argument("--cpu-iterations", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="also time a plain copy and a plain read of 1 GiB")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(solution intervals shard across GPUs without communication; "
              "the one-GPU number is the per-GPU throughput)", file=sys.stderr)
        sys.exit(2)

    w = S.LOFAR
    solver, dp, geom, sky, gains, sol_index = build_interval(w, args.seed)
    solver.set_schedule(waves_per_tile=args.waves, rows_per_trip=args.rows,
                        use_graph=int(not args.no_graph), poll_interval=args.poll)
    n_cb = w.n_channel_blocks
    for _ in range(args.warmup):
        sol = np.ones((n_cb, w.n_stations * w.n_sub * 2), np.complex128)
        solver.solve_problem(dp, sol)
    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sol = np.ones((n_cb, w.n_stations * w.n_sub * 2), np.complex128)
        r = solver.solve_problem(dp, sol)
        iters.append(r.iterations)
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    # Correlate launch duration: one more solve with HIP events recorded on the
    # solver's stream around every correlate launch (graph replay off).
    solver.set_profiling(True)
    prof_sol = np.ones((n_cb, w.n_stations * w.n_sub * 2), np.complex128)
    solver.solve_problem(dp, prof_sol)
    corr_ms, corr_n, _ = solver.last_timing()
    solver.set_profiling(False)
    # solution quality against the generating gains (phase referenced)
    err = float(np.max(np.abs(S.phase_referenced(sol, w.n_stations, w.n_sub) -
                              S.phase_referenced(gains, w.n_stations, w.n_sub))))
    bytes_per = algorithmic_bytes(w)
    avg_ms = corr_ms / max(corr_n, 1)
    achieved = bytes_per / (avg_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded station layout, HIP-predicted sky, drawn gains and noise)",
        "config": {
            "workload": f"{w.name}: 62 stations/1891 baselines, 8 directions x 20 "
                        "point+Gaussian components, 240 channels in 24 blocks, solint 10, "
                        "max 100 iterations, tol 1e-6, step 0.2",
            "form": "plain",
            "graph": not args.no_graph,
            "poll_interval": args.poll,
            "waves_per_tile": args.waves,
            "rows_per_trip": args.rows,
            "model_storage": "complex64",
            "iterations_per_solve": iters,
            "max_gain_error_vs_truth": err,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "correlate",
            "algorithmic_bytes_per_launch": bytes_per,
            "mean_launch_ms": round(avg_ms, 5),
        },
    }
    if args.calibrate:
        copy_gbs, read_gbs = H.calibrate_bandwidth(1 << 30, 10)
        line["calibration"] = {"copy_GBps": round(copy_gbs, 1), "read_GBps": round(read_gbs, 1),
                               "bytes": 1 << 30}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(dp, w, int(np.median(iters)), args.cpu_iterations,
                                            args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
