"""Warm-start network on the GPU: bf16 MFMA GEMM against numpy, forward and
backward against the FP64 oracle (same bf16 rounding points), training, the
one-step estimate and chi^2 against the CPU oracle, and the warm start's gain
over starting from the previous solutions (propagatesolutions) and from unity."""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from dp3_amd import warmstart as WS
from oracle import mlp_oracle as MO
from oracle import pyoracle as O
from problems import small_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 16, 0, 0), (100, 37, 53, 0, 0),
                                         (33, 256, 256, 1, 0), (256, 4, 300, 0, 1),
                                         (5, 7, 3, 1, 1), (256, 256, 4096, 1, 0),
                                         (16, 256, 2000, 1, 0), (256, 4, 1100, 1, 0),
                                         (4096, 256, 256, 0, 1)])
def test_gemm_bf16_matches_numpy(M, N, K, ta, tb):
    print(f"case: gemm {M}x{N}x{K} ta={ta} tb={tb}", flush=True)
    rng = np.random.default_rng(M * 1000 + N)
    a = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    b = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    c = H.gemm_bf16(a, b, bool(ta), bool(tb))
    ar, br = MO.bf16(a), MO.bf16(b)
    ref = (ar.T if ta else ar) @ (br.T if tb else br)
    scale = np.abs(ar).max() * np.abs(br).max() * K
    assert np.abs(c - ref).max() <= 1e-6 * scale + 1e-5 * np.abs(ref).max()


def test_mlp_gradients_match_fp64_oracle():
    print("case: mlp gradients", flush=True)
    n_in, h, n_out, n = 16, 64, 4, 333
    net = H.MLP(n_in, h, n_out, seed=3)
    rng = np.random.default_rng(0)
    params = (0.2 * rng.normal(size=net.num_params())).astype(np.float32)
    net.set_params(params)
    x = rng.normal(size=(n, n_in)).astype(np.float32)
    t = rng.normal(size=(n, n_out)).astype(np.float32)
    loss, g = net.loss_and_grads(x, t)
    ref_loss, ref_g, ref_y = MO.loss_and_grads(params, x, t, n_in, h, n_out)
    np.testing.assert_allclose(net.forward(x), ref_y, rtol=0, atol=1e-4 * np.abs(ref_y).max())
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    # per parameter block, relative to the block's scale
    sizes = [n_in * h, h, h * h, h, h * n_out, n_out]
    o = 0
    for sz in sizes:
        gb, rb = g[o:o + sz], ref_g[o:o + sz]
        assert np.abs(gb - rb).max() <= 2e-3 * np.abs(rb).max() + 1e-9, o
        o += sz


@pytest.mark.parametrize("h,n", [(32, 1), (96, 45), (256, 4099), (256, 32), (40, 70)])
def test_fused_forward_equals_layered(h, n):
    """The one-kernel forward pass gives the layer-by-layer GEMM path's bits
    (same MFMA steps over the same bf16 operands); 40 hidden units run layer
    by layer either way. Partial last row blocks (n not a multiple of 32)
    included. After set_params the packed weights follow the new values."""
    print(f"case: fused forward h={h} n={n}", flush=True)
    net = H.MLP(16, h, 4, seed=7)
    rng = np.random.default_rng(h + n)
    x = (2.0 * rng.normal(size=(n, 16))).astype(np.float32)
    for k in range(2):
        params = (0.1 * rng.normal(size=net.num_params())).astype(np.float32)
        net.set_params(params)
        net.set_fused_forward(True)
        fused = net.forward(x)
        net.set_fused_forward(False)
        layered = net.forward(x)
        np.testing.assert_array_equal(fused, layered)
        # FP64 oracle with the same bf16 rounding points: FP32 sums in another
        # order can move a hidden activation across a bf16 rounding boundary
        _, _, ref = MO.loss_and_grads(params, x, np.zeros((n, 4), np.float32), 16, h, 4)
        assert np.abs(fused - ref).max() <= 1e-3 * np.abs(ref).max()
    net.close()


def test_mlp_training_fits_a_smooth_map():
    print("case: mlp training", flush=True)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(20000, 16)).astype(np.float32)
    t = np.stack([np.tanh(x[:, 0]), 0.5 * x[:, 1] * x[:, 2], np.sin(x[:, 3]), x[:, 4] ** 2 / 4],
                 axis=1).astype(np.float32)
    net = H.MLP(16, 128, 4, seed=0)
    l0, _ = net.loss_and_grads(x[:4096], t[:4096])
    loss = net.train(x, t, epochs=20, batch=1024, lr=3e-3, seed=0)
    l1, _ = net.loss_and_grads(x[:4096], t[:4096])
    print(f"loss {l0:.4f} -> {l1:.4f} (last epoch {loss:.4f})")
    assert l1 < 0.1 * l0


@pytest.fixture(scope="module")
def small():
    return small_problem()


def test_one_step_and_chi2_match_oracle(small):
    print("case: one step + chi2", flush=True)
    w, geom, sky, gains, P = small
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks)
    dp = H.DeviceProblem.from_problem(s, P)
    dp.upload(P)
    rng = np.random.default_rng(4)
    start = gains * (1 + 0.05 * rng.normal(size=gains.shape))
    nxt, den = H.one_step(s, dp, start)
    ref = O.iterate_once(P, start, fp64=True)
    assert np.max(np.abs(nxt - ref) / np.abs(ref)) <= 1e-4
    assert np.all(den > 0)
    c = H.chi2(s, dp, gains)
    assert abs(c - O.chi2(P, gains)) <= 1e-9 * O.chi2(P, gains)


def test_warm_start_reduces_iterations():
    """Consecutive intervals with drifting gains: train on the first intervals'
    converged solver outputs, evaluate on the rest. Every start converges to
    the same chi^2 (the plain solver decides when to stop); the learned start
    needs fewer iterations than the previous solutions, which need fewer than
    unity."""
    print("case: warm start sequence", flush=True)
    w = S.Workload("ws", 16, 3, 2, 1, 16, 4, 4, 400, 1e-6, seed=5)
    seq = WS.IntervalSequence(w, 36, seed=5)
    ws = WS.WarmStart(hidden=256, seed=0)
    x, t, prev = WS.training_set(seq, ws, 0, 26)
    loss = ws.fit(x, t, epochs=30, batch=1024, lr=1e-3)
    print(f"train samples {len(x)}, last epoch loss {loss:.3e}, target var {t.var():.3e}")
    its = {"unit": [], "previous": [], "learned": []}
    chis = {k: [] for k in its}
    for i in range(26, 36):
        dp = seq.problems[i]
        starts = {"unit": seq.unit(), "previous": prev.copy(),
                  "learned": ws.initial(seq.solver, dp, prev)}
        finals = {}
        for k, sol in starts.items():
            r = seq.solver.solve_problem(dp, sol)
            its[k].append(r.iterations)
            chis[k].append(H.chi2(seq.solver, dp, sol))
            finals[k] = sol
        prev = finals["previous"]
    for k in its:
        print(f"{k:9s} mean iterations {np.mean(its[k]):6.1f}  mean chi2 {np.mean(chis[k]):.6e}")
    for k in ("unit", "learned"):
        np.testing.assert_allclose(chis[k], chis["previous"], rtol=1e-3)
    # the learned start includes 2 x stretch damped iterations of its own
    learned_total = np.mean(its["learned"]) + 2 * ws.stretch + 1
    print(f"learned incl. feature iterations {learned_total:.1f}")
    assert learned_total < np.mean(its["previous"]) < np.mean(its["unit"])


def test_batched_warm_start_matches_host_sequence():
    """ddh_batch_solve_warm (features, network and start for every interval on
    the device) gives the same starts and solves as the per-interval host
    sequence one_step -> 2 x stretch iterations -> features -> forward ->
    apply, followed by a batched solve. Same schedule (4 waves per tile, one
    row per trip) on both sides, so the arithmetic is the same."""
    print("case: batched warm start vs host sequence", flush=True)
    w = S.Workload("wsb", 12, 3, 2, 1, 12, 3, 4, 60, 1e-5, seed=9)
    seqs = [WS.IntervalSequence(w, 2, seed=20 + k) for k in range(3)]
    ws = WS.WarmStart(hidden=64, seed=1, stretch=3)
    prevs = []
    for q in seqs:
        g = q.unit()
        q.solver.solve_problem(q.problems[0], g)
        prevs.append(g)
    solvers = [q.solver for q in seqs]
    dps = [q.problems[1] for q in seqs]
    host = [ws.initial(q.solver, q.problems[1], prevs[k]) for k, q in enumerate(seqs)]
    b1 = H.Batch(solvers, dps)
    b1.set_schedule(4, 1, 1, 8)
    r_host = b1.solve(host)
    dev = [p.copy() for p in prevs]
    b2 = H.Batch(solvers, dps)
    b2.set_schedule(4, 1, 1, 8)
    r_dev = b2.solve_warm(ws.net, dev, stretch=ws.stretch)
    assert b2.last_warm_ms() > 0.0
    for k in range(len(seqs)):
        assert r_dev[k].iterations == r_host[k].iterations
        np.testing.assert_allclose(dev[k], host[k], rtol=1e-9, atol=1e-12)
    # a second call on the same batch (rebound to the same problems) repeats exactly
    again = [p.copy() for p in prevs]
    b2.rebind(dps)
    b2.solve_warm(ws.net, again, stretch=ws.stretch)
    for k in range(len(seqs)):
        np.testing.assert_array_equal(again[k], dev[k])


def test_concurrent_batches_share_a_network():
    """Two batches (each its own HIP stream) driven from two host threads at
    once, sharing one warm-start network, give the same solutions and
    iteration counts bit for bit as the same solves run one after the other:
    the batches share nothing but the network's weights, and the forward pass
    reads and writes each batch's own rows."""
    import threading
    print("case: concurrent batches, one network", flush=True)
    w = S.Workload("wsc", 12, 3, 2, 1, 12, 3, 4, 60, 1e-5, seed=11)
    seqs = [WS.IntervalSequence(w, 2, seed=40 + k, solver_kwargs=dict(iteration=H.ITERATION_GRAM))
            for k in range(4)]
    ws = WS.WarmStart(hidden=64, seed=2, stretch=3)
    prevs = []
    for q in seqs:
        g = q.unit()
        q.solver.solve_problem(q.problems[0], g)
        prevs.append(g)
    halves = [[0, 1], [2, 3]]
    batches = [H.Batch([seqs[i].solver for i in h], [seqs[i].problems[1] for i in h])
               for h in halves]
    for b in batches:
        b.set_schedule(4, 1, 1, 8)

    def run(k, out, res):
        sols = [prevs[i].copy() for i in halves[k]]
        for rep in range(3):  # repeated, so the two threads overlap
            batches[k].rebind([seqs[i].problems[1] for i in halves[k]])
            sols = [prevs[i].copy() for i in halves[k]]
            res[k] = batches[k].solve_warm(ws.net, sols, stretch=ws.stretch)
        out[k] = sols

    seq_out, seq_res = [None, None], [None, None]
    for k in range(2):
        run(k, seq_out, seq_res)
    con_out, con_res = [None, None], [None, None]
    threads = [threading.Thread(target=run, args=(k, con_out, con_res)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for k in range(2):
        for m in range(2):
            assert con_res[k][m].iterations == seq_res[k][m].iterations
            np.testing.assert_array_equal(con_out[k][m], seq_out[k][m])
    for b in batches:
        b.close()


FJ = dict(n_correlations=4, n_solution_polarizations=4)


def pairs_to_matrices(g):
    """[n_cb, n*2] diagonal pairs -> [n_cb, n*4] row-major 2x2 matrices."""
    g = np.asarray(g, np.complex128).reshape(g.shape[0], -1, 2)
    m = np.zeros(g.shape[:2] + (4,), np.complex128)
    m[..., 0], m[..., 3] = g[..., 0], g[..., 1]
    return m.reshape(g.shape[0], -1)


def test_full_jones_one_step_matches_oracle():
    """ddh_one_step on the full-Jones solver: the undamped 2x2 update equals
    one oracle iteration at step size 1 (IterativeFullJonesSolver.cc:178-189,
    then SolverBase::Step with step 1), per matrix within 1e-4 (Frobenius);
    the denominators (diagonal of a sum of M^H M-type products) are positive."""
    print("case: full-Jones one step", flush=True)
    w, geom, sky, gains, P = small_problem(n_corr=4)
    s = H.Solver(P.n_antennas, P.n_solutions_per_direction, P.n_channel_blocks, **FJ)
    dp = H.DeviceProblem.from_problem(s, P)
    dp.upload(P)
    rng = np.random.default_rng(6)
    start = pairs_to_matrices(gains) * (1 + 0.05 * rng.normal(size=(gains.shape[0],
                                                                      gains.shape[1] * 2)))
    start = start.reshape(gains.shape[0], -1, 4)
    start[..., 1:3] += 0.02 * (rng.normal(size=start[..., 1:3].shape)
                               + 1j * rng.normal(size=start[..., 1:3].shape))
    start = np.ascontiguousarray(start.reshape(gains.shape[0], -1))
    nxt, den = H.one_step(s, dp, start)
    ref = start.copy()
    O.solve(P, ref, full_jones=True, max_iterations=1, step_size=1.0, accuracy=0.0,
            detect_stalling=False, fp64=True)
    d = np.linalg.norm((nxt - ref).reshape(-1, 4), axis=1)
    r = np.linalg.norm(ref.reshape(-1, 4), axis=1)
    assert np.max(d / r) <= 1e-4, np.max(d / r)
    assert den.shape == (P.n_channel_blocks, P.n_antennas * P.n_sub * 2)
    assert np.all(den > 0)
    s.close()


def batched_warm_matches_host(solver_kwargs, seed0):
    """Batched warm start (ddh_batch_solve_warm) against the per-interval host
    sequence followed by a batched solve under the same schedule: equal
    iteration counts and solutions to FP64 rounding."""
    w = S.Workload("wsx", 12, 3, 2, 1, 12, 3, 4, 60, 1e-5, seed=9)
    seqs = [WS.IntervalSequence(w, 2, seed=seed0 + k, solver_kwargs=solver_kwargs)
            for k in range(3)]
    ws = WS.WarmStart(hidden=64, seed=1, stretch=3)
    prevs = []
    for q in seqs:
        g = q.unit()
        q.solver.solve_problem(q.problems[0], g)
        prevs.append(g)
    solvers = [q.solver for q in seqs]
    dps = [q.problems[1] for q in seqs]
    host = [ws.initial(q.solver, q.problems[1], prevs[k]) for k, q in enumerate(seqs)]
    b1 = H.Batch(solvers, dps)
    b1.set_schedule(4, 1, 1, 8)
    r_host = b1.solve(host)
    dev = [p.copy() for p in prevs]
    b2 = H.Batch(solvers, dps)
    b2.set_schedule(4, 1, 1, 8)
    r_dev = b2.solve_warm(ws.net, dev, stretch=ws.stretch)
    for k in range(len(seqs)):
        assert r_dev[k].iterations == r_host[k].iterations
        np.testing.assert_allclose(dev[k], host[k], rtol=1e-9, atol=1e-12)
    b1.close()
    b2.close()
    return dev


@pytest.mark.parametrize("iteration", [H.ITERATION_STREAM, H.ITERATION_GRAM])
def test_full_jones_batched_warm_start_matches_host_sequence(iteration):
    """ddh_batch_solve_warm on full-Jones intervals: the same starts and solves
    as the per-interval host sequence one_step -> 2 x stretch iterations ->
    ddh_warmstart_features_2x2 -> forward -> ddh_warmstart_apply_2x2,
    followed by a batched solve (same schedule on both sides), with the
    streaming and the Gram iteration."""
    print(f"case: full-Jones batched warm start vs host sequence, iteration {iteration}",
          flush=True)
    batched_warm_matches_host(dict(FJ, iteration=iteration), 60)


def test_scalar_batched_warm_start_matches_host_sequence():
    """The scalar solver's warm start: each gain goes through the features as a
    pair holding it twice and starts from the mean of the two polarizations'
    corrections, so the solve runs on one gain per solution; batched equals
    the host sequence, and the solutions are finite."""
    print("case: scalar batched warm start vs host sequence", flush=True)
    dev = batched_warm_matches_host(dict(n_solution_polarizations=1), 80)
    for d in dev:
        assert d.shape[1] == 12 * 3 and np.all(np.isfinite(d))


def test_full_jones_warm_start_apply_rows():
    """ddh_warmstart_apply_2x2 scales row p of each matrix by exp(pred_p)
    (log-magnitude, phase pairs) and starts a matrix with a non-finite entry
    from the identity; the features of a 2x2 solution equal those of the
    pair of its diagonal entries."""
    print("case: full-Jones warm start apply + features", flush=True)
    rng = np.random.default_rng(3)
    n_cb, n_ant, n_sub = 2, 5, 3
    n = n_cb * n_ant * n_sub
    m = rng.normal(size=(n, 4)) + 1j * rng.normal(size=(n, 4))
    m[7, 2] = np.nan
    pred = rng.normal(scale=0.2, size=(n, 4)).astype(np.float32)
    out = H.warmstart_apply(m, pred, full_jones=True).reshape(n, 4)
    f = np.exp(pred[:, 0::2].astype(np.float64) + 1j * pred[:, 1::2].astype(np.float64))
    ref = m.copy()
    ref[7] = [1, 0, 0, 1]
    ref[:, :2] *= f[:, :1]
    ref[:, 2:] *= f[:, 1:]
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-14)
    # features from the diagonal entries
    its = [rng.normal(size=(n, 4)) + 1j * rng.normal(size=(n, 4)) for _ in range(4)]
    den = rng.uniform(0.5, 2.0, size=(n, 2))
    x4 = H.warmstart_features(*its, den, n_cb, n_ant, n_sub, full_jones=True)
    x2 = H.warmstart_features(*[a[:, ::3] for a in its], den, n_cb, n_ant, n_sub)
    np.testing.assert_array_equal(x4, x2)


def test_full_jones_warm_start_converges_sooner():
    """Consecutive full-Jones intervals: trained on the first intervals, the
    learned start (from the second stretch's iterate) reaches the same chi^2
    as the previous solutions in fewer solve iterations."""
    print("case: full-Jones warm start sequence", flush=True)
    w = S.Workload("wsfs", 12, 3, 2, 1, 12, 3, 4, 400, 1e-6, seed=7)
    seq = WS.IntervalSequence(w, 20, seed=7, solver_kwargs=FJ)
    ws = WS.WarmStart(hidden=128, seed=0)
    x, t, prev = WS.training_set(seq, ws, 0, 14, converge_iterations=400)
    loss = ws.fit(x, t, epochs=30, batch=512, lr=1e-3)
    print(f"train samples {len(x)}, last epoch loss {loss:.3e}, target var {t.var():.3e}")
    its = {"previous": [], "learned": []}
    chis = {k: [] for k in its}
    for i in range(14, 20):
        dp = seq.problems[i]
        starts = {"previous": prev.copy(), "learned": ws.initial(seq.solver, dp, prev)}
        finals = {}
        for k, sol in starts.items():
            r = seq.solver.solve_problem(dp, sol)
            its[k].append(r.iterations)
            chis[k].append(H.chi2(seq.solver, dp, sol))
            finals[k] = sol
        prev = finals["previous"]
    for k in its:
        print(f"{k:9s} mean iterations {np.mean(its[k]):6.1f}  mean chi2 {np.mean(chis[k]):.6e}")
    print(f"learned incl. feature iterations {np.mean(its['learned']) + 2 * ws.stretch + 1:.1f}")
    np.testing.assert_allclose(chis["learned"], chis["previous"], rtol=1e-3)
    assert np.mean(its["learned"]) < np.mean(its["previous"])
