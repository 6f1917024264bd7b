"""AssignAndWeight + SolveData packing on the GPU against the oracle
(ddecal/gain_solvers/SolverTools.cc:81-167, SolveData.cc:77-188).

Raw DPBuffer-order buffers with auto-correlations, uneven channel blocks,
direction-dependent solution intervals, set flags, zero weights and
non-finite values in correlations that Duo drops. The device planes must
equal, bit for bit, the oracle's weighted and packed SolveData in the device
layout (complex64), and the planes that an upload of that packed SolveData
gives (FP16: same scale rule and rounding)."""
import types

import numpy as np
import pytest

from dp3_amd import hip as H
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

N_SOL = (1, 2, 3)


def raw_interval(seed, n_t=3, n_ant=7, n_ch=11, n_dir=len(N_SOL), bad=True):
    rng = np.random.default_rng(seed)
    bl = [(a, b) for a in range(n_ant) for b in range(a, n_ant)]  # DPBuffer order, autos in
    a1 = np.array([x for x, _ in bl], np.int32)
    a2 = np.array([y for _, y in bl], np.int32)
    n_bl = len(bl)

    def cplx(*s):
        return (rng.normal(size=s) + 1j * rng.normal(size=s)).astype(np.complex64)

    data = cplx(n_t, n_bl, n_ch, 4)
    model = cplx(n_dir, n_t, n_bl, n_ch, 4)
    flags = rng.random((n_t, n_bl, n_ch, 4)) < 0.04
    weights = rng.uniform(0.0, 2.0, (n_t, n_bl, n_ch, 4)).astype(np.float32)
    weights[rng.random(weights.shape) < 0.05] = 0.0
    if bad:
        # non-finite values flag every correlation and direction of their
        # visibility, also in correlations 1 and 2, which Duo does not keep
        data[0, 3, 2, 1] = np.nan
        data[1, 9, 5, 0] = np.inf
        model[2, 2, 12, 7, 2] = complex(np.nan, 0.0)
        model[0, 0, 20, 10, 3] = complex(0.0, -np.inf)
    return data, model, flags, weights, a1, a2


def channel_blocks(n_ch, n_cb):
    edges = [(cb * n_ch) // n_cb for cb in range(n_cb + 1)]  # SolveData.cc:112-113
    return np.diff(edges).astype(np.int32)


def sol_index(n_t, n_sol=N_SOL):
    out, start = [], 0
    for n in n_sol:
        out.append([t * n // n_t + start for t in range(n_t)])  # SolveData.cc:167-168
        start += n
    return np.array(out, np.uint32)


def make_problem(data, a1, a2, n_ch, n_cb, n_corr, storage, n_sol=N_SOL):
    n_t = data.shape[0]
    n_ant = int(max(a1.max(), a2.max())) + 1
    s = H.Solver(n_ant, list(n_sol), n_cb, n_correlations=n_corr, model_storage=storage)
    cross = a1 != a2
    ch = channel_blocks(n_ch, n_cb)
    return s, H.DeviceProblem(s, n_t, a1[cross], a2[cross], ch, sol_index(n_t, n_sol))


def oracle_packed(data, model, flags, weights, a1, a2, n_cb, n_corr, n_sol=N_SOL):
    d, m = data.copy(), model.copy()
    O.assign_and_weight(d, flags, weights, m)
    n_ant = int(max(a1.max(), a2.max())) + 1
    pk = O.pack_solve_data(d, m, a1, a2, n_ant, n_cb, list(n_sol), n_corr=n_corr)
    return types.SimpleNamespace(n_channel_blocks=n_cb, n_directions=m.shape[0],
                                 vis_offsets=pk["vis_offsets"], data=pk["data"],
                                 model=pk["model"])


def device_layout(packed, cb, plane, n_t, n_bl, ch, bl_stride, n_corr):
    """One block's plane of packed SolveData ([t][bl][ch] order) in the device
    layout [t][ch][bl_stride] (zero padding)."""
    v0, v1 = int(packed.vis_offsets[cb]), int(packed.vis_offsets[cb + 1])
    n_dir = packed.n_directions
    if plane == 0:
        x = packed.data[v0:v1]
    else:
        x = packed.model[v0 * n_dir:v1 * n_dir].reshape(n_dir, v1 - v0, n_corr)[plane - 1]
    x = x.reshape(n_t, n_bl, ch, n_corr).transpose(0, 2, 1, 3)
    out = np.zeros((n_t, ch, bl_stride, n_corr), np.complex64)
    out[:, :, :n_bl] = x
    return out.reshape(n_t * ch, bl_stride, n_corr)


@pytest.mark.parametrize("n_corr", [2, 4])
def test_assign_and_weight_matches_oracle_c64(n_corr):
    print(f"case: ingest complex64 n_corr={n_corr}", flush=True)
    n_ch, n_cb = 11, 3
    data, model, flags, weights, a1, a2 = raw_interval(1)
    s, p = make_problem(data, a1, a2, n_ch, n_cb, n_corr, H.MODEL_C64)
    ms = p.assign_and_weight(data, model, a1, a2, flags=flags, weights=weights)
    assert ms >= 0.0
    ref = oracle_packed(data, model, flags, weights, a1, a2, n_cb, n_corr)
    n_t, n_bl = data.shape[0], p.n_baselines
    for cb, ch in enumerate(p.channels_per_block):
        for plane in range(model.shape[0] + 1):
            got = p.download_plane(cb, plane, n_corr=n_corr)
            want = device_layout(ref, cb, plane, n_t, n_bl, int(ch), p.bl_stride, n_corr)
            np.testing.assert_array_equal(got, want, err_msg=f"block {cb} plane {plane}")
    # the flagged visibilities are zero in every plane
    assert np.count_nonzero(ref.data == 0) > 0
    p.close()
    s.close()


@pytest.mark.parametrize("n_corr", [2, 4])
def test_assign_and_weight_f16_equals_packed_upload(n_corr):
    print(f"case: ingest FP16 n_corr={n_corr}", flush=True)
    n_ch, n_cb = 11, 3
    data, model, flags, weights, a1, a2 = raw_interval(2)
    s, p = make_problem(data, a1, a2, n_ch, n_cb, n_corr, H.MODEL_F16)
    p.assign_and_weight(data, model, a1, a2, flags=flags, weights=weights)
    ref = oracle_packed(data, model, flags, weights, a1, a2, n_cb, n_corr)
    s2, q = make_problem(data, a1, a2, n_ch, n_cb, n_corr, H.MODEL_F16)
    q.upload(ref)
    for cb in range(n_cb):
        for plane in range(model.shape[0] + 1):
            got = p.download_plane(cb, plane, n_corr=n_corr, f16=True)
            want = q.download_plane(cb, plane, n_corr=n_corr, f16=True)
            np.testing.assert_array_equal(got.view(np.uint16), want.view(np.uint16),
                                          err_msg=f"block {cb} plane {plane}")
    for x in (p, q, s, s2):
        x.close()


@pytest.mark.parametrize("n_dir", [6, 11, 16])
@pytest.mark.parametrize("storage", [H.MODEL_C64, H.MODEL_F16])
def test_assign_and_weight_direction_counts(n_dir, storage):
    """Direction counts of every register bucket (up to 4, 8, 16 directions)."""
    print(f"case: ingest n_dir={n_dir} storage={storage}", flush=True)
    n_ch, n_cb, n_corr = 7, 2, 2
    n_sol = tuple(1 + d % 2 for d in range(n_dir))
    data, model, flags, weights, a1, a2 = raw_interval(7 + n_dir, n_t=2, n_ant=5, n_ch=n_ch,
                                                       n_dir=n_dir, bad=False)
    model[n_dir - 1, 1, 3, 4, 0] = np.nan  # the last direction flags its visibility
    s, p = make_problem(data, a1, a2, n_ch, n_cb, n_corr, storage, n_sol)
    p.assign_and_weight(data, model, a1, a2, flags=flags, weights=weights)
    ref = oracle_packed(data, model, flags, weights, a1, a2, n_cb, n_corr, n_sol)
    f16 = storage == H.MODEL_F16
    if f16:
        s2, q = make_problem(data, a1, a2, n_ch, n_cb, n_corr, storage, n_sol)
        q.upload(ref)
    n_t, n_bl = data.shape[0], p.n_baselines
    for cb, ch in enumerate(p.channels_per_block):
        for plane in range(n_dir + 1):
            got = p.download_plane(cb, plane, n_corr=n_corr, f16=f16)
            if f16:
                want = q.download_plane(cb, plane, n_corr=n_corr, f16=True)
                np.testing.assert_array_equal(got.view(np.uint16), want.view(np.uint16))
            else:
                want = device_layout(ref, cb, plane, n_t, n_bl, int(ch), p.bl_stride, n_corr)
                np.testing.assert_array_equal(got, want)
    if f16:
        q.close()
        s2.close()
    p.close()
    s.close()


@pytest.mark.parametrize("storage", [H.MODEL_C64, H.MODEL_F16])
def test_assign_and_weight_single_correlation(storage):
    """Single-correlation problems (scalar solver, n_correlations 1): both Duo
    slots of every plane hold 0.5 (XX + YY) of the weighted visibility
    (SolveData.cc:29-31), bit for bit as the oracle's one-correlation packing
    (complex64) and as the upload of that packing (FP16)."""
    print(f"case: ingest single correlation storage={storage}", flush=True)
    n_ch, n_cb = 11, 3
    data, model, flags, weights, a1, a2 = raw_interval(5)
    n_t = data.shape[0]
    n_ant = int(max(a1.max(), a2.max())) + 1
    cross = a1 != a2

    def problem():
        s = H.Solver(n_ant, list(N_SOL), n_cb, n_correlations=1, n_solution_polarizations=1,
                     model_storage=storage)
        return s, H.DeviceProblem(s, n_t, a1[cross], a2[cross], channel_blocks(n_ch, n_cb),
                                  sol_index(n_t))

    s, p = problem()
    p.assign_and_weight(data, model, a1, a2, flags=flags, weights=weights)
    ref = oracle_packed(data, model, flags, weights, a1, a2, n_cb, 1)
    f16 = storage == H.MODEL_F16
    if f16:
        s2, q = problem()
        q.upload(ref)
    n_bl = p.n_baselines
    for cb, ch in enumerate(p.channels_per_block):
        for plane in range(model.shape[0] + 1):
            got = p.download_plane(cb, plane, n_corr=2, f16=f16 and plane > 0)
            if f16 and plane > 0:
                want = q.download_plane(cb, plane, n_corr=2, f16=True)
                np.testing.assert_array_equal(got.view(np.uint16), want.view(np.uint16))
            else:
                one = device_layout(ref, cb, plane, n_t, n_bl, int(ch), p.bl_stride, 1)
                np.testing.assert_array_equal(got, np.repeat(one, 2, axis=2),
                                              err_msg=f"block {cb} plane {plane}")
    if f16:
        q.close()
        s2.close()
    p.close()
    s.close()


def test_assign_and_weight_defaults_and_linear_mode():
    """No flags and unit weights; linear weighting multiplies by w itself
    (SolverTools.cc:131-136)."""
    print("case: ingest defaults and linear weighting", flush=True)
    n_ch, n_cb, n_corr = 11, 3, 2
    data, model, flags, weights, a1, a2 = raw_interval(3, bad=False)
    s, p = make_problem(data, a1, a2, n_ch, n_cb, n_corr, H.MODEL_C64)
    p.assign_and_weight(data, model, a1, a2)
    ref = oracle_packed(data, model, np.zeros_like(flags), np.ones_like(weights), a1, a2, n_cb,
                        n_corr)
    n_t, n_bl = data.shape[0], p.n_baselines
    for cb, ch in enumerate(p.channels_per_block):
        for plane in range(model.shape[0] + 1):
            got = p.download_plane(cb, plane, n_corr=n_corr)
            want = device_layout(ref, cb, plane, n_t, n_bl, int(ch), p.bl_stride, n_corr)
            np.testing.assert_array_equal(got, want)
    # linear mode with weights w equals sqrt mode with weights w^2 when w^2
    # is exact: powers of two
    w = np.exp2(np.random.default_rng(4).integers(-3, 3, weights.shape)).astype(np.float32)
    p.assign_and_weight(data, model, a1, a2, weights=w, linear_weighting=True)
    ref = oracle_packed(data, model, np.zeros_like(flags), w * w, a1, a2, n_cb, n_corr)
    for cb, ch in enumerate(p.channels_per_block):
        got = p.download_plane(cb, 1, n_corr=n_corr)
        want = device_layout(ref, cb, 1, n_t, n_bl, int(ch), p.bl_stride, n_corr)
        np.testing.assert_array_equal(got, want)
    p.close()
    s.close()


def test_assign_and_weight_then_solve_equals_packed_solve():
    """Solutions from an ingested problem equal those from the packed upload."""
    print("case: ingest then solve", flush=True)
    n_ch, n_cb, n_corr = 11, 3, 2
    data, model, flags, weights, a1, a2 = raw_interval(5, bad=False)
    s, p = make_problem(data, a1, a2, n_ch, n_cb, n_corr, H.MODEL_C64)
    p.assign_and_weight(data, model, a1, a2, flags=flags, weights=weights)
    ref = oracle_packed(data, model, flags, weights, a1, a2, n_cb, n_corr)
    s2, q = make_problem(data, a1, a2, n_ch, n_cb, n_corr, H.MODEL_C64)
    q.upload(ref)
    n_ant = int(a1.max()) + 1
    g0 = np.ones((n_cb, n_ant * sum(N_SOL) * 2), np.complex128)
    g1, g2 = g0.copy(), g0.copy()
    r1 = s.solve_problem(p, g1)
    r2 = s2.solve_problem(q, g2)
    assert r1.iterations == r2.iterations
    np.testing.assert_array_equal(g1, g2)
    for x in (p, q, s, s2):
        x.close()


def test_assign_and_weight_rejects_mismatched_baselines():
    print("case: ingest baseline mismatch", flush=True)
    n_ch, n_cb = 11, 3
    data, model, flags, weights, a1, a2 = raw_interval(6, bad=False)
    s, p = make_problem(data, a1, a2, n_ch, n_cb, 2, H.MODEL_C64)
    b1, b2 = a1.copy(), a2.copy()
    b1[[1, 2]] = b1[[2, 1]]
    b2[[1, 2]] = b2[[2, 1]]
    with pytest.raises(H.HipError, match="differ from the problem"):
        p.assign_and_weight(data, model, b1, b2)
    p.close()
    s.close()
