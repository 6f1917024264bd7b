"""GPU predict (point + Gaussian components) against the CPU oracle's
restatement of Simulator::visit, in the solver's device layout."""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from oracle import pyoracle as O
from problems import oracle_model

pytestmark = pytest.mark.gpu


def device_problem(w, geom, n_corr=2, storage=H.MODEL_C64):
    s = H.Solver(w.n_stations, w.sols_per_dir(), w.n_channel_blocks, n_correlations=n_corr,
                 model_storage=storage)
    ch = np.diff(S.channel_begin(w.n_channels, w.n_channel_blocks))
    sol_index = np.zeros((w.n_directions, geom.uvw.shape[0]), np.uint32)
    start = np.concatenate([[0], np.cumsum(w.sols_per_dir())[:-1]])
    for d in range(w.n_directions):
        sol_index[d] = (np.arange(geom.uvw.shape[0]) * w.sols_per_dir()[d]) // geom.uvw.shape[0] + start[d]
    dp = H.DeviceProblem(s, geom.uvw.shape[0], geom.baselines[:, 0], geom.baselines[:, 1], ch,
                         sol_index)
    return s, dp, sol_index


def device_model(dp, w, n_corr=2, f16=False, scale=None):
    """Model planes back in [n_dir, n_t, n_bl, n_ch, n_corr]."""
    cbe = S.channel_begin(w.n_channels, w.n_channel_blocks)
    out = np.zeros((w.n_directions, dp.n_times, dp.n_baselines, w.n_channels, n_corr), np.complex128)
    for cb in range(w.n_channel_blocks):
        nc = cbe[cb + 1] - cbe[cb]
        for d in range(w.n_directions):
            pl = dp.download_plane(cb, d + 1, n_corr=n_corr, f16=f16)
            if f16:
                pl = (pl[..., 0].astype(np.float64) + 1j * pl[..., 1].astype(np.float64)) * scale[cb, d]
            pl = pl.reshape(dp.n_times, nc, dp.bl_stride, n_corr)[:, :, : dp.n_baselines]
            out[d, :, :, cbe[cb]:cbe[cb + 1]] = pl.transpose(0, 2, 1, 3)
    return out


@pytest.fixture(scope="module")
def lofar_like():
    w = S.Workload("pred", 12, 3, 3, 3, 20, 4, 3, 10, 1e-5)
    rng = np.random.default_rng(7)
    geom = S.make_geometry(w, rng)
    sky = S.make_sky(w, geom, rng)
    return w, geom, sky


def test_predict_matches_oracle_duo(lofar_like):
    w, geom, sky = lofar_like
    print("case: predict duo c64", flush=True)
    ref = oracle_model(w, geom, sky)  # [d, t, bl, ch]
    s, dp, _ = device_problem(w, geom)
    dp.predict(geom.ref, geom.uvw, geom.freq, geom.chan_width, H.components_from_sky(sky))
    got = device_model(dp, w)
    for d in range(w.n_directions):
        flux = sky.flux[sky.direction == d].sum()
        err = np.abs(got[d, ..., 0] - ref[d]).max()
        assert err <= 2e-6 * flux, (d, err, flux)
        np.testing.assert_array_equal(got[d, ..., 0], got[d, ..., 1])  # XX = YY = I


def test_predict_full_stokes_spectra_and_smearing():
    """Polarized components with spectral terms, a rotation measure, absolute
    position angles and frequency smearing, 4 correlations."""
    print("case: predict full stokes", flush=True)
    w = S.Workload("predfull", 8, 2, 2, 2, 12, 3, 2, 10, 1e-5)
    rng = np.random.default_rng(9)
    geom = S.make_geometry(w, rng)
    sky = S.make_sky(w, geom, rng)
    comps = H.components_from_sky(sky)
    oc = [[] for _ in range(w.n_directions)]
    for i in range(len(comps)):
        c = comps[i]
        c.Q, c.U, c.V = 0.1 * c.I, -0.05 * c.I, 0.02 * c.I
        c.ref_freq = 140e6
        c.n_spectral_terms = 2
        c.log_si = int(i % 2)
        c.spectral_terms[0], c.spectral_terms[1] = -0.7, 0.05
        if i % 3 == 0:
            c.has_rm, c.pol_fraction, c.pol_angle, c.rm = 1, 0.2, 0.3, 4.0
        c.pa_is_absolute = int(i % 2)
        oc[c.direction].append(O.make_component(
            c.ra, c.dec, c.I, c.Q, c.U, c.V, gaussian=bool(c.is_gaussian), major=c.major_axis,
            minor=c.minor_axis, pa=c.position_angle, pa_absolute=bool(c.pa_is_absolute),
            spectral_terms=[-0.7, 0.05], ref_freq=140e6, log_si=bool(c.log_si),
            rm=(0.2, 0.3, 4.0) if c.has_rm else None))
    s, dp, _ = device_problem(w, geom, n_corr=4)
    dp.predict(geom.ref, geom.uvw, geom.freq, geom.chan_width, comps, stokes_i_only=False,
               correct_freq_smearing=True)
    got = device_model(dp, w, n_corr=4)
    for d in range(w.n_directions):
        scale = sum(abs(c.I) for c in comps if c.direction == d) * 2
        for t in range(geom.uvw.shape[0]):
            ref = O.simulate(geom.ref, geom.uvw[t], geom.baselines, geom.freq, geom.chan_width,
                             oc[d], stokes_i_only=False, correct_freq_smearing=True)
            err = np.abs(got[d, t] - ref).max()
            assert err <= 2e-6 * scale, (d, t, err)


def test_predict_fp16_model(lofar_like):
    w, geom, sky = lofar_like
    print("case: predict fp16", flush=True)
    ref = oracle_model(w, geom, sky)
    s, dp, _ = device_problem(w, geom, storage=H.MODEL_F16)
    dp.predict(geom.ref, geom.uvw, geom.freq, geom.chan_width, H.components_from_sky(sky))
    # the scale bound: sum of |flux| per direction / 2^14
    scale = np.zeros((w.n_channel_blocks, w.n_directions))
    for d in range(w.n_directions):
        scale[:, d] = np.float32(np.float32(np.abs(sky.flux[sky.direction == d]).astype(np.float32).sum()) / 16384.0)
    got = device_model(dp, w, f16=True, scale=scale)
    for d in range(w.n_directions):
        flux = sky.flux[sky.direction == d].sum()
        err = np.abs(got[d, ..., 0] - ref[d]).max()
        # FP16 keeps 11 significant bits: |err| <= 2^-11 * max|V| <= 2^-11 * flux
        assert err <= 2.0 ** -11 * flux * 1.01, (d, err, flux)


def test_synthesize_data_matches_host_assembly(lofar_like):
    w, geom, sky = lofar_like
    print("case: synthesize data", flush=True)
    s, dp, sol_index = device_problem(w, geom)
    dp.predict(geom.ref, geom.uvw, geom.freq, geom.chan_width, H.components_from_sky(sky))
    model = device_model(dp, w)[..., 0]
    rng = np.random.default_rng(1)
    gains = S.true_gains(w, rng)
    dp.synthesize_data(gains, sol_index, 0.0)
    P = S.assemble(w, geom, model, gains, rng, noise_sigma=0.0)
    cbe = S.channel_begin(w.n_channels, w.n_channel_blocks)
    for cb in range(w.n_channel_blocks):
        nc = cbe[cb + 1] - cbe[cb]
        pl = dp.download_plane(cb, 0).reshape(dp.n_times, nc, dp.bl_stride, 2)[:, :, :dp.n_baselines]
        ref = P.data[P.vis_offsets[cb]:P.vis_offsets[cb + 1]].reshape(dp.n_times, dp.n_baselines, nc, 2)
        np.testing.assert_allclose(pl.transpose(0, 2, 1, 3), ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    # noise statistics
    dp.synthesize_data(np.zeros_like(gains), sol_index, 0.25, seed=3)
    pl = dp.download_plane(0, 0).reshape(dp.n_times, -1, dp.bl_stride, 2)[:, :, :dp.n_baselines]
    x = np.concatenate([pl.real.ravel(), pl.imag.ravel()])
    assert abs(x.mean()) < 0.01 and abs(x.std() - 0.25) < 0.01
