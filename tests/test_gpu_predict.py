"""GPU predict (point + Gaussian components) against the CPU oracle's
restatement of Simulator::visit, in the solver's device layout.

The device predict is FP64 throughout and rounds each direction's sum once,
as Simulator.cc accumulates in double and OnePredict casts to complex64. So
every complex64 model value must lie within half an ulp (2^-24 relative, per
real and imaginary part) of the oracle's FP64 sum, up to FP64-level
differences of the transcendentals (slack 1e-12 of the direction's flux).
The smearing factors are sinc of a float argument in float arithmetic, as the
reference's sinc(float) (Simulator.cc:472-474); the device and host sinf may
differ by an ulp or two, which the smeared tests allow (2^-21 of the
direction's flux)."""
import numpy as np
import pytest

from dp3_amd import hip as H
from dp3_amd import synth as S
from oracle import pyoracle as O
from problems import oracle_components, oracle_model

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


def within_one_rounding(got, ref, slack):
    """|got - ref| per real and imaginary part within half an ulp of complex64
    (2^-24 relative) plus `slack`; returns the worst excess ratio."""
    worst = 0.0
    for g, r in ((got.real, ref.real), (got.imag, ref.imag)):
        bound = 2.0 ** -24 * np.abs(r) + slack
        worst = max(worst, float(np.max(np.abs(g - r) / bound)))
    return worst


def test_predict_matches_oracle_duo(lofar_like):
    w, geom, sky = lofar_like
    print("case: predict duo c64", flush=True)
    ref = oracle_model(w, geom, sky)  # [d, t, bl, ch]
    s, dp, _ = device_problem(w, geom)
    dp.predict(geom.ref, geom.uvw, geom.freq, geom.chan_width, H.components_from_sky(sky))
    got = device_model(dp, w)
    for d in range(w.n_directions):
        flux = sky.flux[sky.direction == d].sum()
        worst = within_one_rounding(got[d, ..., 0], ref[d], 1e-12 * flux)
        assert worst <= 1.0, (d, worst)
        np.testing.assert_array_equal(got[d, ..., 0], got[d, ..., 1])  # XX = YY = I


def tsimulator_problem():
    """The fixture of tSimulator.cc:22-60: 4 stations with UVW (5000 st,
    1000 st, 0) m, 2 channels at 130 and 131 MHz of 1 MHz, phase centre
    (0.5, 0.1), a 1-Jy Stokes-I point at (+0.02, +0.02) rad, one time step."""
    n_st = 4
    bl = np.array([(a, b) for a in range(n_st - 1) for b in range(a + 1, n_st)], np.int32)
    uvw = np.zeros((1, n_st, 3))
    uvw[0, :, 0] = np.arange(n_st) * 5000.0
    uvw[0, :, 1] = np.arange(n_st) * 1000.0
    freq = np.array([130e6, 131e6])
    width = np.full(2, 1e6)
    s = H.Solver(n_st, [1], 1)
    dp = H.DeviceProblem(s, 1, bl[:, 0], bl[:, 1], np.array([2], np.int32),
                         np.zeros((1, 1), np.uint32))
    comps = (H.Component * 1)()
    comps[0].ra, comps[0].dec, comps[0].I = 0.52, 0.12, 1.0
    comps[0].direction = 0
    return s, dp, uvw, freq, width, comps, bl


def test_predict_smearing_known_answers():
    """tSimulator.cc:78-155 through the HIP predict: |V| = 1 without smearing,
    0.759154 with frequency smearing and 0.893578 / 0.891990 with time
    smearing (scaled_ncp_uvw = (0, 0.3, 0)) on baseline 3 (stations 1-2),
    BOOST_CHECK_CLOSE 1e-3 percent."""
    print("case: predict tSimulator known answers", flush=True)
    s, dp, uvw, freq, width, comps, bl = tsimulator_problem()
    ref = (0.5, 0.1)

    def magnitudes():
        pl = dp.download_plane(0, 1).reshape(1, 2, dp.bl_stride, 2)[0, :, 3, 0]
        return np.abs(pl)

    dp.predict(ref, uvw, freq, width, comps)
    np.testing.assert_allclose(magnitudes(), [1.0, 1.0], rtol=1e-5)
    dp.predict(ref, uvw, freq, width, comps, correct_freq_smearing=True)
    np.testing.assert_allclose(magnitudes(), [0.759154, 0.759154], rtol=1e-5)
    dp.predict(ref, uvw, freq, width, comps, scaled_ncp_uvw=[0.0, 0.3, 0.0])
    np.testing.assert_allclose(magnitudes(), [0.893578, 0.891990], rtol=1e-5)
    dp.close()
    s.close()


def test_predict_time_and_freq_smearing_matches_oracle(lofar_like):
    """Points and Gaussians with time and frequency smearing against the
    oracle's Simulator restatement, per time step, with a LOFAR-like scaled
    NCP vector (angular speed over a 10 s integration at the phase centre's
    declination, FastPredict.cc:447-462)."""
    w, geom, sky = lofar_like
    print("case: predict time+freq smearing", flush=True)
    omega = 2 * np.pi * w.integration_time / 86164.0905
    ncp = [0.0, omega * np.cos(geom.ref[1]), omega * np.sin(geom.ref[1])]
    s, dp, _ = device_problem(w, geom)
    dp.predict(geom.ref, geom.uvw, geom.freq, geom.chan_width, H.components_from_sky(sky),
               correct_freq_smearing=True, scaled_ncp_uvw=ncp)
    got = device_model(dp, w)
    plain = np.abs(oracle_model(w, geom, sky))
    smeared_any = False
    for d in range(w.n_directions):
        comps = oracle_components(sky, d)
        flux = sky.flux[sky.direction == d].sum()
        for t in range(geom.uvw.shape[0]):
            ref = O.simulate(geom.ref, geom.uvw[t], geom.baselines, geom.freq, geom.chan_width,
                             comps, correct_freq_smearing=True, correct_time_smearing=True,
                             scaled_ncp_uvw=ncp)[:, :, 0]
            worst = within_one_rounding(got[d, t, :, :, 0], ref, 2.0 ** -21 * flux)
            assert worst <= 1.0, (d, t, worst)
            smeared_any = smeared_any or bool(np.any(np.abs(ref) < plain[d, t] * (1 - 1e-9)))
    assert smeared_any  # the factors changed the model


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
            worst = within_one_rounding(got[d, t], ref, 2.0 ** -21 * scale)
            assert worst <= 1.0, (d, t, worst)


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
