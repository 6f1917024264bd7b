"""Seeded synthetic calibration problems of the shapes the solver is measured on.

Geometry (random station layout, station UVW with earth rotation), sky model
(point and Gaussian components per direction), true gains and thermal noise.
The model visibilities themselves come from a predict (the HIP predict on the
GPU, or any other predict a caller supplies); this module only draws the
random quantities and assembles data = sum_d G_a M_d G_b^H + noise in the
SolveData layout of DP3 (ddecal/gain_solvers/SolveData.cc:77-188: per channel
block, visibilities ordered time -> cross baseline -> channel in block).

Choices that the workload description leaves open, fixed here:
  * station radius: lognormal with median 400 m and sigma 1, clipped at 3 km;
    uniform azimuth; array latitude 52.91 deg (LOFAR core);
  * phase centre RA 2.0 rad, Dec 0.9 rad; hour angle starts at -0.3 rad and
    advances by the earth rotation rate over 10 s integrations;
  * frequencies: n_ch channels evenly covering 120-150 MHz (channel centres);
  * direction centres: (l, m) uniform in +-2 deg; extra components of a
    direction are scattered uniformly within 0.5 deg of its centre;
  * component flux: lognormal(mu=0, sigma=1) Jy, Stokes I only;
  * Gaussian axes uniform in 5-60 arcsec (major >= minor), PA uniform [0, pi);
  * gains: amplitude uniform 1 +- 0.1, phase uniform +-pi, independent per
    (channel block, antenna, direction, polarization);
  * noise: sigma = 0.01 Jy on the real and on the imaginary part separately.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

EARTH_ROTATION = 7.2921150e-5  # rad/s
ARCSEC = np.pi / (180.0 * 3600.0)
DEG = np.pi / 180.0


@dataclass
class Workload:
    name: str
    n_stations: int
    n_directions: int
    n_points_per_direction: int
    n_gaussians_per_direction: int
    n_channels: int
    n_channel_blocks: int
    solint: int
    max_iterations: int
    tolerance: float
    step_size: float = 0.2
    seed: int = 0
    noise_sigma: float = 0.01
    n_solutions_per_direction: tuple | None = None
    f_lo: float = 120e6
    f_hi: float = 150e6
    integration_time: float = 10.0

    @property
    def n_baselines(self):
        return self.n_stations * (self.n_stations - 1) // 2

    @property
    def n_sub(self):
        return int(np.sum(self.sols_per_dir()))

    def sols_per_dir(self):
        if self.n_solutions_per_direction is None:
            return np.ones(self.n_directions, dtype=np.int32)
        return np.minimum(np.asarray(self.n_solutions_per_direction, np.int32), self.solint)


# The CPU-runnable parity shape and the LOFAR-scale shape the throughput is
# quoted on.
SMALL = Workload("24st-4dir-32ch-4cb-solint4", 24, 4, 1, 0, 32, 4, 4, 50, 1e-5)
LOFAR = Workload("62st-8dir-240ch-24cb-solint10", 62, 8, 10, 10, 240, 24, 10, 100, 1e-6)


@dataclass
class Geometry:
    ref: tuple                    # phase centre (ra, dec)
    station_enu: np.ndarray       # [n_st, 3]
    uvw: np.ndarray               # [n_t, n_st, 3] station UVW (m)
    baselines: np.ndarray         # [n_bl, 2] cross pairs (p < q)
    freq: np.ndarray              # [n_ch]
    chan_width: np.ndarray        # [n_ch]


@dataclass
class Sky:
    # one entry per component
    direction: np.ndarray   # [n_comp] direction index
    ra: np.ndarray
    dec: np.ndarray
    flux: np.ndarray        # Stokes I (Jy)
    is_gaussian: np.ndarray
    major: np.ndarray       # FWHM (rad)
    minor: np.ndarray
    pa: np.ndarray          # rad


@dataclass
class SolveProblem:
    """One solution interval in the reference SolveData layout."""
    n_antennas: int
    n_directions: int
    n_channel_blocks: int
    n_correlations: int
    n_solutions_per_direction: np.ndarray
    vis_offsets: np.ndarray   # [n_cb+1] int64
    data: np.ndarray          # [n_vis, n_corr] complex64
    model: np.ndarray         # per block [n_dir][n_vis_cb] -> [n_vis*n_dir, n_corr] c64
    ant1: np.ndarray          # [n_vis] uint32
    ant2: np.ndarray
    sol_map: np.ndarray       # per block [n_dir][n_vis_cb] -> [n_vis*n_dir] uint32
    n_times: int = 0
    n_baselines: int = 0
    channel_begin: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))

    @property
    def n_sub(self):
        return int(np.sum(self.n_solutions_per_direction))

    def unit_solutions(self):
        return np.ones((self.n_channel_blocks, self.n_antennas * self.n_sub * 2),
                       np.complex128)


def lm_to_radec(ref, l, m):
    ra0, dec0 = ref
    n = np.sqrt(1.0 - l * l - m * m)
    dec = np.arcsin(m * np.cos(dec0) + n * np.sin(dec0))
    ra = ra0 + np.arctan2(l, n * np.cos(dec0) - m * np.sin(dec0))
    return ra, dec


def make_geometry(w: Workload, rng: np.random.Generator, n_times=None, start_step=0,
                  station_enu=None) -> Geometry:
    """Station UVW for n_times integrations starting `start_step` integrations
    after the first. A given station_enu keeps the layout of another interval
    (no random draws are then made)."""
    n_times = w.solint if n_times is None else n_times
    if station_enu is None:
        r = np.minimum(np.exp(rng.normal(np.log(400.0), 1.0, w.n_stations)), 3000.0)
        az = rng.uniform(0, 2 * np.pi, w.n_stations)
        enu = np.stack([r * np.cos(az), r * np.sin(az), np.zeros(w.n_stations)], axis=1)
    else:
        enu = np.asarray(station_enu, np.float64)
    lat = 52.91 * DEG
    # ENU -> local equatorial XYZ
    X = -np.sin(lat) * enu[:, 1] + np.cos(lat) * enu[:, 2]
    Y = enu[:, 0]
    Z = np.cos(lat) * enu[:, 1] + np.sin(lat) * enu[:, 2]
    ref = (2.0, 0.9)
    dec0 = ref[1]
    uvw = np.zeros((n_times, w.n_stations, 3))
    for t in range(n_times):
        H = -0.3 + EARTH_ROTATION * w.integration_time * (start_step + t)
        sh, ch = np.sin(H), np.cos(H)
        uvw[t, :, 0] = sh * X + ch * Y
        uvw[t, :, 1] = -np.sin(dec0) * ch * X + np.sin(dec0) * sh * Y + np.cos(dec0) * Z
        uvw[t, :, 2] = np.cos(dec0) * ch * X - np.cos(dec0) * sh * Y + np.sin(dec0) * Z
    bl = np.array([(p, q) for p in range(w.n_stations) for q in range(p + 1, w.n_stations)],
                  dtype=np.int32)
    width = (w.f_hi - w.f_lo) / w.n_channels
    freq = w.f_lo + (np.arange(w.n_channels) + 0.5) * width
    return Geometry(ref, enu, uvw, bl, freq, np.full(w.n_channels, width))


def make_sky(w: Workload, geom: Geometry, rng: np.random.Generator) -> Sky:
    n_per = w.n_points_per_direction + w.n_gaussians_per_direction
    lc = rng.uniform(-2 * DEG, 2 * DEG, w.n_directions)
    mc = rng.uniform(-2 * DEG, 2 * DEG, w.n_directions)
    d_idx, ls, ms, gauss = [], [], [], []
    for d in range(w.n_directions):
        for c in range(n_per):
            if c == 0:
                l, m = lc[d], mc[d]
            else:
                rad = 0.5 * DEG * np.sqrt(rng.uniform())
                ang = rng.uniform(0, 2 * np.pi)
                l, m = lc[d] + rad * np.cos(ang), mc[d] + rad * np.sin(ang)
            d_idx.append(d)
            ls.append(l)
            ms.append(m)
            gauss.append(c >= w.n_points_per_direction)
    n = len(d_idx)
    ra, dec = lm_to_radec(geom.ref, np.array(ls), np.array(ms))
    flux = np.exp(rng.normal(0.0, 1.0, n))
    ax = np.sort(rng.uniform(5 * ARCSEC, 60 * ARCSEC, (n, 2)), axis=1)
    pa = rng.uniform(0, np.pi, n)
    g = np.array(gauss, dtype=bool)
    return Sky(np.array(d_idx, np.int32), ra, dec, flux, g,
               np.where(g, ax[:, 1], 0.0), np.where(g, ax[:, 0], 0.0), np.where(g, pa, 0.0))


def channel_begin(n_ch, n_cb):
    """Block edges (SolveData.cc:112-113)."""
    return np.array([(cb * n_ch) // n_cb for cb in range(n_cb + 1)], dtype=np.int64)


def true_gains(w: Workload, rng: np.random.Generator, n_sub=None):
    n_sub = w.n_sub if n_sub is None else n_sub
    amp = rng.uniform(0.9, 1.1, (w.n_channel_blocks, w.n_stations, n_sub, 2))
    ph = rng.uniform(-np.pi, np.pi, (w.n_channel_blocks, w.n_stations, n_sub, 2))
    return (amp * np.exp(1j * ph)).reshape(w.n_channel_blocks, -1)


def pack_block(vis, c0, c1):
    """vis [n_t, n_bl, n_ch, ncorr] -> [(n_t*n_bl*(c1-c0)), ncorr] in t->bl->ch order."""
    x = vis[:, :, c0:c1, :]
    return np.ascontiguousarray(x).reshape(-1, vis.shape[-1])


def assemble(w: Workload, geom: Geometry, model_vis: np.ndarray, gains: np.ndarray,
             rng: np.random.Generator, noise_sigma=None, n_corr=2) -> SolveProblem:
    """Builds the SolveData-layout problem from per-direction Stokes-I model
    visibilities model_vis [n_dir, n_t, n_bl, n_ch] (complex), gains
    [n_cb, n_ant*n_sub*2] and thermal noise. XX = YY = I (OnePredict.cc:440-452);
    the Duo form keeps XX and YY (SolveData.cc:25-27)."""
    sigma = w.noise_sigma if noise_sigma is None else noise_sigma
    n_dir, n_t, n_bl, n_ch = model_vis.shape
    n_ant = w.n_stations
    nsd = w.sols_per_dir()
    n_sub = int(nsd.sum())
    start = np.concatenate([[0], np.cumsum(nsd)[:-1]])
    cbe = channel_begin(n_ch, w.n_channel_blocks)
    p = geom.baselines[:, 0]
    q = geom.baselines[:, 1]
    g = gains.reshape(w.n_channel_blocks, n_ant, n_sub, 2)
    vis_off = [0]
    datas, models, a1s, a2s, sms = [], [], [], [], []
    for cb in range(w.n_channel_blocks):
        c0, c1 = cbe[cb], cbe[cb + 1]
        nc = c1 - c0
        m_cb = model_vis[:, :, :, c0:c1].astype(np.complex64)          # [d, t, bl, ch]
        sol_t = (np.arange(n_t)[None, :] * nsd[:, None]) // n_t + start[:, None]  # [d, t]
        data = np.zeros((n_t, n_bl, nc, 2), np.complex64)
        for d in range(n_dir):
            for pol in range(2):
                ga = g[cb][:, sol_t[d], pol]                          # [ant, t]
                gap = ga[p].T.astype(np.complex64)                    # [t, bl]
                gbq = np.conj(ga[q].T).astype(np.complex64)
                data[..., pol] += (gap[:, :, None] * m_cb[d]) * gbq[:, :, None]
        if sigma > 0:
            data += (rng.normal(0, sigma, data.shape) +
                     1j * rng.normal(0, sigma, data.shape)).astype(np.complex64)
        if n_corr == 2:
            dvis = data
            mvis = np.stack([m_cb, m_cb], axis=-1)
        else:
            dvis = np.zeros((n_t, n_bl, nc, 4), np.complex64)
            dvis[..., 0] = data[..., 0]
            dvis[..., 3] = data[..., 1]
            mvis = np.zeros((n_dir, n_t, n_bl, nc, 4), np.complex64)
            mvis[..., 0] = m_cb
            mvis[..., 3] = m_cb
        datas.append(dvis.reshape(-1, n_corr))
        models.append(mvis.reshape(-1, n_corr))
        nv = n_t * n_bl * nc
        a1s.append(np.repeat(np.tile(p, n_t), nc).astype(np.uint32))
        a2s.append(np.repeat(np.tile(q, n_t), nc).astype(np.uint32))
        sm = np.repeat(sol_t, n_bl * nc, axis=1).astype(np.uint32)  # [d, nv]
        sms.append(sm.reshape(-1))
        vis_off.append(vis_off[-1] + nv)
    return SolveProblem(n_ant, n_dir, w.n_channel_blocks, n_corr, nsd.astype(np.int32),
                        np.array(vis_off, np.int64), np.concatenate(datas),
                        np.concatenate(models), np.concatenate(a1s), np.concatenate(a2s),
                        np.concatenate(sms), n_t, n_bl, cbe)


def problem_from_solvertester(data, model, n_sol_per_dir, n_ant, n_cb, packed):
    """Wraps packed SolveData arrays (from the oracle's packer) as a SolveProblem."""
    n_dir = model.shape[0]
    return SolveProblem(n_ant, n_dir, n_cb, packed["data"].shape[1],
                        np.asarray(n_sol_per_dir, np.int32), packed["vis_offsets"],
                        packed["data"], packed["model"], packed["ant1"], packed["ant2"],
                        packed["sol_map"], data.shape[0], data.shape[1])


def download_problem(dp, w: Workload, baselines=None):
    """The SolveData-layout copy (SolveProblem) of a device problem built for
    workload w with the regular baseline order (a < b), complex64 planes:
    each plane is read back and reordered time -> baseline -> channel."""
    cbe = channel_begin(w.n_channels, w.n_channel_blocks)
    n_t, n_bl = dp.n_times, dp.n_baselines
    if baselines is None:
        baselines = np.array([(a, b) for a in range(w.n_stations)
                              for b in range(a + 1, w.n_stations)])
    datas, models, a1s, a2s, sms, offs = [], [], [], [], [], [0]
    spd = w.sols_per_dir()
    start = np.concatenate([[0], np.cumsum(spd)[:-1]])
    for cb in range(w.n_channel_blocks):
        nc = int(cbe[cb + 1] - cbe[cb])
        d = dp.download_plane(cb, 0).reshape(n_t, nc, dp.bl_stride, 2)[:, :, :n_bl]
        datas.append(np.ascontiguousarray(d.transpose(0, 2, 1, 3)).reshape(-1, 2))
        for di in range(w.n_directions):
            m = dp.download_plane(cb, di + 1).reshape(n_t, nc, dp.bl_stride, 2)[:, :, :n_bl]
            models.append(np.ascontiguousarray(m.transpose(0, 2, 1, 3)).reshape(-1, 2))
        nv = n_t * n_bl * nc
        offs.append(offs[-1] + nv)
        t_of_vis = np.repeat(np.arange(n_t), n_bl * nc)
        sms.append(np.concatenate([(t_of_vis * spd[di]) // n_t + start[di]
                                   for di in range(w.n_directions)]).astype(np.uint32))
        a1s.append(np.repeat(np.tile(baselines[:, 0], n_t), nc).astype(np.uint32))
        a2s.append(np.repeat(np.tile(baselines[:, 1], n_t), nc).astype(np.uint32))
    return SolveProblem(w.n_stations, w.n_directions, w.n_channel_blocks, 2, spd,
                        np.array(offs, np.int64), np.concatenate(datas), np.concatenate(models),
                        np.concatenate(a1s), np.concatenate(a2s), np.concatenate(sms), n_t, n_bl,
                        cbe)


def phase_referenced(solutions, n_ant, n_sub):
    """Solutions [n_cb, n_ant*n_sub*2] -> g * conj(g_ant0) / |g_ant0| per
    (cb, sub-solution, pol), removing the unconstrained common phase."""
    s = solutions.reshape(solutions.shape[0], n_ant, n_sub, 2)
    ref = s[:, 0:1, :, :]
    return s * np.conj(ref) / np.abs(ref)
