// Point / Gaussian component predict, written directly into a problem's
// device model planes, plus the on-device synthesis of data visibilities
// (sum of gain-corrupted model planes + Gaussian noise) used to build
// synthetic solution intervals.
//
// Arithmetic follows base/Simulator.cc (lofar-astron/DP3 6.6.0):
//   radec2lmn             Simulator.h:35-60
//   station phases        Simulator.cc:477-524 (phase = 2 pi/c (u l + v m + w (n-1)) f)
//   Stokes spectrum       Simulator.cc:527-549, PointSource.cc:51-99
//   point visibility      Simulator.cc:157-281: shift_q conj(shift_p) S
//   Gaussian envelope     Simulator.cc:296-405: exp(-2 pi^2 (u'^2 + v'^2) f^2/c^2),
//                         (u', v') rotated by pi/2 + PA + pi and scaled by
//                         FWHM / (2 sqrt(2 ln 2)); absolute PA rotates the
//                         station UVW with the Euler matrices of
//                         steps/PhaseShift.cc:197-212
//   frequency smearing    sinc(0.5 (phi_q - phi_p) width), Simulator.cc:190-199
//   time smearing         sinc(0.5 (rot_q - rot_p) f), Simulator.cc:177-188, with the
//                         station earth-rotation phase of Simulator.cc:488-503
//                         (scaled_ncp_uvw[0] taken as zero, as there)
//   Stokes-I copy         steps/OnePredict.cc:440-452 (XX = YY = I, XY = YX = 0)
// As in Simulator.cc, everything is FP64 (station phasors, Gaussian envelope,
// per-baseline products and the sum over a direction's components) and the
// sum is rounded once to the model storage; sinc takes a float argument, as
// the reference's sinc(float) does (Simulator.cc:472-474).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/ddecal_hip.h"
#include "common.h"
#include "solver_kernels.h"


namespace ddh {
namespace {

constexpr double kSpeedOfLight = 299792458.0;

struct CompGeom {
  double l, m, nm1;                // lmn, n - 1
  double cos_phi, sin_phi;         // Gaussian rotation
  double u_scale, v_scale;         // FWHM -> sigma
  double rot[9];                   // Euler rotation (identity unless absolute PA)
  int32_t is_gaussian;
  int32_t direction;
};

__device__ void Radec2Lmn(double ra0, double dec0, double ra, double dec, double* lmn) {
  double sdr, cdr, sd, cd, sd0, cd0;
  sincos(ra - ra0, &sdr, &cdr);
  sincos(dec, &sd, &cd);
  sincos(dec0, &sd0, &cd0);
  lmn[0] = cd * sdr;
  lmn[1] = sd * cd0 - cd * sd0 * cdr;
  lmn[2] = sd * sd0 + cd * cd0 * cdr;
}

__device__ void Euler(double ra, double dec, double m[3][3]) {
  double sra, cra, sdec, cdec;
  sincos(ra, &sra, &cra);
  sincos(dec, &sdec, &cdec);
  m[0][0] = cra;
  m[1][0] = -sra;
  m[2][0] = 0;
  m[0][1] = -sra * sdec;
  m[1][1] = -cra * sdec;
  m[2][1] = cdec;
  m[0][2] = sra * cdec;
  m[1][2] = cra * cdec;
  m[2][2] = sdec;
}

// One thread per component: lmn and Gaussian constants.
__global__ void ComponentGeometryKernel(const ddh_component* comps, int n_comp, double ra0,
                                        double dec0, CompGeom* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_comp) return;
  const ddh_component& s = comps[c];
  CompGeom g{};
  double lmn[3];
  Radec2Lmn(ra0, dec0, s.ra, s.dec, lmn);
  g.l = lmn[0];
  g.m = lmn[1];
  g.nm1 = lmn[2] - 1.0;
  g.is_gaussian = s.is_gaussian;
  g.direction = s.direction;
  for (int i = 0; i < 9; ++i) g.rot[i] = (i % 4 == 0) ? 1.0 : 0.0;
  if (s.is_gaussian) {
    const double phi = M_PI_2 + s.position_angle + M_PI;
    double sp, cp;
    sincos(phi, &sp, &cp);
    g.cos_phi = cp;
    g.sin_phi = sp;
    const double fwhm2sigma = 1.0 / (2.0 * sqrt(2.0 * log(2.0)));
    g.u_scale = s.major_axis * fwhm2sigma;
    g.v_scale = s.minor_axis * fwhm2sigma;
    if (s.pa_is_absolute) {
      double ep[3][3], es[3][3];
      Euler(ra0, dec0, ep);
      Euler(s.ra, s.dec, es);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double acc = 0.0;
          for (int k = 0; k < 3; ++k) acc += es[k][i] * ep[k][j];
          g.rot[i * 3 + j] = acc;
        }
    }
  }
  out[c] = g;
}

// Stokes (I, Q, U, V) per (component, channel) (PointSource.cc:51-99), FP64.
struct Stokes4 {
  double I, Q, U, V;
};
__global__ void SpectrumKernel(const ddh_component* comps, int n_comp, const double* freq,
                               int n_ch, Stokes4* spec) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_comp * n_ch) return;
  const int c = idx / n_ch;
  const int ch = idx - c * n_ch;
  const ddh_component& s = comps[c];
  const double f = freq[ch];
  double I = s.I, Q = s.Q, U = s.U, V = s.V;
  const int nt = min(s.n_spectral_terms, 8);
  if (nt > 0) {
    if (s.log_si) {
      const double base = log10(f) - log10(s.ref_freq);
      double e = 0.0;
      for (int i = nt - 1; i >= 0; --i) e = e * base + s.spectral_terms[i];
      I *= pow(10.0, base * e);
    } else {
      const double x = f / s.ref_freq - 1.0;
      double val = 0.0;
      for (int i = nt - 1; i >= 0; --i) val = val * x + s.spectral_terms[i];
      I += val * x;
    }
  }
  if (s.has_rm) {
    const double lambda = kSpeedOfLight / f;
    const double chi = 2.0 * (s.pol_angle + s.rm * lambda * lambda);
    const double qu = I * s.pol_fraction;
    double sc, cc;
    sincos(chi, &sc, &cc);
    Q = qu * cc;
    U = qu * sc;
  }
  spec[idx] = Stokes4{I, Q, U, V};
}

__device__ __forceinline__ float Sinc(float x) { return fabsf(x) < 1e-6f ? 1.0f : sinf(x) / x; }

typedef _Float16 half2_v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t HalfBits(float re, float im) {
  half2_v h;
  h.x = static_cast<_Float16>(re);
  h.y = static_cast<_Float16>(im);
  return __builtin_bit_cast(uint32_t, h);
}

// ---- the per-visibility predict ------------------------------------------
// One workgroup per (channel, time step, direction, baseline chunk of
// kPredictThreads * kPredictBaselines): each thread sums the direction's
// components for up to kPredictBaselines baselines in FP64 registers and
// rounds the sum once into the model plane (Simulator.cc accumulates in
// double, OnePredict casts to complex64).
//
// A direction's components come points first, then Gaussians (the host
// orders them so; the sum is the same up to FP64 rounding), and are taken in
// chunks that fit the LDS. For a chunk the workgroup first forms, per
// (component, station), the terms every baseline of the station shares:
//   phase   phi_s = 2 pi / c (u l + v m + w (n - 1))           (Simulator.cc:487-488)
//   shift   exp(i phi_s f) (cos, sin)                           (:510-521)
//   rot_s   the earth-rotation phase of time smearing           (:496-503)
//   uv'_s   the station's (u, v) rotated to the source (absolute position
//           angle) and by the Gaussian's position angle, scaled by the axes
//           (:312-342, :360-366)
// and the components' spectra at the channel, so a baseline's Gaussian
// (u', v') is uv'_q - uv'_p (the rotation and scaling are linear; the
// reference forms u_q - u_p first, which differs by FP64 rounding only).
// Then every thread runs over the chunk's points and Gaussians for each of
// its baselines: shift_q conj(shift_p) times the spectrum, times the
// envelope exp(f^2 / c^2 (-2 pi^2 (u'^2 + v'^2))) and the smearing factors,
// with no per-component branch.
// ---------------------------------------------------------------------------
constexpr int kPredictThreads = 512;
constexpr int kPredictBaselines = 4;  // baselines per thread
constexpr size_t kPredictLds = 48 * 1024;

// Per global channel: where its rows start in the model planes of its block.
struct PredictChannel {
  int64_t base;      // model_off(cb) + ch_in_block * bl_stride (visibilities)
  int64_t plane;     // visibilities of one (block, direction) plane
  int64_t t_stride;  // n_chb * bl_stride
  int32_t cb;
  int32_t pad;
};

struct PredictArgs {
  const double* uvw;           // [n_t][n_st][3]
  const CompGeom* geom;        // [n_comp], by direction, points before Gaussians
  const Stokes4* spec;         // [n_comp][n_ch_total]
  const int32_t* c_off;        // [n_dir + 1] component ranges per direction
  const int32_t* g_off;        // [n_dir] first Gaussian of each direction
  const double* freq;          // [n_ch_total]
  const double* width;         // [n_ch_total]
  const int32_t* ant1;         // [n_bl]
  const int32_t* ant2;
  const PredictChannel* chan;  // [n_ch_total]
  const float* scale;          // [n_cb][n_dir] (FP16 planes)
  void* model;
  int32_t n_st, n_bl, bl_stride, n_t, n_dir, n_ch_total, chunk;
  int32_t freq_smear, time_smear;
  double ncp_v, ncp_w;
};

// LDS doubles per (component, station) of a chunk: shift (2) and Gaussian
// uv' (2), with smearing also rot and phase; then 4 per component (Stokes).
constexpr int PredictWords(bool smear) { return smear ? 6 : 4; }

// SI: Stokes I only (XX = YY = I, XY = YX = 0); SMEAR: time or frequency
// smearing on.
template <int NC, bool F16, bool SI, bool SMEAR>
__device__ __forceinline__ void PredictBody(const PredictArgs& a) {
  extern __shared__ double s_st[];
  constexpr int W = PredictWords(SMEAR);
  const int ch = blockIdx.x;
  const int t = blockIdx.y / a.n_dir;
  const int d = blockIdx.y - t * a.n_dir;
  const int bl0 = blockIdx.z * (kPredictThreads * kPredictBaselines);
  const PredictChannel pc = a.chan[ch];
  const int c_begin = a.c_off[d], c_end = a.c_off[d + 1], g_begin = a.g_off[d];
  const double f = a.freq[ch];
  const double f2c = f * f * (1.0 / (kSpeedOfLight * kSpeedOfLight));
  const double width = (SMEAR && a.freq_smear) ? a.width[ch] : 0.0;
  const int n_st = a.n_st;
  constexpr double cinv = 2.0 * M_PI / kSpeedOfLight;
  constexpr int NV = SI ? 1 : (NC == 2 ? 2 : 4);  // sums kept: I, (XX, YY) or all four
  double2 v[kPredictBaselines][NV];
  int pq[kPredictBaselines][2];
#pragma unroll
  for (int j = 0; j < kPredictBaselines; ++j) {
    const int bl = bl0 + int(threadIdx.x) + j * kPredictThreads;
    const bool ok = bl < a.n_bl;
    pq[j][0] = ok ? a.ant1[bl] : 0;
    pq[j][1] = ok ? a.ant2[bl] : 0;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[j][k] = make_double2(0.0, 0.0);
  }
  const double* uvw_t = a.uvw + size_t(t) * n_st * 3;
  for (int c0 = c_begin; c0 < c_end; c0 += a.chunk) {
    const int nc = min(a.chunk, c_end - c0);
    const int np = min(max(g_begin - c0, 0), nc);  // the chunk's points
    double* s_spec = s_st + size_t(nc) * W * n_st;
    __syncthreads();  // the previous chunk's terms are no longer read
    for (int e = threadIdx.x; e < nc * n_st; e += kPredictThreads) {
      const int ci = e / n_st;
      const int st = e - ci * n_st;
      const CompGeom& g = a.geom[c0 + ci];
      const double u = uvw_t[st * 3], vv = uvw_t[st * 3 + 1], w = uvw_t[st * 3 + 2];
      const double phi = cinv * (u * g.l + vv * g.m + w * g.nm1);
      double sn, cs;
      sincos(phi * f, &sn, &cs);
      double* row = s_st + size_t(ci) * W * n_st + st;
      row[0] = cs;
      row[n_st] = sn;
      if (ci >= np) {
        const double ur = g.rot[0] * u + g.rot[1] * vv + g.rot[2] * w;
        const double vr = g.rot[3] * u + g.rot[4] * vv + g.rot[5] * w;
        row[2 * n_st] = g.u_scale * (ur * g.cos_phi - vr * g.sin_phi);
        row[3 * n_st] = g.v_scale * (ur * g.sin_phi + vr * g.cos_phi);
      }
      if constexpr (SMEAR) {
        row[4 * n_st] = cinv * (g.l * (vv * a.ncp_w - w * a.ncp_v) + g.m * (-u * a.ncp_w) +
                                g.nm1 * (u * a.ncp_v));
        row[5 * n_st] = phi;
      }
    }
    for (int e = threadIdx.x; e < nc; e += kPredictThreads) {
      const Stokes4 s = a.spec[size_t(c0 + e) * a.n_ch_total + ch];
      s_spec[4 * e] = s.I;
      s_spec[4 * e + 1] = s.Q;
      s_spec[4 * e + 2] = s.U;
      s_spec[4 * e + 3] = s.V;
    }
    __syncthreads();
    // one component's contribution to baseline j: shift_q conj(shift_p)
    // (Simulator.cc:212-213) times the amplitude and the spectrum
    auto add = [&](const int j, const int ci, const double amp_in, const bool gauss) {
      const int p = pq[j][0], q = pq[j][1];
      const double* row = s_st + size_t(ci) * W * n_st;
      const double xp = row[p], yp = row[n_st + p], xq = row[q], yq = row[n_st + q];
      const double re = xp * xq + yp * yq;
      const double im = xp * yq - xq * yp;
      double amp = amp_in;
      if (gauss) {
        const double du = row[2 * n_st + q] - row[2 * n_st + p];
        const double dv = row[3 * n_st + q] - row[3 * n_st + p];
        amp = exp(f2c * ((-2.0 * M_PI * M_PI) * (du * du + dv * dv)));
      }
      if constexpr (SMEAR) {
        if (a.time_smear)
          amp *= double(Sinc(float(0.5 * (row[4 * n_st + q] - row[4 * n_st + p]) * f)));
        if (a.freq_smear)
          amp *= double(Sinc(float(0.5 * (row[5 * n_st + q] - row[5 * n_st + p]) * width)));
      }
      const double* sp = s_spec + 4 * ci;
      if constexpr (SI) {
        const double I = sp[0];
        v[j][0].x += amp * (re * I);
        v[j][0].y += amp * (im * I);
      } else {
        // XX = I + Q, XY = U + iV, YX = U - iV, YY = I - Q
        const double I = sp[0], Q = sp[1], U = sp[2], V = sp[3];
        const double2 sv[4] = {make_double2(I + Q, 0.0), make_double2(U, V),
                               make_double2(U, -V), make_double2(I - Q, 0.0)};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int kk = NC == 2 ? (k == 0 ? 0 : (k == 3 ? 1 : -1)) : k;
          if (kk < 0) continue;
          v[j][kk].x += amp * (re * sv[k].x - im * sv[k].y);
          v[j][kk].y += amp * (re * sv[k].y + im * sv[k].x);
        }
      }
    };
#pragma unroll
    for (int j = 0; j < kPredictBaselines; ++j) {
#pragma unroll 2
      for (int ci = 0; ci < np; ++ci) add(j, ci, 1.0, false);
#pragma unroll 2
      for (int ci = np; ci < nc; ++ci) add(j, ci, 1.0, true);
    }
  }
  // correlations stored: Duo (XX, YY) or Full (XX, XY, YX, YY); Stokes-I
  // only: XX = YY = I, XY = YX = 0
  const double inv = F16 ? 1.0 / double(a.scale[size_t(pc.cb) * a.n_dir + d]) : 1.0;
#pragma unroll
  for (int j = 0; j < kPredictBaselines; ++j) {
    const int bl = bl0 + int(threadIdx.x) + j * kPredictThreads;
    if (bl >= a.bl_stride) break;
    double2 out[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      if (bl >= a.n_bl)  // padding lanes hold zeros
        out[k] = make_double2(0.0, 0.0);
      else if constexpr (SI)
        out[k] = (NC == 2 || k == 0 || k == 3) ? v[j][0] : make_double2(0.0, 0.0);
      else
        out[k] = v[j][k];
    }
    const int64_t idx = pc.base + d * pc.plane + int64_t(t) * pc.t_stride + bl;
    if constexpr (!F16) {
      float4* o = reinterpret_cast<float4*>(a.model) + idx * (NC / 2);
#pragma unroll
      for (int h = 0; h < NC / 2; ++h)
        o[h] = make_float4(float(out[2 * h].x), float(out[2 * h].y), float(out[2 * h + 1].x),
                           float(out[2 * h + 1].y));
    } else {
      uint32_t bits[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k)
        bits[k] = HalfBits(float(out[k].x * inv), float(out[k].y * inv));
      if constexpr (NC == 2) {
        reinterpret_cast<uint2*>(a.model)[idx] = make_uint2(bits[0], bits[1]);
      } else {
        reinterpret_cast<uint4*>(a.model)[idx] = make_uint4(bits[0], bits[1], bits[2], bits[3]);
      }
    }
  }
}

// FP16 scale bound per direction: max over channels of sum_c |S_c| (every
// model value of the direction is bounded by it); largest maps to 2^14.
__global__ void DirectionScaleKernel(const Stokes4* spec, const CompGeom* geom, int n_comp,
                                     int n_ch_total, int n_dir, int n_cb, float* scale) {
  const int d = blockIdx.x;
  __shared__ float red[256];
  float m = 0.0f;
  for (int ch = threadIdx.x; ch < n_ch_total; ch += blockDim.x) {
    float acc = 0.0f;
    for (int c = 0; c < n_comp; ++c) {
      if (geom[c].direction != d) continue;
      const Stokes4 s = spec[int64_t(c) * n_ch_total + ch];
      acc += float(fabs(s.I) + fabs(s.Q) + fabs(s.U) + fabs(s.V));
    }
    m = fmaxf(m, acc);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  // the same scale for every block: [n_cb][n_dir]
  const float b = red[0] > 0.0f ? red[0] / 16384.0f : 1.0f;
  for (int cb = threadIdx.x; cb < n_cb; cb += blockDim.x) scale[size_t(cb) * n_dir + d] = b;
}

// ---- data synthesis -------------------------------------------------------
__device__ __forceinline__ uint32_t Hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ float Uniform01(uint64_t key, uint32_t stream) {
  const uint32_t h = Hash(uint32_t(key) ^ Hash(uint32_t(key >> 32) ^ Hash(stream * 0x9e3779b9U)));
  return (float(h >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

// data = sum_d Diag(g_a) M_d Diag(g_b)^H + complex Gaussian noise of sigma
// per real and imaginary part (Box-Muller on a counter-based hash).
template <int NC, bool F16>
__global__ void SynthesizeKernel(const void* model, void* data, const float* scale,
                                 const double2* gains, const int32_t* ant1, const int32_t* ant2,
                                 const int32_t* sol_of_t, int n_t, int n_chb, int n_bl,
                                 int bl_stride, int n_dir, int n_ant, int n_sub, int cb,
                                 int64_t model_off, int64_t data_off, float sigma,
                                 uint32_t seed) {
  const int64_t rows = int64_t(n_t) * n_chb;
  const int64_t total = rows * bl_stride;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int bl = int(e % bl_stride);
  const int t = int(e / bl_stride / n_chb);
  float2 acc[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) acc[k] = make_float2(0, 0);
  if (bl < n_bl) {
    const int a = ant1[bl], b = ant2[bl];
    const double2* g = gains + size_t(cb) * n_ant * n_sub * 2;
    for (int d = 0; d < n_dir; ++d) {
      const int s = sol_of_t[d * n_t + t];
      float2 m[NC];
      const int64_t idx = model_off + int64_t(d) * total + e;
      if constexpr (!F16) {
        const float4* p = reinterpret_cast<const float4*>(model) + idx * (NC / 2);
#pragma unroll
        for (int h = 0; h < NC / 2; ++h) {
          const float4 x = p[h];
          m[2 * h] = make_float2(x.x, x.y);
          m[2 * h + 1] = make_float2(x.z, x.w);
        }
      } else {
        const float sc = scale[cb * n_dir + d];
        const uint32_t* p = reinterpret_cast<const uint32_t*>(model) + idx * NC;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          const half2_v h = __builtin_bit_cast(half2_v, p[k]);
          m[k] = make_float2(float(h.x) * sc, float(h.y) * sc);
        }
      }
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const int i = NC == 2 ? k : (k >> 1);
        const int j = NC == 2 ? k : (k & 1);
        const double2 ga = g[(size_t(a) * n_sub + s) * 2 + i];
        const double2 gb = g[(size_t(b) * n_sub + s) * 2 + j];
        const float2 gaf = make_float2(float(ga.x), float(ga.y));
        const float2 gbf = make_float2(float(gb.x), float(gb.y));
        const float2 c = cmulc(cmul(gaf, m[k]), gbf);
        acc[k].x += c.x;
        acc[k].y += c.y;
      }
    }
    if (sigma > 0.0f) {
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const uint64_t key = (uint64_t(cb) << 48) ^ (uint64_t(e) * NC + k);
        const float u1 = Uniform01(key, seed * 2u);
        const float u2 = Uniform01(key, seed * 2u + 1u);
        const float r = sigma * sqrtf(-2.0f * logf(u1));
        float sn, cs;
        sincosf(2.0f * float(M_PI) * u2, &sn, &cs);
        acc[k].x += r * cs;
        acc[k].y += r * sn;
      }
    }
  }
  float4* o = reinterpret_cast<float4*>(data) + (data_off + e) * (NC / 2);
#pragma unroll
  for (int h = 0; h < NC / 2; ++h)
    o[h] = make_float4(acc[2 * h].x, acc[2 * h].y, acc[2 * h + 1].x, acc[2 * h + 1].y);
}

template <typename T>
struct Scratch {
  T* ptr = nullptr;
  size_t n = 0;
  ~Scratch() {
    if (ptr) (void)hipFree(ptr);
  }
  T* Get(size_t count) {
    if (count > n) {
      if (ptr) DDH_HIP_CHECK(hipFree(ptr));
      ptr = nullptr;
      DDH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(count, 1) * sizeof(T)));
      n = count;
    }
    return ptr;
  }
};

template <int NC, bool F16, bool SI, bool SMEAR>
__global__ void __launch_bounds__(kPredictThreads) PredictKernel(const PredictArgs a) {
  PredictBody<NC, F16, SI, SMEAR>(a);
}
// Stokes I without smearing (the common case) at four waves per SIMD: two
// workgroups per compute unit instead of one; the registers it gives up cost
// a spill or two in the station terms' sincos, once per (component, station).
template <int NC, bool F16>
__global__ void __launch_bounds__(kPredictThreads) __attribute__((amdgpu_waves_per_eu(4)))
    PredictStokesIKernel(const PredictArgs a) {
  PredictBody<NC, F16, true, false>(a);
}

using PredictFn = void (*)(PredictArgs);
template <int NC, bool F16>
PredictFn PredictVariant(bool si, bool smear) {
  if (si) return smear ? PredictKernel<NC, F16, true, true> : PredictStokesIKernel<NC, F16>;
  return smear ? PredictKernel<NC, F16, false, true> : PredictKernel<NC, F16, false, false>;
}

void LaunchPredict(bool duo, bool f16, bool si, bool smear, dim3 grid, size_t lds,
                   hipStream_t s, const PredictArgs& pa) {
  const PredictFn fn = duo ? (f16 ? PredictVariant<2, true>(si, smear)
                                  : PredictVariant<2, false>(si, smear))
                           : (f16 ? PredictVariant<4, true>(si, smear)
                                  : PredictVariant<4, false>(si, smear));
  hipLaunchKernelGGL(fn, grid, dim3(kPredictThreads), lds, s, pa);
  DDH_HIP_CHECK(hipGetLastError());
}

}  // namespace
}  // namespace ddh

using namespace ddh;

extern "C" int ddh_predict(ddh_problem* problem, double ref_ra, double ref_dec,
                           const double* station_uvw, int32_t n_stations, const double* freq,
                           const double* chan_width, const ddh_component* components,
                           int32_t n_components, int32_t stokes_i_only,
                           int32_t correct_freq_smearing, const double* scaled_ncp_uvw) {
  return Guard([&] {
    DDH_REQUIRE(problem && station_uvw && freq && components, -1, "NULL argument");
    DDH_REQUIRE(n_components >= 1 && n_stations >= 1, -1, "empty predict");
    DDH_REQUIRE(!correct_freq_smearing || chan_width, -1, "frequency smearing needs widths");
    ProblemView v = GetProblemView(problem);
    DDH_REQUIRE(n_stations >= v.n_ant, -1, "station count smaller than the antenna count");
    DDH_REQUIRE(!v.uni, -4, "the predict writes XX and YY; single-correlation problems take "
                            "their models through upload or ingest");
    std::vector<ddh_component> comps(components, components + n_components);
    for (const ddh_component& c : comps) {
      DDH_REQUIRE(c.direction >= 0 && c.direction < v.n_dir, -1, "component direction out of range");
      DDH_REQUIRE(c.n_spectral_terms <= 8, -1, "at most 8 spectral terms");
    }
    // group components by direction, points before Gaussians (stable)
    std::stable_sort(comps.begin(), comps.end(),
                     [](const ddh_component& a, const ddh_component& b) {
                       return a.direction != b.direction ? a.direction < b.direction
                                                         : (a.is_gaussian != 0) < (b.is_gaussian != 0);
                     });
    std::vector<int32_t> c_off(v.n_dir + 1, 0), g_off(v.n_dir, 0);
    for (const ddh_component& c : comps) c_off[c.direction + 1]++;
    for (int d = 0; d < v.n_dir; ++d) c_off[d + 1] += c_off[d];
    for (int d = 0; d < v.n_dir; ++d) {
      g_off[d] = c_off[d];
      while (g_off[d] < c_off[d + 1] && !comps[g_off[d]].is_gaussian) ++g_off[d];
    }
    int n_ch_total = 0;
    for (int x : v.ch_per_cb) n_ch_total += x;
    hipStream_t s = v.stream;
    static thread_local Scratch<ddh_component> d_comps;
    static thread_local Scratch<CompGeom> d_geom;
    static thread_local Scratch<Stokes4> d_spec;
    static thread_local Scratch<double> d_uvw, d_freq, d_width;
    ddh_component* dc = d_comps.Get(comps.size());
    DDH_HIP_CHECK(hipMemcpyAsync(dc, comps.data(), comps.size() * sizeof(ddh_component),
                                 hipMemcpyHostToDevice, s));
    const size_t n_uvw = size_t(v.n_t) * n_stations * 3;
    double* du = d_uvw.Get(n_uvw);
    DDH_HIP_CHECK(hipMemcpyAsync(du, station_uvw, n_uvw * sizeof(double), hipMemcpyHostToDevice, s));
    double* df = d_freq.Get(n_ch_total);
    DDH_HIP_CHECK(hipMemcpyAsync(df, freq, n_ch_total * sizeof(double), hipMemcpyHostToDevice, s));
    double* dw = d_width.Get(n_ch_total);
    if (chan_width)
      DDH_HIP_CHECK(hipMemcpyAsync(dw, chan_width, n_ch_total * sizeof(double),
                                   hipMemcpyHostToDevice, s));
    CompGeom* dg = d_geom.Get(comps.size());
    hipLaunchKernelGGL(ComponentGeometryKernel, dim3((n_components + 127) / 128), dim3(128), 0, s,
                       dc, n_components, ref_ra, ref_dec, dg);
    Stokes4* dspec = d_spec.Get(size_t(n_components) * n_ch_total);
    const int n_spec = n_components * n_ch_total;
    hipLaunchKernelGGL(SpectrumKernel, dim3((n_spec + 255) / 256), dim3(256), 0, s, dc,
                       n_components, df, n_ch_total, dspec);
    DDH_HIP_CHECK(hipGetLastError());
    if (v.f16)  // one scale per direction, shared by every block
      hipLaunchKernelGGL(DirectionScaleKernel, dim3(v.n_dir), dim3(256), 0, s, dspec, dg,
                         n_components, n_ch_total, v.n_dir, v.n_cb, v.scale);
    // where each channel's rows start in its block's model planes
    std::vector<PredictChannel> chan(n_ch_total);
    {
      int ch = 0;
      for (int cb = 0; cb < v.n_cb; ++cb) {
        const int n_chb = v.ch_per_cb[cb];
        for (int k = 0; k < n_chb; ++k, ++ch) {
          PredictChannel& c = chan[ch];
          c.base = v.cb_host[cb].model_off + int64_t(k) * v.bl_stride;
          c.plane = int64_t(v.n_t) * n_chb * v.bl_stride;
          c.t_stride = int64_t(n_chb) * v.bl_stride;
          c.cb = cb;
          c.pad = 0;
        }
      }
    }
    static thread_local Scratch<PredictChannel> d_chan;
    static thread_local Scratch<int32_t> d_coff;
    PredictChannel* dch = d_chan.Get(chan.size());
    DDH_HIP_CHECK(hipMemcpyAsync(dch, chan.data(), chan.size() * sizeof(PredictChannel),
                                 hipMemcpyHostToDevice, s));
    // c_off then g_off in one buffer
    std::vector<int32_t> offs(c_off);
    offs.insert(offs.end(), g_off.begin(), g_off.end());
    int32_t* dco = d_coff.Get(offs.size());
    DDH_HIP_CHECK(hipMemcpyAsync(dco, offs.data(), offs.size() * sizeof(int32_t),
                                 hipMemcpyHostToDevice, s));
    PredictArgs pa{};
    pa.uvw = du;
    pa.geom = dg;
    pa.spec = dspec;
    pa.c_off = dco;
    pa.g_off = dco + c_off.size();
    pa.freq = df;
    pa.width = dw;
    pa.ant1 = v.ant1;
    pa.ant2 = v.ant2;
    pa.chan = dch;
    pa.scale = v.scale;
    pa.model = v.model;
    pa.n_st = n_stations;
    pa.n_bl = v.n_bl;
    pa.bl_stride = v.bl_stride;
    pa.n_t = v.n_t;
    pa.n_dir = v.n_dir;
    pa.n_ch_total = n_ch_total;
    pa.freq_smear = correct_freq_smearing ? 1 : 0;
    pa.time_smear = scaled_ncp_uvw ? 1 : 0;
    pa.ncp_v = scaled_ncp_uvw ? scaled_ncp_uvw[1] : 0.0;
    pa.ncp_w = scaled_ncp_uvw ? scaled_ncp_uvw[2] : 0.0;
    const bool smear = pa.freq_smear || pa.time_smear;
    // LDS per component: its station terms and its Stokes values
    const size_t per_comp = (size_t(PredictWords(smear)) * n_stations + 4) * sizeof(double);
    DDH_REQUIRE(per_comp <= kPredictLds, -1, "too many stations for the predict's LDS");
    pa.chunk = int(kPredictLds / per_comp);
    const int per_wg = kPredictThreads * kPredictBaselines;
    const dim3 grid(n_ch_total, v.n_t * v.n_dir, (v.bl_stride + per_wg - 1) / per_wg);
    const size_t lds = size_t(pa.chunk) * per_comp;
    LaunchPredict(v.n_corr == 2, v.f16, stokes_i_only != 0, smear, grid, lds, s, pa);
    DDH_HIP_CHECK(hipStreamSynchronize(s));
  });
}

extern "C" int ddh_problem_synthesize_data(ddh_problem* problem, double* const* gains,
                                           int32_t n_sub, const uint32_t* sol_index,
                                           float noise_sigma, uint32_t seed) {
  return Guard([&] {
    DDH_REQUIRE(problem && gains && sol_index, -1, "NULL argument");
    ProblemView v = GetProblemView(problem);
    hipStream_t s = v.stream;
    const size_t per_cb = size_t(v.n_ant) * n_sub * 2;
    static thread_local Scratch<double2> d_g;
    static thread_local Scratch<int32_t> d_sol;
    double2* g = d_g.Get(per_cb * v.n_cb);
    for (int cb = 0; cb < v.n_cb; ++cb)
      DDH_HIP_CHECK(hipMemcpyAsync(g + cb * per_cb, gains[cb], per_cb * sizeof(double2),
                                   hipMemcpyHostToDevice, s));
    std::vector<int32_t> sol(size_t(v.n_dir) * v.n_t);
    for (size_t i = 0; i < sol.size(); ++i) {
      DDH_REQUIRE(int(sol_index[i]) < n_sub, -1, "solution index out of range");
      sol[i] = int32_t(sol_index[i]);
    }
    int32_t* ds = d_sol.Get(sol.size());
    DDH_HIP_CHECK(hipMemcpyAsync(ds, sol.data(), sol.size() * sizeof(int32_t),
                                 hipMemcpyHostToDevice, s));
    for (int cb = 0; cb < v.n_cb; ++cb) {
      const int n_chb = v.ch_per_cb[cb];
      const int64_t total = int64_t(v.n_t) * n_chb * v.bl_stride;
      const dim3 grid(unsigned((total + 255) / 256));
#define DDH_SYN_LAUNCH(NC, F16)                                                              \
  hipLaunchKernelGGL((SynthesizeKernel<NC, F16>), grid, dim3(256), 0, s, v.model, v.data,   \
                     v.scale, g, v.ant1, v.ant2, ds, v.n_t, n_chb, v.n_bl, v.bl_stride,     \
                     v.n_dir, v.n_ant, n_sub, cb, v.cb_host[cb].model_off,                 \
                     v.cb_host[cb].data_off, noise_sigma, seed)
      if (v.n_corr == 2) {
        if (v.f16) DDH_SYN_LAUNCH(2, true); else DDH_SYN_LAUNCH(2, false);
      } else {
        if (v.f16) DDH_SYN_LAUNCH(4, true); else DDH_SYN_LAUNCH(4, false);
      }
#undef DDH_SYN_LAUNCH
      DDH_HIP_CHECK(hipGetLastError());
    }
    DDH_HIP_CHECK(hipStreamSynchronize(s));
  });
}
