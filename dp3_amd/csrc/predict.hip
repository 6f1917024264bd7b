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

// Station phase phi = 2 pi / c (u l + v m + w (n - 1)) per (t, component,
// station) and its per-channel phasor exp(i phi f) for the channels of one
// block: table[(t * n_chb + ch) * n_comp + c][st] (complex128). With
// scaled_ncp_uvw (time smearing), also the station earth-rotation phase
// 2 pi / c (l (v ncp_w - w ncp_v) - m u ncp_w + (n - 1) u ncp_v)
// (Simulator.cc:488-503) per (t, component, station).
__global__ void StationPhasorKernel(const double* uvw, int n_st, const CompGeom* geom,
                                    int n_comp, const double* freq, int ch0, int n_chb,
                                    int n_t, double2* table, double* st_phase, double* st_rot,
                                    double ncp_v, double ncp_w) {
  const int64_t total = int64_t(n_t) * n_chb * n_comp * n_st;
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int st = int(idx % n_st);
  int64_t r = idx / n_st;
  const int c = int(r % n_comp);
  r /= n_comp;
  const int ch = int(r % n_chb);
  const int t = int(r / n_chb);
  const double* u = uvw + (int64_t(t) * n_st + st) * 3;
  const CompGeom& g = geom[c];
  constexpr double cinv = 2.0 * M_PI / kSpeedOfLight;
  const double phi = cinv * (u[0] * g.l + u[1] * g.m + u[2] * g.nm1);
  double s, co;
  sincos(phi * freq[ch0 + ch], &s, &co);
  table[idx] = make_double2(co, s);
  if (ch == 0) {
    const int64_t k = (int64_t(t) * n_comp + c) * n_st + st;
    st_phase[k] = phi;
    st_rot[k] = cinv * (g.l * (u[1] * ncp_w - u[2] * ncp_v) + g.m * (-u[0] * ncp_w) +
                        g.nm1 * (u[0] * ncp_v));
  }
}

__device__ __forceinline__ float Sinc(float x) { return fabsf(x) < 1e-6f ? 1.0f : sinf(x) / x; }

typedef _Float16 half2_v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t HalfBits(float re, float im) {
  half2_v h;
  h.x = static_cast<_Float16>(re);
  h.y = static_cast<_Float16>(im);
  return __builtin_bit_cast(uint32_t, h);
}

// One thread per baseline of row (t, ch); loops over the components of
// direction d in FP64 (Simulator.cc:157-281 for points, :284-468 for
// Gaussians: the envelope, then the time- and frequency-smearing factors, on
// shift_q conj(shift_p) S) and rounds the direction's sum once. Writes the
// model plane of (cb, d).
template <int NC, bool F16>
__global__ void BaselinePredictKernel(const double2* __restrict__ table,
                                      const double* __restrict__ st_phase,
                                      const double* __restrict__ st_rot,
                                      const CompGeom* __restrict__ geom,
                                      const Stokes4* __restrict__ spec,
                                      const double* __restrict__ uvw,
                                      const int32_t* __restrict__ ant1,
                                      const int32_t* __restrict__ ant2,
                                      const double* __restrict__ freq,
                                      const double* __restrict__ chan_width, int n_bl,
                                      int bl_stride, int n_st, int n_comp, int c0, int c1,
                                      int ch0, int n_chb, int n_ch_total, int stokes_i_only,
                                      int freq_smear, int time_smear, const float* scale,
                                      void* plane) {
  const int bl = blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y;
  const int t = blockIdx.z;
  if (bl >= bl_stride) return;
  const int64_t out_idx = (int64_t(t) * n_chb + ch) * bl_stride + bl;
  double2 v[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0),
                  make_double2(0, 0)};
  if (bl < n_bl) {
    const int p = ant1[bl], q = ant2[bl];
    const double* up = uvw + (int64_t(t) * n_st + p) * 3;
    const double* uq = uvw + (int64_t(t) * n_st + q) * 3;
    const double f = freq[ch0 + ch];
    const double f2c = f * f * (1.0 / (kSpeedOfLight * kSpeedOfLight));
    const double width = freq_smear ? chan_width[ch0 + ch] : 0.0;
    const double2* row = table + (int64_t(t) * n_chb + ch) * n_comp * int64_t(n_st);
    const double* phase_t = st_phase + int64_t(t) * n_comp * n_st;
    const double* rot_t = st_rot + int64_t(t) * n_comp * n_st;
    for (int c = c0; c < c1; ++c) {
      const double2 sp = row[int64_t(c) * n_st + p];
      const double2 sq = row[int64_t(c) * n_st + q];
      // shift_q * conj(shift_p)
      const double2 ph = make_double2(sp.x * sq.x + sp.y * sq.y, sp.x * sq.y - sq.x * sp.y);
      const CompGeom& g = geom[c];
      double amp = 1.0;
      if (g.is_gaussian) {
        // station UVW rotated to the source (identity unless absolute PA), then
        // the baseline's (u, v) difference (Simulator.cc:312-342)
        double uvq[2], uvp[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          uvq[i] = g.rot[i * 3] * uq[0] + g.rot[i * 3 + 1] * uq[1] + g.rot[i * 3 + 2] * uq[2];
          uvp[i] = g.rot[i * 3] * up[0] + g.rot[i * 3 + 1] * up[1] + g.rot[i * 3 + 2] * up[2];
        }
        const double u = uvq[0] - uvp[0];
        const double vv = uvq[1] - uvp[1];
        const double up_ = g.u_scale * (u * g.cos_phi - vv * g.sin_phi);
        const double vp_ = g.v_scale * (u * g.sin_phi + vv * g.cos_phi);
        const double uv = (-2.0 * M_PI * M_PI) * (up_ * up_ + vp_ * vp_);
        amp = exp(f2c * uv);
      }
      if (time_smear)
        amp *= double(Sinc(float(0.5 * (rot_t[int64_t(c) * n_st + q] -
                                        rot_t[int64_t(c) * n_st + p]) * f)));
      if (freq_smear)
        amp *= double(Sinc(float(0.5 * (phase_t[int64_t(c) * n_st + q] -
                                        phase_t[int64_t(c) * n_st + p]) * width)));
      const Stokes4 s = spec[int64_t(c) * n_ch_total + ch0 + ch];
      if (stokes_i_only) {
        v[0].x += amp * (ph.x * s.I);
        v[0].y += amp * (ph.y * s.I);
      } else {
        // XX = I + Q, XY = U + iV, YX = U - iV, YY = I - Q
        const double2 sv[4] = {make_double2(s.I + s.Q, 0.0), make_double2(s.U, s.V),
                               make_double2(s.U, -s.V), make_double2(s.I - s.Q, 0.0)};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          v[k].x += amp * (ph.x * sv[k].x - ph.y * sv[k].y);
          v[k].y += amp * (ph.x * sv[k].y + ph.y * sv[k].x);
        }
      }
    }
    if (stokes_i_only) v[3] = v[0];
  }
  // correlations stored: Duo (XX, YY) or Full (XX, XY, YX, YY)
  double2 out[NC];
  if constexpr (NC == 2) {
    out[0] = v[0];
    out[1] = v[3];
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = v[k];
  }
  if constexpr (!F16) {
    float4* o = reinterpret_cast<float4*>(plane) + out_idx * (NC / 2);
#pragma unroll
    for (int h = 0; h < NC / 2; ++h)
      o[h] = make_float4(float(out[2 * h].x), float(out[2 * h].y), float(out[2 * h + 1].x),
                         float(out[2 * h + 1].y));
  } else {
    const double inv = 1.0 / double(*scale);
    uint32_t bits[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) bits[k] = HalfBits(float(out[k].x * inv), float(out[k].y * inv));
    if constexpr (NC == 2) {
      reinterpret_cast<uint2*>(plane)[out_idx] = make_uint2(bits[0], bits[1]);
    } else {
      reinterpret_cast<uint4*>(plane)[out_idx] = make_uint4(bits[0], bits[1], bits[2], bits[3]);
    }
  }
}

// FP16 scale bound per direction: max over channels of sum_c |S_c| (every
// model value of the direction is bounded by it); largest maps to 2^14.
__global__ void DirectionScaleKernel(const Stokes4* spec, const CompGeom* geom, int n_comp,
                                     int n_ch_total, int n_dir, float* bound) {
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
  if (threadIdx.x == 0) bound[d] = red[0] > 0.0f ? red[0] / 16384.0f : 1.0f;
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
    // group components by direction (stable)
    std::stable_sort(comps.begin(), comps.end(),
                     [](const ddh_component& a, const ddh_component& b) {
                       return a.direction < b.direction;
                     });
    std::vector<int> c_off(v.n_dir + 1, 0);
    for (const ddh_component& c : comps) {
      DDH_REQUIRE(c.direction >= 0 && c.direction < v.n_dir, -1, "component direction out of range");
      DDH_REQUIRE(c.n_spectral_terms <= 8, -1, "at most 8 spectral terms");
      c_off[c.direction + 1]++;
    }
    for (int d = 0; d < v.n_dir; ++d) c_off[d + 1] += c_off[d];
    int n_ch_total = 0;
    for (int x : v.ch_per_cb) n_ch_total += x;
    hipStream_t s = v.stream;
    static thread_local Scratch<ddh_component> d_comps;
    static thread_local Scratch<CompGeom> d_geom;
    static thread_local Scratch<Stokes4> d_spec;
    static thread_local Scratch<double> d_uvw, d_freq, d_width, d_phase, d_rot;
    static thread_local Scratch<double2> d_table;
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
    if (v.f16) {
      // one scale per direction, shared by every block
      static thread_local Scratch<float> d_bound;
      float* bound = d_bound.Get(v.n_dir);
      hipLaunchKernelGGL(DirectionScaleKernel, dim3(v.n_dir), dim3(256), 0, s, dspec, dg,
                         n_components, n_ch_total, v.n_dir, bound);
      for (int cb = 0; cb < v.n_cb; ++cb)
        DDH_HIP_CHECK(hipMemcpyAsync(v.scale + size_t(cb) * v.n_dir, bound, v.n_dir * sizeof(float),
                                     hipMemcpyDeviceToDevice, s));
    }
    double* dphase = d_phase.Get(size_t(v.n_t) * n_components * n_stations);
    double* drot = d_rot.Get(size_t(v.n_t) * n_components * n_stations);
    const int time_smear = scaled_ncp_uvw ? 1 : 0;
    const double ncp_v = scaled_ncp_uvw ? scaled_ncp_uvw[1] : 0.0;
    const double ncp_w = scaled_ncp_uvw ? scaled_ncp_uvw[2] : 0.0;
    int ch0 = 0;
    for (int cb = 0; cb < v.n_cb; ++cb) {
      const int n_chb = v.ch_per_cb[cb];
      const int64_t n_tab = int64_t(v.n_t) * n_chb * n_components * n_stations;
      double2* tab = d_table.Get(size_t(n_tab));
      hipLaunchKernelGGL(StationPhasorKernel, dim3(unsigned((n_tab + 255) / 256)), dim3(256), 0, s,
                         du, n_stations, dg, n_components, df, ch0, n_chb, v.n_t, tab, dphase,
                         drot, ncp_v, ncp_w);
      DDH_HIP_CHECK(hipGetLastError());
      const size_t plane = size_t(v.n_t) * n_chb * v.bl_stride;
      const size_t vis_bytes = v.f16 ? size_t(v.n_corr) * 4 : size_t(v.n_corr) * 8;
      for (int d = 0; d < v.n_dir; ++d) {
        char* out = static_cast<char*>(v.model) +
                    (size_t(v.cb_host[cb].model_off) + d * plane) * vis_bytes;
        const float* sc = v.scale + size_t(cb) * v.n_dir + d;
        const dim3 grid((v.bl_stride + 255) / 256, n_chb, v.n_t);
        const int c0 = c_off[d], c1 = c_off[d + 1];
#define DDH_PREDICT_LAUNCH(NC, F16)                                                            \
  hipLaunchKernelGGL((BaselinePredictKernel<NC, F16>), grid, dim3(256), 0, s, tab, dphase, drot, \
                     dg, dspec, du, v.ant1, v.ant2, df, dw, v.n_bl, v.bl_stride, n_stations,   \
                     n_components, c0, c1, ch0, n_chb, n_ch_total, stokes_i_only,              \
                     correct_freq_smearing, time_smear, sc, out)
        if (v.n_corr == 2) {
          if (v.f16) DDH_PREDICT_LAUNCH(2, true); else DDH_PREDICT_LAUNCH(2, false);
        } else {
          if (v.f16) DDH_PREDICT_LAUNCH(4, true); else DDH_PREDICT_LAUNCH(4, false);
        }
#undef DDH_PREDICT_LAUNCH
        DDH_HIP_CHECK(hipGetLastError());
      }
      ch0 += n_chb;
    }
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
