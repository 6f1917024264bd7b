// CPU oracle: restatement of DP3's direct-DFT predict for point and Gaussian
// components. TEST INFRASTRUCTURE ONLY.
//
// Reference (lofar-astron/DP3 6.6.0):
//   base/Simulator.h:35-60          radec2lmn
//   base/Simulator.cc:110-282       visit(PointSource)
//   base/Simulator.cc:284-468       visit(GaussianSource)
//   base/Simulator.cc:472-549       sinc, phases, spectrum
//   base/PointSource.cc:51-99       PointSource::stokes
//   steps/PhaseShift.cc:197-212     fillEulerMatrix
#include <cmath>
#include <complex>
#include <vector>

#include "oracle.h"

namespace {
constexpr double kSpeedOfLight = 299792458.0;
using CD = std::complex<double>;

inline float Sinc(float x) { return std::abs(x) < 1e-6f ? 1.0f : std::sin(x) / x; }

void FillEuler(double m[3][3], double ra, double dec) {
  const double sinra = std::sin(ra), cosra = std::cos(ra);
  const double sindec = std::sin(dec), cosdec = std::cos(dec);
  // casacore Matrix(row, col)
  m[0][0] = cosra;
  m[1][0] = -sinra;
  m[2][0] = 0;
  m[0][1] = -sinra * sindec;
  m[1][1] = -cosra * sindec;
  m[2][1] = cosdec;
  m[0][2] = sinra * cosdec;
  m[1][2] = cosra * cosdec;
  m[2][2] = sindec;
}
}  // namespace

extern "C" void orc_radec2lmn(double ref_ra, double ref_dec, double ra,
                              double dec, double* lmn) {
  const double delta_ra = ra - ref_ra;
  const double sin_delta_ra = std::sin(delta_ra);
  const double cos_delta_ra = std::cos(delta_ra);
  const double sin_dec = std::sin(dec);
  const double cos_dec = std::cos(dec);
  const double sin_dec0 = std::sin(ref_dec);
  const double cos_dec0 = std::cos(ref_dec);
  lmn[0] = cos_dec * sin_delta_ra;
  lmn[1] = sin_dec * cos_dec0 - cos_dec * sin_dec0 * cos_delta_ra;
  lmn[2] = sin_dec * sin_dec0 + cos_dec * cos_dec0 * cos_delta_ra;
}

extern "C" void orc_stokes(const orc_component* c, double freq, double* iquv) {
  double I = c->I, Q = c->Q, U = c->U, V = c->V;
  if (c->n_spectral_terms > 0) {
    if (c->log_si) {
      const double base = std::log10(freq) - std::log10(c->ref_freq);
      double exponent = 0.0;
      for (int i = c->n_spectral_terms - 1; i >= 0; --i)
        exponent = exponent * base + c->spectral_terms[i];
      I *= std::pow(10., base * exponent);
    } else {
      const double x = freq / c->ref_freq - 1.0;
      double val = 0.0;
      for (int i = c->n_spectral_terms - 1; i >= 0; --i)
        val = val * x + c->spectral_terms[i];
      I += val * x;
    }
  }
  if (c->has_rm) {
    const double lambda = kSpeedOfLight / freq;
    const double chi = 2.0 * (c->pol_angle + c->rm * lambda * lambda);
    const double qu = I * c->pol_fraction;
    Q = qu * std::cos(chi);
    U = qu * std::sin(chi);
  }
  iquv[0] = I;
  iquv[1] = Q;
  iquv[2] = U;
  iquv[3] = V;
}

extern "C" void orc_simulate(double ref_ra, double ref_dec, int32_t n_st,
                             int32_t n_bl, const int32_t* bl_p,
                             const int32_t* bl_q, int32_t n_ch,
                             const double* freq, const double* chan_width,
                             const double* uvw, const double* ncp,
                             int32_t time_smear, int32_t freq_smear,
                             int32_t stokes_i_only, int32_t n_comp,
                             const orc_component* comps, double* buffer_d) {
  CD* buffer = reinterpret_cast<CD*>(buffer_d);
  const int n_corr = stokes_i_only ? 1 : 4;
  const double ncp_uvw[3] = {ncp ? ncp[0] : 0.0, ncp ? ncp[1] : 0.0,
                             ncp ? ncp[2] : 0.0};
  std::vector<double> shift_re(size_t(n_st) * n_ch), shift_im(size_t(n_st) * n_ch);
  std::vector<double> st_phase(n_st), st_rot(n_st);
  std::vector<double> spec_re(size_t(n_ch) * 4), spec_im(size_t(n_ch) * 4);
  std::vector<double> smear(n_ch);
  constexpr double cinv = 2.0 * M_PI / kSpeedOfLight;

  for (int32_t ci = 0; ci != n_comp; ++ci) {
    const orc_component& c = comps[ci];
    double lmn[3];
    orc_radec2lmn(ref_ra, ref_dec, c.ra, c.dec, lmn);
    // phases (Simulator.cc:477-524)
    for (int32_t st = 0; st != n_st; ++st) {
      const double* u = &uvw[st * 3];
      st_phase[st] = cinv * (u[0] * lmn[0] + u[1] * lmn[1] + u[2] * (lmn[2] - 1.0));
      st_rot[st] = cinv * (lmn[0] * (u[1] * ncp_uvw[2] - u[2] * ncp_uvw[1]) +
                           lmn[1] * (-u[0] * ncp_uvw[2]) +
                           (lmn[2] - 1.0) * (u[0] * ncp_uvw[1]));
      for (int32_t ch = 0; ch != n_ch; ++ch) {
        const double ph = st_phase[st] * freq[ch];
        shift_im[size_t(st) * n_ch + ch] = std::sin(ph);
        shift_re[size_t(st) * n_ch + ch] = std::cos(ph);
      }
    }
    // spectrum (Simulator.cc:527-549)
    for (int32_t ch = 0; ch != n_ch; ++ch) {
      double s[4];
      orc_stokes(&c, freq[ch], s);
      if (stokes_i_only) {
        spec_re[ch] = s[0];
        spec_im[ch] = 0.0;
      } else {
        spec_re[ch * 4 + 0] = s[0] + s[1];
        spec_im[ch * 4 + 0] = 0.0;
        spec_re[ch * 4 + 1] = s[2];
        spec_im[ch * 4 + 1] = s[3];
        spec_re[ch * 4 + 2] = s[2];
        spec_im[ch * 4 + 2] = -s[3];
        spec_re[ch * 4 + 3] = s[0] - s[1];
        spec_im[ch * 4 + 3] = 0.0;
      }
    }
    // Gaussian set-up (Simulator.cc:296-342)
    double uvw_shift[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double cos_phi = 0, sin_phi = 0, u_scale = 0, v_scale = 0;
    if (c.is_gaussian) {
      const double phi = M_PI_2 + c.position_angle + M_PI;
      cos_phi = std::cos(phi);
      sin_phi = std::sin(phi);
      if (c.pa_is_absolute) {
        double e_pc[3][3], e_src[3][3];
        FillEuler(e_pc, ref_ra, ref_dec);
        FillEuler(e_src, c.ra, c.dec);
        // euler = transpose(e_src) * e_pc
        for (int i = 0; i != 3; ++i)
          for (int j = 0; j != 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k != 3; ++k) acc += e_src[k][i] * e_pc[k][j];
            uvw_shift[i][j] = acc;
          }
      }
      const double fwhm2sigma = 1.0 / (2.0 * std::sqrt(2.0 * std::log(2.0)));
      u_scale = c.major_axis * fwhm2sigma;
      v_scale = c.minor_axis * fwhm2sigma;
    }
    const double inv_c_sqr = 1.0 / (kSpeedOfLight * kSpeedOfLight);

    for (int32_t bl = 0; bl != n_bl; ++bl) {
      const int32_t p = bl_p[bl], q = bl_q[bl];
      if (p == q) continue;
      CD* buf = buffer + size_t(bl) * n_ch * n_corr;
      if (c.is_gaussian) {
        double uq[3], up[3];
        for (int i = 0; i != 3; ++i) {
          uq[i] = 0;
          up[i] = 0;
          for (int k = 0; k != 3; ++k) {
            uq[i] += uvw_shift[i][k] * uvw[q * 3 + k];
            up[i] += uvw_shift[i][k] * uvw[p * 3 + k];
          }
        }
        double u = uq[0], v = uq[1];
        u -= up[0];
        v -= up[1];
        const double u_prime = u_scale * (u * cos_phi - v * sin_phi);
        const double v_prime = v_scale * (u * sin_phi + v * cos_phi);
        const double uv_prime =
            (-2.0 * M_PI * M_PI) * (u_prime * u_prime + v_prime * v_prime);
        for (int32_t ch = 0; ch != n_ch; ++ch)
          smear[ch] = std::exp(freq[ch] * freq[ch] * inv_c_sqr * uv_prime);
        if (time_smear)
          for (int32_t ch = 0; ch != n_ch; ++ch)
            smear[ch] *= Sinc(0.5 * (st_rot[q] - st_rot[p]) * freq[ch]);
      } else {
        for (int32_t ch = 0; ch != n_ch; ++ch)
          smear[ch] = time_smear
                          ? Sinc(0.5 * (st_rot[q] - st_rot[p]) * freq[ch])
                          : 1.0;
      }
      if (freq_smear)
        for (int32_t ch = 0; ch != n_ch; ++ch)
          smear[ch] *= Sinc(0.5 * (st_phase[q] - st_phase[p]) * chan_width[ch]);
      for (int32_t ch = 0; ch != n_ch; ++ch) {
        const double xp = shift_re[size_t(p) * n_ch + ch];
        const double yp = shift_im[size_t(p) * n_ch + ch];
        const double xq = shift_re[size_t(q) * n_ch + ch];
        const double yq = shift_im[size_t(q) * n_ch + ch];
        const double qcp_re = xp * xq + yp * yq;
        const double qcp_im = xp * yq - xq * yp;
        for (int k = 0; k != n_corr; ++k) {
          const double xc = spec_re[ch * n_corr + k];
          const double yc = spec_im[ch * n_corr + k];
          const double re = qcp_re * xc - qcp_im * yc;
          const double im = qcp_re * yc + qcp_im * xc;
          buf[ch * n_corr + k] += CD(smear[ch] * re, smear[ch] * im);
        }
      }
    }
  }
}
