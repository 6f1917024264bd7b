// Oracle restatement of the smoothness constraint. TEST INFRASTRUCTURE ONLY
// (see oracle.h).
//
//   ddecal/constraints/SmoothnessConstraint.cc:50-97   (Apply: one series per
//       antenna x sub-solution x polarization over the channel blocks,
//       non-finite values enter with weight 0, kernel size factor
//       (1 / dd_factor) / antenna_factor)
//   ddecal/constraints/KernelSmoother.h                (Gaussian kernel,
//       truncation windows, frequency-dependent bandwidth, weighted mean,
//       NaN where the weights sum to zero)
#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "oracle.h"

namespace {

using DComplex = std::complex<double>;

// Gaussian kernel at `distance` (Hz), unnormalised, cut at |x| > 1 when
// truncating (KernelSmoother::Kernel, GaussianKernel case).
double Gaussian(double distance, double bandwidth, bool truncate) {
  const double x = distance / bandwidth;
  if (truncate && (x < -1.0 || x > 1.0)) return 0.0;
  return std::exp(-x * x * 9.0);
}

// The weighted kernel mean of one series over the channel grid `f`
// (ascending), in place (KernelSmoother::Smooth).
void SmoothSeries(const std::vector<double>& f, double bandwidth, double ref_frequency,
                  double spectral_exponent, bool truncate, DComplex* data,
                  const double* weight, double size_factor, std::vector<DComplex>& out) {
  const size_t n = f.size();
  // The truncation window [left, right) moves monotonically up the grid.
  size_t left = 0;
  size_t right = 0;
  while (right != n && f[right] < f[0] + bandwidth * 0.5) ++right;
  right = std::min(right + 1, n);
  for (size_t i = 0; i != n; ++i) {
    const double local_bw = ref_frequency == 0.0 ? bandwidth : bandwidth * f[i] / ref_frequency;
    size_t first = 0, last = n;
    if (truncate) {
      while (f[left] < f[i] - local_bw * 0.5) ++left;
      while (right != n && f[right] < f[i] + local_bw * 0.5) ++right;
      first = left > 0 ? left - 1 : 0;
      last = right < n ? right + 1 : n;
    }
    const double correction = std::pow(local_bw / bandwidth, spectral_exponent) * size_factor;
    DComplex sum(0.0, 0.0);
    double weight_sum = 0.0;
    for (size_t j = first; j != last; ++j) {
      const double w = Gaussian((f[i] - f[j]) * correction, bandwidth, truncate) * weight[j];
      sum += data[j] * w;
      weight_sum += w;
    }
    out[i] = weight_sum == 0.0 ? DComplex(std::numeric_limits<double>::quiet_NaN(),
                                          std::numeric_limits<double>::quiet_NaN())
                               : sum / weight_sum;
  }
  std::copy(out.begin(), out.end(), data);
}

}  // namespace

extern "C" int orc_smooth(const orc_smoothness* c, int32_t n_cb, int32_t n_ant, int32_t n_sub,
                          int32_t n_pol, double* solutions) {
  if (!c || !c->frequencies || n_cb <= 0 || n_ant <= 0 || n_sub <= 0 || n_pol <= 0 ||
      !(c->bandwidth_hz > 0.0))
    return -1;
  const std::vector<double> f(c->frequencies, c->frequencies + n_cb);
  const size_t per_cb = size_t(n_ant) * n_sub * n_pol;  // series per block
  DComplex* sol = reinterpret_cast<DComplex*>(solutions);
  std::vector<DComplex> series(n_cb), out(n_cb);
  std::vector<double> weight(n_cb);
  for (size_t idx = 0; idx != per_cb; ++idx) {
    const size_t sub = (idx / n_pol) % n_sub;
    const size_t ant = idx / (size_t(n_sub) * n_pol);
    const double* w_ant = c->sub_solution_weights
                              ? c->sub_solution_weights + sub * size_t(n_ant) * n_cb
                              : c->weights;
    for (int32_t cb = 0; cb != n_cb; ++cb) {
      const DComplex v = sol[cb * per_cb + idx];
      const bool finite = std::isfinite(v.real()) && std::isfinite(v.imag());
      series[cb] = finite ? v : DComplex(0.0, 0.0);
      weight[cb] = finite ? (w_ant ? w_ant[ant * n_cb + cb] : 1.0) : 0.0;
    }
    const double dd = c->dd_factors ? 1.0 / c->dd_factors[sub] : 1.0;
    const double ant_factor = c->antenna_factors ? c->antenna_factors[ant] : 1.0;
    SmoothSeries(f, c->bandwidth_hz, c->bandwidth_ref_frequency_hz, c->spectral_exponent,
                 c->kernel_truncation != 0, series.data(), weight.data(), dd / ant_factor, out);
    for (int32_t cb = 0; cb != n_cb; ++cb) sol[cb * per_cb + idx] = series[cb];
  }
  return 0;
}
