// CPU oracle: applying the solutions to the model data and subtracting them
// from the data after a solve. TEST INFRASTRUCTURE ONLY.
//
// Reference (lofar-astron/DP3 6.6.0):
//   steps/DDECal.cc:992-1020   ApplyFullJonesSolution / ApplyDiagonalSolution /
//                              ApplyScalarSolution (2x2 products in double)
//   steps/DDECal.cc:1031-1072  ApplySolutionAndSubtract: per (baseline,
//                              channel) the sum over directions of the
//                              corrected model, subtracted from the data;
//                              the corrected model kept when asked
//   steps/DDECal.cc:1074-1124  CorrectAndSubtractModels: over the time steps
//                              of the interval, channel block of each channel
#include <complex>
#include <vector>

#include "oracle.h"

namespace {
using CF = std::complex<float>;
using CD = std::complex<double>;

// 2x2 products on (XX, XY, YX, YY) in row-major order.
void Mul(const CD* a, const CD* b, CD* out) {
  out[0] = a[0] * b[0] + a[1] * b[2];
  out[1] = a[0] * b[1] + a[1] * b[3];
  out[2] = a[2] * b[0] + a[3] * b[2];
  out[3] = a[2] * b[1] + a[3] * b[3];
}
// a * b^H
void MulHerm(const CD* a, const CD* b, CD* out) {
  out[0] = a[0] * std::conj(b[0]) + a[1] * std::conj(b[1]);
  out[1] = a[0] * std::conj(b[2]) + a[1] * std::conj(b[3]);
  out[2] = a[2] * std::conj(b[0]) + a[3] * std::conj(b[1]);
  out[3] = a[2] * std::conj(b[2]) + a[3] * std::conj(b[3]);
}
}  // namespace

extern "C" void orc_correct_and_subtract(int32_t n_t, int32_t n_bl, int32_t n_ch,
                                         int32_t n_dir, int32_t n_ant, int32_t n_cb,
                                         int32_t n_pol, const int32_t* antenna1,
                                         const int32_t* antenna2,
                                         const int32_t* channel_block_start,
                                         const double* solutions, float* data, float* model,
                                         int32_t subtract, int32_t keep_model) {
  CF* d = reinterpret_cast<CF*>(data);
  CF* m = reinterpret_cast<CF*>(model);
  const CD* sol = reinterpret_cast<const CD*>(solutions);
  const size_t per_cb = size_t(n_ant) * n_dir * n_pol;
  for (int t = 0; t < n_t; ++t)
    for (int bl = 0; bl < n_bl; ++bl) {
      const int a1 = antenna1[bl], a2 = antenna2[bl];
      int cb = 0;
      for (int ch = 0; ch < n_ch; ++ch) {
        if (ch == channel_block_start[cb + 1]) ++cb;  // DDECal.cc:1106-1108
        const CD* s = sol + size_t(cb) * per_cb;
        CD sum[4] = {};
        for (int dir = 0; dir < n_dir; ++dir) {
          const size_t i1 = size_t(a1) * n_dir + dir, i2 = size_t(a2) * n_dir + dir;
          CF* mv = m + ((((size_t(dir) * n_t + t) * n_bl + bl) * n_ch + ch) * 4);
          const CD md[4] = {CD(mv[0]), CD(mv[1]), CD(mv[2]), CD(mv[3])};
          CD out[4];
          if (n_pol == 4) {
            CD tmp[4];
            Mul(s + i1 * 4, md, tmp);
            MulHerm(tmp, s + i2 * 4, out);
          } else if (n_pol == 2) {
            // diag(s1) * M * diag(s2)^H
            const CD* s1 = s + i1 * 2;
            const CD* s2 = s + i2 * 2;
            out[0] = s1[0] * md[0] * std::conj(s2[0]);
            out[1] = s1[0] * md[1] * std::conj(s2[1]);
            out[2] = s1[1] * md[2] * std::conj(s2[0]);
            out[3] = s1[1] * md[3] * std::conj(s2[1]);
          } else {
            const CD f = s[i1] * std::conj(s[i2]);
            for (int c = 0; c < 4; ++c) out[c] = md[c] * f;
          }
          for (int c = 0; c < 4; ++c) sum[c] += out[c];
          if (keep_model)
            for (int c = 0; c < 4; ++c) mv[c] = CF(out[c]);
        }
        if (subtract) {
          CF* dv = d + (((size_t(t) * n_bl + bl) * n_ch + ch) * 4);
          // complex<float> -= complex<double> (DDECal.cc:1070): promoted to
          // double, subtracted, rounded once
          for (int c = 0; c < 4; ++c) dv[c] = CF(CD(dv[c]) - sum[c]);
        }
      }
    }
}
