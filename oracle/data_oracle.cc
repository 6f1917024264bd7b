// CPU oracle: input preparation (SolveData packing, AssignAndWeight) and the
// SolverTester fixture generator. TEST INFRASTRUCTURE ONLY.
//
// Reference (lofar-astron/DP3 6.6.0):
//   ddecal/gain_solvers/SolveData.cc:18-31,44-72,77-188
//   ddecal/gain_solvers/SolverTools.cc:48-167
//   ddecal/test/unit/SolverTester.cc:95-202,348-363
#include <cmath>
#include <complex>
#include <random>
#include <vector>

#include "oracle.h"

namespace {
using CF = std::complex<float>;
using CD = std::complex<double>;

// Corr selection of ToSpecificMatrix (SolveData.cc:20-27): Duo takes XX, YY.
inline int CorrIndex(int n_corr, int k) { return n_corr == 2 ? 3 * k : k; }

// One packed visibility (ToSpecificMatrix, SolveData.cc:19-31): the selected
// correlations, or for single-correlation data 0.5f * (XX + YY) in float.
inline void PackVisibility(const float* in, int n_corr, float* out) {
  if (n_corr == 1) {
    const CF v = 0.5f * (CF(in[0], in[1]) + CF(in[6], in[7]));
    out[0] = v.real();
    out[1] = v.imag();
    return;
  }
  for (int k = 0; k != n_corr; ++k) {
    const size_t c = CorrIndex(n_corr, k);
    out[2 * k] = in[c * 2];
    out[2 * k + 1] = in[c * 2 + 1];
  }
}
}  // namespace

extern "C" int orc_pack_solve_data(
    int32_t n_t, int32_t n_bl, int32_t n_ch, int32_t n_dir, int32_t n_cb,
    int32_t n_corr, int32_t n_ant, const int32_t* n_sol_dir,
    const int32_t* antennas1, const int32_t* antennas2, const float* data,
    const float* model, int64_t* vis_offsets, float* out_data,
    float* out_model, uint32_t* out_ant1, uint32_t* out_ant2,
    uint32_t* out_sol_map) {
  if (n_corr != 1 && n_corr != 2 && n_corr != 4) return -2;
  // Cross-correlation count (SolveData.cc:95-100).
  size_t n_cross = 0;
  for (int32_t bl = 0; bl != n_bl; ++bl) {
    if (antennas1[bl] < 0 || antennas2[bl] < 0 || antennas1[bl] >= n_ant ||
        antennas2[bl] >= n_ant)
      return -1;
    if (antennas1[bl] != antennas2[bl]) ++n_cross;
  }
  // LimitNSolutionsPerDirection (SolveData.cc:62-72) and start indices
  // (SolutionStartIndices, :45-56).
  std::vector<uint32_t> n_sol(n_dir), start(n_dir);
  uint32_t acc = 0;
  for (int32_t d = 0; d != n_dir; ++d) {
    n_sol[d] = std::min<uint32_t>(n_sol_dir[d], n_t);
    start[d] = acc;
    acc += n_sol[d];
  }
  std::vector<size_t> ch_begin(n_cb + 1, 0);
  vis_offsets[0] = 0;
  for (int32_t cb = 0; cb != n_cb; ++cb) {
    ch_begin[cb + 1] = size_t(cb + 1) * n_ch / n_cb;
    const size_t cb_size = ch_begin[cb + 1] - ch_begin[cb];
    vis_offsets[cb + 1] = vis_offsets[cb] + int64_t(n_t * n_cross * cb_size);
  }
  if (!data) return 0;

  std::vector<size_t> vis_index(n_cb, 0);
  for (int32_t t = 0; t != n_t; ++t) {
    for (int32_t bl = 0; bl != n_bl; ++bl) {
      const int32_t a1 = antennas1[bl], a2 = antennas2[bl];
      if (a1 == a2) continue;
      for (int32_t cb = 0; cb != n_cb; ++cb) {
        const size_t v0 = vis_offsets[cb];
        const size_t nv = vis_offsets[cb + 1] - v0;
        size_t& vi = vis_index[cb];
        const size_t first = ch_begin[cb];
        const size_t cb_size = ch_begin[cb + 1] - first;
        for (size_t i = 0; i != cb_size; ++i) {
          const size_t src = ((size_t(t) * n_bl + bl) * n_ch + first + i) * 4;
          PackVisibility(&data[src * 2], n_corr, &out_data[(v0 + vi + i) * n_corr * 2]);
          out_ant1[v0 + vi + i] = a1;
          out_ant2[v0 + vi + i] = a2;
        }
        for (int32_t d = 0; d != n_dir; ++d) {
          const uint32_t si = uint32_t(size_t(t) * n_sol[d] / n_t) + start[d];
          for (size_t i = 0; i != cb_size; ++i) {
            const size_t src =
                (((size_t(d) * n_t + t) * n_bl + bl) * n_ch + first + i) * 4;
            const size_t dst = v0 * n_dir + size_t(d) * nv + vi + i;
            PackVisibility(&model[src * 2], n_corr, &out_model[dst * n_corr * 2]);
            out_sol_map[dst] = si;
          }
        }
        vi += cb_size;
      }
    }
  }
  return 0;
}

// AssignAndWeight with linear_weighting_mode == false (SolverTools.cc:81-167):
// one flag per (bl, ch) = OR over correlations of the input flag, a non-finite
// data value or a non-finite value in any direction's model; data and models
// are multiplied by sqrt(weight) (Weigh, :48-79) and flagged entries zeroed.
extern "C" void orc_assign_and_weight(int32_t n_t, int32_t n_bl, int32_t n_ch,
                                      int32_t n_dir, float* data,
                                      const uint8_t* flags,
                                      const float* weights, float* model) {
  const size_t n_elem = size_t(n_bl) * n_ch;  // per time step
  for (int32_t t = 0; t != n_t; ++t) {
    float* d_t = data + size_t(t) * n_elem * 8;
    const uint8_t* f_t = flags + size_t(t) * n_elem * 4;
    const float* w_t = weights + size_t(t) * n_elem * 4;
    std::vector<int> flag(n_elem, 0);
    for (size_t e = 0; e != n_elem; ++e) {
      for (int c = 0; c != 4; ++c) {
        const float* v = d_t + (e * 4 + c) * 2;
        flag[e] |= f_t[e * 4 + c] | !(std::isfinite(v[0]) && std::isfinite(v[1]));
        for (int32_t d = 0; d != n_dir; ++d) {
          const float* m =
              model + ((size_t(d) * n_t + t) * n_elem * 4 + e * 4 + c) * 2;
          flag[e] |= int(!(std::isfinite(m[0]) && std::isfinite(m[1])));
        }
      }
    }
    for (size_t e = 0; e != n_elem; ++e) {
      for (int c = 0; c != 4; ++c) {
        const float sw = std::sqrt(w_t[e * 4 + c]);
        float* v = d_t + (e * 4 + c) * 2;
        if (flag[e]) {
          v[0] = 0.0f;
          v[1] = 0.0f;
        } else {
          v[0] *= sw;
          v[1] *= sw;
        }
        for (int32_t d = 0; d != n_dir; ++d) {
          float* m = model + ((size_t(d) * n_t + t) * n_elem * 4 + e * 4 + c) * 2;
          if (flag[e]) {
            m[0] = 0.0f;
            m[1] = 0.0f;
          } else {
            m[0] *= sw;
            m[1] *= sw;
          }
        }
      }
    }
  }
}

// SolverTester fixture constants (SolverTester.h:103-125).
namespace {
constexpr int kNAntennas = 50;
constexpr int kNDirections = 3;
constexpr int kNChannels = 16;
constexpr int kNChannelBlocks = 4;
constexpr int kNRegularTimes = 50;
constexpr int kNBaselines = kNAntennas * (kNAntennas - 1) / 2;
}  // namespace

extern "C" void orc_solvertester_shape(int32_t* n_ant, int32_t* n_dir,
                                       int32_t* n_ch, int32_t* n_t,
                                       int32_t* n_cb) {
  *n_ant = kNAntennas;
  *n_dir = kNDirections;
  *n_ch = kNChannels;
  *n_t = kNRegularTimes;
  *n_cb = kNChannelBlocks;
}

namespace {

// FillDdIntervalData (SolverTester.cc:95-202) for the input solutions
// `input` ([ant][sub][2]): diagonal gains, or with `leakage` the leakage
// terms of J = [[1, l_0], [l_1, 1]]. Weights are 1 and nothing is flagged,
// so the AssignAndWeight pass at :199 is the identity.
void FillDdIntervalData(const int32_t* n_sol_dir, const std::vector<CF>& input,
                        size_t n_solutions, float* data, float* model, bool leakage = false) {
  std::uniform_real_distribution<float> uniform_data(-1.0, 1.0);
  std::mt19937 mt(0);
  std::vector<std::vector<size_t>> solution_indices(kNDirections);
  size_t first = 0;
  for (size_t d = 0; d != kNDirections; ++d) {
    solution_indices[d].resize(kNRegularTimes);
    for (size_t t = 0; t != kNRegularTimes; ++t)
      solution_indices[d][t] = first + t * n_sol_dir[d] / kNRegularTimes;
    first += n_sol_dir[d];
  }
  const size_t n_elem = size_t(kNBaselines) * kNChannels * 4;
  for (size_t t = 0; t != kNRegularTimes; ++t) {
    CF* time_data = reinterpret_cast<CF*>(data) + t * n_elem;
    for (size_t d = 0; d != kNDirections; ++d) {
      CF* this_direction =
          reinterpret_cast<CF*>(model) + (d * kNRegularTimes + t) * n_elem;
      for (size_t bl = 0; bl != kNBaselines; ++bl) {
        for (size_t ch = 0; ch != kNChannels; ++ch) {
          const size_t mi = (bl * kNChannels + ch) * 4;
          // Same expression shapes as the fixture, so that the compiler picks
          // the same draw order.
          this_direction[mi + 0] = CF(uniform_data(mt), uniform_data(mt));
          this_direction[mi + 1] = CF(uniform_data(mt), uniform_data(mt)) * 0.1f;
          this_direction[mi + 2] = CF(uniform_data(mt), uniform_data(mt)) * 0.1f;
          this_direction[mi + 3] = CF(uniform_data(mt), uniform_data(mt)) * 1.5f;
        }
      }
    }
    size_t bl = 0;
    for (size_t a1 = 0; a1 != kNAntennas; ++a1) {
      for (size_t a2 = a1 + 1; a2 != kNAntennas; ++a2) {
        for (size_t ch = 0; ch != kNChannels; ++ch) {
          CD sum[4] = {};
          for (size_t d = 0; d != kNDirections; ++d) {
            const size_t si = solution_indices[d][t];
            const CF* val = reinterpret_cast<const CF*>(model) +
                            (d * kNRegularTimes + t) * n_elem +
                            (bl * kNChannels + ch) * 4;
            const CD ga[2] = {CD(input[(a1 * n_solutions + si) * 2 + 0]),
                              CD(input[(a1 * n_solutions + si) * 2 + 1])};
            const CD gb[2] = {CD(input[(a2 * n_solutions + si) * 2 + 0]),
                              CD(input[(a2 * n_solutions + si) * 2 + 1])};
            if (leakage) {
              // left * val.MultiplyHerm(right) with left = [[1, ga0], [ga1, 1]],
              // right = [[1, gb0], [gb1, 1]]: val * right^H, then left * that
              const CD left[4] = {1.0, ga[0], ga[1], 1.0};
              const CD right_h[4] = {1.0, std::conj(gb[1]), std::conj(gb[0]), 1.0};
              CD vr[4];
              for (int i = 0; i != 2; ++i)
                for (int j = 0; j != 2; ++j)
                  vr[2 * i + j] = CD(val[2 * i]) * right_h[j] + CD(val[2 * i + 1]) * right_h[2 + j];
              for (int i = 0; i != 2; ++i)
                for (int j = 0; j != 2; ++j)
                  sum[2 * i + j] += left[2 * i] * vr[j] + left[2 * i + 1] * vr[2 + j];
              continue;
            }
            // left * val.MultiplyHerm(right) = Diag(ga) * val * Diag(gb)^H
            for (int i = 0; i != 2; ++i)
              for (int j = 0; j != 2; ++j)
                sum[2 * i + j] += ga[i] * (CD(val[2 * i + j]) * std::conj(gb[j]));
          }
          for (int p = 0; p != 4; ++p)
            time_data[(bl * kNChannels + ch) * 4 + p] = CF(sum[p]);
        }
        ++bl;
      }
    }
  }
}

void CopyOut(const std::vector<CF>& input, double* input_solutions) {
  for (size_t i = 0; i != input.size(); ++i) {
    input_solutions[2 * i] = input[i].real();
    input_solutions[2 * i + 1] = input[i].imag();
  }
}

size_t CountSolutions(const int32_t* n_sol_dir) {
  size_t n = 0;
  for (int d = 0; d != kNDirections; ++d) n += n_sol_dir[d];
  return n;
}

}  // namespace

// SetDiagonalSolutions (SolverTester.cc:348-363), then FillDdIntervalData.
extern "C" void orc_solvertester_fill(const int32_t* n_sol_dir, float* data,
                                      float* model, double* input_solutions) {
  const size_t n_solutions = CountSolutions(n_sol_dir);
  std::vector<CF> input(kNAntennas * n_solutions * 2);
  std::mt19937 mt;
  std::uniform_real_distribution<float> uniform_sols(1.0, 2.0);
  for (size_t ant = 0; ant != kNAntennas; ++ant) {
    for (size_t pol = 0; pol != 2; ++pol) {
      for (size_t s = 0; s != n_solutions; ++s) {
        const size_t index = (ant * n_solutions + s) * 2 + pol;
        input[index] = {uniform_sols(mt), uniform_sols(mt)};
        if (s != 0) input[index] *= 0.5f;
      }
    }
  }
  CopyOut(input, input_solutions);
  FillDdIntervalData(n_sol_dir, input, n_solutions, data, model);
}

// SetScalarSolutions (SolverTester.cc:331-346: one draw per (antenna,
// solution), copied to the second polarization), then FillDdIntervalData.
extern "C" void orc_solvertester_fill_scalar(const int32_t* n_sol_dir, float* data,
                                             float* model, double* input_solutions) {
  const size_t n_solutions = CountSolutions(n_sol_dir);
  std::vector<CF> input(kNAntennas * n_solutions * 2);
  std::mt19937 mt;
  std::uniform_real_distribution<float> uniform_sols(1.0, 2.0);
  for (size_t ant = 0; ant != kNAntennas; ++ant) {
    for (size_t s = 0; s != n_solutions; ++s) {
      const size_t index = (ant * n_solutions + s) * 2;
      input[index] = {uniform_sols(mt), uniform_sols(mt)};
      if (s != 0) input[index] *= 0.5f;
      input[index + 1] = input[index];
    }
  }
  CopyOut(input, input_solutions);
  FillDdIntervalData(n_sol_dir, input, n_solutions, data, model);
}

// SetLeakageSolutions (SolverTester.cc:365-379), then FillDdIntervalData with
// leakage_only.
extern "C" void orc_solvertester_fill_leakage(const int32_t* n_sol_dir, float* data,
                                              float* model, double* input_solutions) {
  const size_t n_solutions = CountSolutions(n_sol_dir);
  std::vector<CF> input(kNAntennas * n_solutions * 2);
  std::mt19937 mt;
  std::uniform_real_distribution<float> uniform_sols(0.0, 0.5);
  for (size_t ant = 0; ant != kNAntennas; ++ant) {
    for (size_t pol = 0; pol != 2; ++pol) {
      for (size_t s = 0; s != n_solutions; ++s) {
        const size_t index = (ant * n_solutions + s) * 2 + pol;
        input[index] = {uniform_sols(mt), uniform_sols(mt)};
        if (s != 0) input[index] *= 0.5f;
      }
    }
  }
  CopyOut(input, input_solutions);
  FillDdIntervalData(n_sol_dir, input, n_solutions, data, model, true);
}
