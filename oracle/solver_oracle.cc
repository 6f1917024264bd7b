// CPU oracle: restatement of DP3's IterativeDiagonalSolver and the SolverBase
// logic it uses. TEST INFRASTRUCTURE ONLY (the checker; never the product).
//
// Reference (lofar-astron/DP3 6.6.0):
//   ddecal/gain_solvers/IterativeDiagonalSolver.cc:20-222
//   ddecal/gain_solvers/IterativeScalarSolver.cc:30-226 (scalar mode: one
//       solution polarization, Trace / Norm sums in double accumulators)
//   ddecal/gain_solvers/IterativeFullJonesSolver.cc:18-224 (full-Jones mode:
//       2x2 solutions, float numerator and denominator matrices, float
//       inverse; SolverBase.cc:269-283 and :373-405 for its convergence term
//       and MakeSolutionsFinite4Pol)
//   ddecal/gain_solvers/SolveData.h:204-244 (DiagonalAddOrSubtractDirection)
//   ddecal/gain_solvers/SolverBase.cc:63-371, SolverBase.h:274-285
//
// Precision follows the reference: visibilities, residuals, numerators and
// denominators in float; solutions, step, convergence and the final division in
// double. With settings->fp64 the same loop runs with double everywhere (the
// "FP64 oracle").
//
// Threading follows SolverBase::NSubThreads (SolverBase.cc:416-420): channel
// blocks are split over threads first (contiguous static chunks), remaining
// threads split each block's visibility range. Partial sums from visibility
// ranges are merged in range order (the reference merges in lock order), so
// results are deterministic for any thread count.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

using DComplex = std::complex<double>;

template <typename T>
struct Uni {
  static constexpr int kCorr = 1;
  std::complex<T> v[1];
};
template <typename T>
struct Duo {
  static constexpr int kCorr = 2;
  std::complex<T> v[2];
};
template <typename T>
struct Full {
  static constexpr int kCorr = 4;
  std::complex<T> v[4];
};

// ---------------------------------------------------------------------------
// 2x2 algebra (aocommon semantics, closed forms):
//   Diag(a) * M * Diag(b)^H : entry (i,j) = a_i * M_ij * conj(b_j)
//   (SolveData.h:230-237; row/column scaling as in the CUDA helper
//   kernels/MatrixComplex2x2.h:104-108).
// ---------------------------------------------------------------------------
// model * solution_1 * solution_2_conj (IterativeScalarSolver.cc:218-219)
template <typename T>
inline Uni<T> Contribution(const std::complex<T> a[2], const Uni<T>& m,
                           const std::complex<T> b[2]) {
  Uni<T> r;
  r.v[0] = (m.v[0] * a[0]) * std::conj(b[0]);
  return r;
}
template <typename T>
inline Duo<T> Contribution(const std::complex<T> a[2], const Duo<T>& m,
                           const std::complex<T> b[2]) {
  Duo<T> r;
  r.v[0] = (a[0] * m.v[0]) * std::conj(b[0]);
  r.v[1] = (a[1] * m.v[1]) * std::conj(b[1]);
  return r;
}
template <typename T>
inline Full<T> Contribution(const std::complex<T> a[2], const Full<T>& m,
                            const std::complex<T> b[2]) {
  Full<T> r;
  for (int i = 0; i != 2; ++i)
    for (int j = 0; j != 2; ++j)
      r.v[2 * i + j] = (a[i] * m.v[2 * i + j]) * std::conj(b[j]);
  return r;
}

// Full 2x2 matrices, row-major (00, 01, 10, 11) as the visibility layout.
// Products follow the textbook definitions; the inverse is the adjugate over
// the determinant and fails when the determinant is exactly zero (the
// aocommon Matrix2x2::Invert contract the reference relies on).
template <typename T>
struct Mat2 {
  std::complex<T> m[4];
};
template <typename T>
inline Mat2<T> Mul(const Mat2<T>& a, const Mat2<T>& b) {
  return {{a.m[0] * b.m[0] + a.m[1] * b.m[2], a.m[0] * b.m[1] + a.m[1] * b.m[3],
           a.m[2] * b.m[0] + a.m[3] * b.m[2], a.m[2] * b.m[1] + a.m[3] * b.m[3]}};
}
template <typename T>
inline Mat2<T> HermTranspose(const Mat2<T>& a) {
  return {{std::conj(a.m[0]), std::conj(a.m[2]), std::conj(a.m[1]), std::conj(a.m[3])}};
}
// a * b^H (MC2x2::MultiplyHerm)
template <typename T>
inline Mat2<T> MultiplyHerm(const Mat2<T>& a, const Mat2<T>& b) {
  return Mul(a, HermTranspose(b));
}
template <typename T>
inline bool Invert(Mat2<T>& a) {
  const std::complex<T> det = a.m[0] * a.m[3] - a.m[1] * a.m[2];
  if (det == std::complex<T>(0)) return false;
  const std::complex<T> r = std::complex<T>(1) / det;
  const std::complex<T> t = a.m[3] * r;
  a.m[1] = -a.m[1] * r;
  a.m[2] = -a.m[2] * r;
  a.m[3] = a.m[0] * r;
  a.m[0] = t;
  return true;
}
template <typename T>
inline void AddTo(Mat2<T>& x, const Mat2<T>& c) {
  for (int k = 0; k != 4; ++k) x.m[k] += c.m[k];
}

template <typename V>
inline void AddTo(V& x, const V& c) {
  for (int k = 0; k != V::kCorr; ++k) x.v[k] += c.v[k];
}
template <typename V>
inline void SubFrom(V& x, const V& c) {
  for (int k = 0; k != V::kCorr; ++k) x.v[k] -= c.v[k];
}

// SolveDirection's per-visibility accumulation for one baseline
// (IterativeDiagonalSolver.cc:172-197, AddNormToDenominator :20-35).
// ant1 side uses Diag(g_b) * M^H; ant2 side uses Diag(g_a) * M.
template <typename T>
inline void AccumulateAnt1(const Duo<T>& r, const Duo<T>& m,
                           const std::complex<T> gb[2], std::complex<T> num[2],
                           T den[2]) {
  for (int p = 0; p != 2; ++p) {
    const std::complex<T> cor = gb[p] * std::conj(m.v[p]);
    num[p] += r.v[p] * cor;
    den[p] += std::norm(cor);
  }
}
template <typename T>
inline void AccumulateAnt2(const Duo<T>& r, const Duo<T>& m,
                           const std::complex<T> ga[2], std::complex<T> num[2],
                           T den[2]) {
  for (int p = 0; p != 2; ++p) {
    const std::complex<T> cor = ga[p] * m.v[p];
    num[p] += std::conj(r.v[p]) * cor;
    den[p] += std::norm(cor);
  }
}
template <typename T>
inline void AccumulateAnt1(const Full<T>& r, const Full<T>& m,
                           const std::complex<T> gb[2], std::complex<T> num[2],
                           T den[2]) {
  // C = Diag(gb) * M^H ; M^H = [conj(M0) conj(M2); conj(M1) conj(M3)]
  const std::complex<T> c0 = gb[0] * std::conj(m.v[0]);
  const std::complex<T> c1 = gb[0] * std::conj(m.v[2]);
  const std::complex<T> c2 = gb[1] * std::conj(m.v[1]);
  const std::complex<T> c3 = gb[1] * std::conj(m.v[3]);
  // Diagonal(R * C)
  num[0] += r.v[0] * c0 + r.v[1] * c2;
  num[1] += r.v[2] * c1 + r.v[3] * c3;
  den[0] += std::norm(c0) + std::norm(c2);
  den[1] += std::norm(c1) + std::norm(c3);
}
template <typename T>
inline void AccumulateAnt2(const Full<T>& r, const Full<T>& m,
                           const std::complex<T> ga[2], std::complex<T> num[2],
                           T den[2]) {
  // C = Diag(ga) * M
  const std::complex<T> c0 = ga[0] * m.v[0];
  const std::complex<T> c1 = ga[0] * m.v[1];
  const std::complex<T> c2 = ga[1] * m.v[2];
  const std::complex<T> c3 = ga[1] * m.v[3];
  // Diagonal(R^H * C) ; R^H = [conj(R0) conj(R2); conj(R1) conj(R3)]
  num[0] += std::conj(r.v[0]) * c0 + std::conj(r.v[2]) * c2;
  num[1] += std::conj(r.v[1]) * c1 + std::conj(r.v[3]) * c3;
  den[0] += std::norm(c0) + std::norm(c2);
  den[1] += std::norm(c1) + std::norm(c3);
}

// IterativeScalarSolver::SolveDirection (IterativeScalarSolver.cc:142-173):
// with C1 = M^H g_b and C2 = M g_a, antenna 1 adds Trace(R C1) and Norm(C1),
// antenna 2 adds Trace(R^H C2) and Norm(C2); the products are float, the
// sums double.
// Single correlation: Trace and Norm of 1x1 matrices (IterativeScalarSolver.cc:20-27).
template <typename T>
inline void AccumulateScalar(const Uni<T>& r, const Uni<T>& m, std::complex<T> g1,
                             std::complex<T> g2, DComplex& num1, double& den1, DComplex& num2,
                             double& den2) {
  const std::complex<T> c1 = std::conj(m.v[0]) * g2;
  const std::complex<T> c2 = m.v[0] * g1;
  num1 += DComplex(r.v[0] * c1);
  den1 += double(std::norm(c1));
  num2 += DComplex(std::conj(r.v[0]) * c2);
  den2 += double(std::norm(c2));
}
template <typename T>
inline void AccumulateScalar(const Duo<T>& r, const Duo<T>& m, std::complex<T> g1,
                             std::complex<T> g2, DComplex& num1, double& den1, DComplex& num2,
                             double& den2) {
  std::complex<T> c1[2], c2[2];
  for (int p = 0; p != 2; ++p) {
    c1[p] = std::conj(m.v[p]) * g2;
    c2[p] = m.v[p] * g1;
  }
  num1 += DComplex(r.v[0] * c1[0] + r.v[1] * c1[1]);
  den1 += double(std::norm(c1[0]) + std::norm(c1[1]));
  num2 += DComplex(std::conj(r.v[0]) * c2[0] + std::conj(r.v[1]) * c2[1]);
  den2 += double(std::norm(c2[0]) + std::norm(c2[1]));
}
template <typename T>
inline void AccumulateScalar(const Full<T>& r, const Full<T>& m, std::complex<T> g1,
                             std::complex<T> g2, DComplex& num1, double& den1, DComplex& num2,
                             double& den2) {
  // C1 = M^H g_b = [conj(M0) conj(M2); conj(M1) conj(M3)] g_b, C2 = M g_a
  const std::complex<T> c1[4] = {std::conj(m.v[0]) * g2, std::conj(m.v[2]) * g2,
                                 std::conj(m.v[1]) * g2, std::conj(m.v[3]) * g2};
  const std::complex<T> c2[4] = {m.v[0] * g1, m.v[1] * g1, m.v[2] * g1, m.v[3] * g1};
  // Trace(R C1) = R0 C1_0 + R1 C1_2 + R2 C1_1 + R3 C1_3
  num1 += DComplex((r.v[0] * c1[0] + r.v[1] * c1[2]) + (r.v[2] * c1[1] + r.v[3] * c1[3]));
  den1 += double((std::norm(c1[0]) + std::norm(c1[1])) + (std::norm(c1[2]) + std::norm(c1[3])));
  // Trace(R^H C2), R^H = [conj(R0) conj(R2); conj(R1) conj(R3)]
  num2 += DComplex((std::conj(r.v[0]) * c2[0] + std::conj(r.v[2]) * c2[2]) +
                   (std::conj(r.v[1]) * c2[1] + std::conj(r.v[3]) * c2[3]));
  den2 += double((std::norm(c2[0]) + std::norm(c2[1])) + (std::norm(c2[2]) + std::norm(c2[3])));
}

template <typename V>
struct ChannelBlock {
  size_t n_vis = 0;
  std::vector<V> data;                // [n_vis]
  std::vector<V> model;               // [n_dir][n_vis]
  const uint32_t* ant1 = nullptr;     // [n_vis]
  const uint32_t* ant2 = nullptr;     // [n_vis]
  const uint32_t* sol_map = nullptr;  // [n_dir][n_vis]
};

template <typename V>
V LoadVis(const float* p) {
  using T = decltype(V().v[0].real());
  V r;
  for (int k = 0; k != V::kCorr; ++k)
    r.v[k] = std::complex<T>(static_cast<T>(p[2 * k]), static_cast<T>(p[2 * k + 1]));
  return r;
}

// Calls f(begin, end) for `n_parts` contiguous parts of [0, n), in parallel.
template <typename F>
void StaticFor(size_t n, size_t n_parts, F&& f) {
  n_parts = std::max<size_t>(1, std::min(n_parts, n));
  if (n_parts <= 1) {
    if (n) f(size_t(0), n, size_t(0));
    return;
  }
  std::vector<std::thread> threads;
  threads.reserve(n_parts);
  for (size_t i = 0; i != n_parts; ++i) {
    const size_t b = i * n / n_parts;
    const size_t e = (i + 1) * n / n_parts;
    threads.emplace_back([&f, b, e, i]() { f(b, e, i); });
  }
  for (std::thread& t : threads) t.join();
}

struct StallState {
  size_t n_var_count = 0;
  double step_mean = 0.0;
  double step_var = 0.0;
};

template <typename V>
class IterativeDiagonalOracle {
 public:
  using T = decltype(V().v[0].real());
  using C = std::complex<T>;

  IterativeDiagonalOracle(const orc_settings& s, int n_pol = 2)
      : s_(s),
        np_(n_pol),
        n_ant_(s.n_antennas),
        n_dir_(s.n_directions),
        n_cb_(s.n_channel_blocks),
        n_sol_dir_(s.n_solutions_per_direction,
                   s.n_solutions_per_direction + s.n_directions) {
    n_sub_ = std::accumulate(n_sol_dir_.begin(), n_sol_dir_.end(), size_t(0));
    sol_start_.resize(n_dir_);
    size_t acc = 0;
    for (size_t d = 0; d != n_dir_; ++d) {
      sol_start_[d] = acc;
      acc += n_sol_dir_[d];
    }
    const size_t n_threads = std::max<int32_t>(1, s.n_threads);
    n_sub_threads_ = std::max<size_t>(1, n_threads / n_cb_);
    n_block_threads_ = std::min<size_t>(n_threads, n_cb_);
  }

  size_t SolSize() const { return n_ant_ * n_sub_ * np_; }

  // IterativeDiagonalSolver.cc:100-127
  void PerformIteration(const ChannelBlock<V>& cb, std::vector<V>& residual,
                        const DComplex* solutions, DComplex* next) const {
    std::copy(cb.data.begin(), cb.data.end(), residual.begin());
    for (size_t d = 0; d != n_dir_; ++d)
      AddOrSubtract<false>(cb, residual, d, solutions);
    const std::vector<V> copy = residual;
    for (size_t d = 0; d != n_dir_; ++d) {
      if (d != 0) residual = copy;
      AddOrSubtract<true>(cb, residual, d, solutions);
      SolveDirection(cb, residual, d, solutions, next);
    }
  }

  // SolveData.h:204-244
  template <bool Add>
  void AddOrSubtract(const ChannelBlock<V>& cb, std::vector<V>& residual,
                     size_t d, const DComplex* solutions) const {
    StaticFor(cb.n_vis, n_sub_threads_, [&](size_t b, size_t e, size_t) {
      for (size_t v = b; v != e; ++v) {
        const uint32_t a1 = cb.ant1[v];
        const uint32_t a2 = cb.ant2[v];
        const uint32_t si = cb.sol_map[d * cb.n_vis + v];
        const DComplex* s1 = &solutions[(a1 * n_sub_ + si) * np_];
        const DComplex* s2 = &solutions[(a2 * n_sub_ + si) * np_];
        if constexpr (V::kCorr == 4) {
          if (np_ == 4) {
            // solution_1 * model.MultiplyHerm(solution_2)
            // (IterativeFullJonesSolver.cc:196-212), in float
            const Mat2<T> term = Mul(ToMat(s1), MultiplyHerm(ToMat(cb.model[d * cb.n_vis + v]),
                                                             ToMat(s2)));
            for (int k = 0; k != 4; ++k) {
              if (Add)
                residual[v].v[k] += term.m[k];
              else
                residual[v].v[k] -= term.m[k];
            }
            continue;
          }
        }
        // scalar: model * g_a * conj(g_b) per element (IterativeScalarSolver.cc:
        // 209-222), the same products as Diag(g_a) M Diag(g_b)^H with equal pols
        const C ga[2] = {static_cast<C>(s1[0]), static_cast<C>(s1[np_ - 1])};
        const C gb[2] = {static_cast<C>(s2[0]), static_cast<C>(s2[np_ - 1])};
        const V contribution = Contribution(ga, cb.model[d * cb.n_vis + v], gb);
        if (Add)
          AddTo(residual[v], contribution);
        else
          SubFrom(residual[v], contribution);
      }
    });
  }

  static Mat2<T> ToMat(const DComplex* s) {
    return {{C(s[0]), C(s[1]), C(s[2]), C(s[3])}};
  }
  static Mat2<T> ToMat(const V& x) {
    Mat2<T> r;
    for (int k = 0; k != 4; ++k) r.m[k] = x.v[k];
    return r;
  }

  // IterativeFullJonesSolver.cc:99-188: per antenna and solution the float
  // sums numerator = sum D J_b M^H (or D^H J_a M) and denominator =
  // sum (J_b M^H)^H (J_b M^H) (or (J_a M)^H (J_a M)); the solution is
  // numerator * denominator^-1, NaN when the denominator is singular.
  void SolveDirectionFullJones(const ChannelBlock<V>& cb, const std::vector<V>& residual,
                               size_t d, const DComplex* solutions, DComplex* next) const {
    const size_t n_dir_sol = n_sol_dir_[d];
    const uint32_t sol0 = cb.n_vis ? cb.sol_map[d * cb.n_vis] : sol_start_[d];
    const size_t n_parts = std::max<size_t>(1, std::min(n_sub_threads_, cb.n_vis));
    const Mat2<T> zero{};
    std::vector<std::vector<Mat2<T>>> num(n_parts,
                                          std::vector<Mat2<T>>(n_ant_ * n_dir_sol, zero));
    std::vector<std::vector<Mat2<T>>> den(n_parts,
                                          std::vector<Mat2<T>>(n_ant_ * n_dir_sol, zero));
    StaticFor(cb.n_vis, n_parts, [&](size_t b, size_t e, size_t part) {
      Mat2<T>* lnum = num[part].data();
      Mat2<T>* lden = den[part].data();
      for (size_t v = b; v != e; ++v) {
        const uint32_t a1 = cb.ant1[v];
        const uint32_t a2 = cb.ant2[v];
        const uint32_t si = cb.sol_map[d * cb.n_vis + v];
        const Mat2<T> j1 = ToMat(&solutions[(a1 * n_sub_ + si) * 4]);
        const Mat2<T> j2 = ToMat(&solutions[(a2 * n_sub_ + si) * 4]);
        const Mat2<T> data = ToMat(residual[v]);
        const Mat2<T> model = ToMat(cb.model[d * cb.n_vis + v]);
        const uint32_t rel = si - sol0;
        const Mat2<T> cor_model_herm_1 = MultiplyHerm(j2, model);
        AddTo(lnum[a1 * n_dir_sol + rel], Mul(data, cor_model_herm_1));
        AddTo(lden[a1 * n_dir_sol + rel], Mul(HermTranspose(cor_model_herm_1), cor_model_herm_1));
        const Mat2<T> cor_model_2 = Mul(j1, model);
        AddTo(lnum[a2 * n_dir_sol + rel], Mul(HermTranspose(data), cor_model_2));
        AddTo(lden[a2 * n_dir_sol + rel], Mul(HermTranspose(cor_model_2), cor_model_2));
      }
    });
    for (size_t part = 1; part < n_parts; ++part)
      for (size_t i = 0; i != num[0].size(); ++i) {
        AddTo(num[0][i], num[part][i]);
        AddTo(den[0][i], den[part][i]);
      }
    const T nan = std::numeric_limits<T>::quiet_NaN();
    for (size_t ant = 0; ant != n_ant_; ++ant)
      for (size_t rel = 0; rel != n_dir_sol; ++rel) {
        const size_t index = ant * n_dir_sol + rel;
        Mat2<T> inv = den[0][index];
        Mat2<T> result;
        if (Invert(inv))
          result = Mul(num[0][index], inv);
        else
          for (int k = 0; k != 4; ++k) result.m[k] = std::complex<T>(nan, nan);
        DComplex* out = &next[(ant * n_sub_ + rel + sol0) * 4];
        for (int k = 0; k != 4; ++k) out[k] = DComplex(result.m[k]);
      }
  }

  // IterativeScalarSolver.cc:117-194
  void SolveDirectionScalar(const ChannelBlock<V>& cb, const std::vector<V>& residual,
                            size_t d, const DComplex* solutions, DComplex* next) const {
    const size_t n_dir_sol = n_sol_dir_[d];
    const uint32_t sol0 = cb.n_vis ? cb.sol_map[d * cb.n_vis] : sol_start_[d];
    const size_t n_parts = std::max<size_t>(1, std::min(n_sub_threads_, cb.n_vis));
    std::vector<std::vector<DComplex>> num(n_parts, std::vector<DComplex>(n_ant_ * n_dir_sol));
    std::vector<std::vector<double>> den(n_parts, std::vector<double>(n_ant_ * n_dir_sol));
    StaticFor(cb.n_vis, n_parts, [&](size_t b, size_t e, size_t part) {
      DComplex* lnum = num[part].data();
      double* lden = den[part].data();
      for (size_t v = b; v != e; ++v) {
        const uint32_t a1 = cb.ant1[v];
        const uint32_t a2 = cb.ant2[v];
        const uint32_t si = cb.sol_map[d * cb.n_vis + v];
        const C g1 = static_cast<C>(solutions[a1 * n_sub_ + si]);
        const C g2 = static_cast<C>(solutions[a2 * n_sub_ + si]);
        const uint32_t rel = si - sol0;
        const size_t i1 = a1 * n_dir_sol + rel, i2 = a2 * n_dir_sol + rel;
        AccumulateScalar(residual[v], cb.model[d * cb.n_vis + v], g1, g2, lnum[i1], lden[i1],
                         lnum[i2], lden[i2]);
      }
    });
    for (size_t part = 1; part < n_parts; ++part)
      for (size_t i = 0; i != num[0].size(); ++i) {
        num[0][i] += num[part][i];
        den[0][i] += den[part][i];
      }
    for (size_t ant = 0; ant != n_ant_; ++ant)
      for (size_t rel = 0; rel != n_dir_sol; ++rel) {
        const size_t index = ant * n_dir_sol + rel;
        DComplex& out = next[ant * n_sub_ + rel + sol0];
        // a zero denominator gives NaN (IterativeScalarSolver.cc:184-189)
        out = den[0][index] == 0.0 ? DComplex(std::numeric_limits<float>::quiet_NaN(), 0.0)
                                   : num[0][index] / den[0][index];
      }
  }

  // IterativeDiagonalSolver.cc:130-222
  void SolveDirection(const ChannelBlock<V>& cb, const std::vector<V>& residual,
                      size_t d, const DComplex* solutions,
                      DComplex* next) const {
    if (np_ == 1) {
      SolveDirectionScalar(cb, residual, d, solutions, next);
      return;
    }
    if constexpr (V::kCorr == 4) {
      if (np_ == 4) {
        SolveDirectionFullJones(cb, residual, d, solutions, next);
        return;
      }
    }
    if constexpr (V::kCorr >= 2) SolveDirectionDiagonal(cb, residual, d, solutions, next);
  }

  // IterativeDiagonalSolver.cc:130-222
  void SolveDirectionDiagonal(const ChannelBlock<V>& cb, const std::vector<V>& residual,
                              size_t d, const DComplex* solutions, DComplex* next) const {
    const size_t n_dir_sol = n_sol_dir_[d];
    const uint32_t sol0 = cb.n_vis ? cb.sol_map[d * cb.n_vis] : sol_start_[d];
    const size_t n_parts =
        std::max<size_t>(1, std::min(n_sub_threads_, cb.n_vis));
    std::vector<std::vector<C>> num(n_parts,
                                    std::vector<C>(n_ant_ * n_dir_sol * 2));
    std::vector<std::vector<T>> den(n_parts,
                                    std::vector<T>(n_ant_ * n_dir_sol * 2));
    StaticFor(cb.n_vis, n_parts, [&](size_t b, size_t e, size_t part) {
      C* lnum = num[part].data();
      T* lden = den[part].data();
      for (size_t v = b; v != e; ++v) {
        const uint32_t a1 = cb.ant1[v];
        const uint32_t a2 = cb.ant2[v];
        const uint32_t si = cb.sol_map[d * cb.n_vis + v];
        const DComplex* s1 = &solutions[(a1 * n_sub_ + si) * 2];
        const DComplex* s2 = &solutions[(a2 * n_sub_ + si) * 2];
        const V& r = residual[v];
        const V& m = cb.model[d * cb.n_vis + v];
        const uint32_t rel = si - sol0;
        const C g2[2] = {static_cast<C>(s2[0]), static_cast<C>(s2[1])};
        const size_t i1 = a1 * n_dir_sol + rel;
        AccumulateAnt1(r, m, g2, &lnum[i1 * 2], &lden[i1 * 2]);
        const C g1[2] = {static_cast<C>(s1[0]), static_cast<C>(s1[1])};
        const size_t i2 = a2 * n_dir_sol + rel;
        AccumulateAnt2(r, m, g1, &lnum[i2 * 2], &lden[i2 * 2]);
      }
    });
    // Merge in range order.
    for (size_t part = 1; part < n_parts; ++part)
      for (size_t i = 0; i != num[0].size(); ++i) {
        num[0][i] += num[part][i];
        den[0][i] += den[part][i];
      }
    for (size_t ant = 0; ant != n_ant_; ++ant) {
      for (size_t rel = 0; rel != n_dir_sol; ++rel) {
        const size_t si = rel + sol0;
        const size_t index = ant * n_dir_sol + rel;
        for (size_t pol = 0; pol != 2; ++pol) {
          DComplex& out = next[(ant * n_sub_ + si) * 2 + pol];
          if (den[0][index * 2 + pol] == T(0))
            out = DComplex(std::numeric_limits<double>::quiet_NaN(), 0.0);
          else
            out = DComplex(num[0][index * 2 + pol]) /
                  double(den[0][index * 2 + pol]);
        }
      }
    }
  }

  // SolverBase.cc:342-371 (note: pol-0 magnitude is used for both pols, :354)
  static void MakeSolutionsFinite2Pol(std::vector<DComplex>& sv) {
    size_t count = 0;
    double average[2] = {0.0, 0.0};
    for (size_t i = 0; i < sv.size(); i += 2) {
      if (IsFinite(sv[i]) && IsFinite(sv[i + 1])) {
        for (size_t p = 0; p != 2; ++p) average[p] += std::abs(sv[i]);
        ++count;
      }
    }
    if (count == 0) {
      average[0] = 1.0;
      average[1] = 1.0;
    } else {
      for (size_t p = 0; p != 2; ++p) average[p] /= count;
    }
    for (size_t i = 0; i < sv.size(); i += 2) {
      if (!IsFinite(sv[i]) || !IsFinite(sv[i + 1])) {
        for (size_t p = 0; p != 2; ++p) sv[i + p] = average[p];
      }
    }
  }

  // SolverBase.cc:308-333
  static void MakeSolutionsFinite1Pol(std::vector<DComplex>& sv) {
    size_t count = 0;
    double average = 0.0;
    for (const DComplex& v : sv)
      if (IsFinite(v)) {
        average += std::abs(v);
        ++count;
      }
    average = count == 0 ? 1.0 : average / count;
    for (DComplex& v : sv)
      if (!IsFinite(v)) v = average;
  }

  // SolverBase.cc:373-405. As in the reference, the average of every
  // polarization accumulates |entry 0| (:384), so a non-finite matrix becomes
  // the mean |J_00| in all four entries; with no finite matrix, the identity.
  static void MakeSolutionsFinite4Pol(std::vector<DComplex>& sv) {
    size_t count = 0;
    double average[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t i = 0; i < sv.size(); i += 4) {
      if (IsFinite(sv[i]) && IsFinite(sv[i + 1]) && IsFinite(sv[i + 2]) && IsFinite(sv[i + 3])) {
        for (size_t p = 0; p != 4; ++p) average[p] += std::abs(sv[i]);
        ++count;
      }
    }
    if (count == 0) {
      average[0] = 1.0;
      average[1] = 0.0;
      average[2] = 0.0;
      average[3] = 1.0;
    } else {
      for (size_t p = 0; p != 4; ++p) average[p] /= count;
    }
    for (size_t i = 0; i < sv.size(); i += 4) {
      if (!(IsFinite(sv[i]) && IsFinite(sv[i + 1]) && IsFinite(sv[i + 2]) && IsFinite(sv[i + 3])))
        for (size_t p = 0; p != 4; ++p) sv[i + p] = average[p];
    }
  }

  static bool IsFinite(const DComplex& c) {
    return std::isfinite(c.real()) && std::isfinite(c.imag());
  }

  // SolverBase.cc:88-158
  void Step(const std::vector<std::vector<DComplex>>& solutions,
            std::vector<DComplex>& next) const {
    const size_t n = SolSize();
    const double step = s_.step_size;
    for (size_t cb = 0; cb != n_cb_; ++cb) {
      const DComplex* sol = solutions[cb].data();
      DComplex* nxt = &next[cb * n];
      if (!s_.phase_only) {
        for (size_t i = 0; i != n; ++i)
          nxt[i] = sol[i] * (1.0 - step) + nxt[i] * step;
      } else {
        for (size_t i = 0; i != n; ++i) {
          const double phase_from = std::arg(sol[i]);
          double distance = std::arg(nxt[i]) - phase_from;
          if (distance > M_PI)
            distance -= 2.0 * M_PI;
          else if (distance < -M_PI)
            distance += 2.0 * M_PI;
          nxt[i] = std::polar(1.0, phase_from + step * distance);
        }
      }
    }
  }

  // SolverBase.cc:215-307 for two solution polarizations.
  bool AssignSolutions(std::vector<std::vector<DComplex>>& solutions,
                       const std::vector<DComplex>& next,
                       bool use_constraint_accuracy,
                       std::vector<double>& step_magnitudes) const {
    double avg_abs_diff = 0.0;
    size_t n = 0;
    const size_t size = SolSize();
    for (size_t cb = 0; cb != n_cb_; ++cb) {
      const DComplex* ns_cb = &next[cb * size];
      if (np_ == 1) {
        // one solution polarization (SolverBase.cc:244-253)
        for (size_t i = 0; i != size; ++i) {
          const DComplex s0 = solutions[cb][i];
          if (s0 != 0.0) {
            const double a = std::abs((ns_cb[i] - s0) / s0);
            if (std::isfinite(a)) {
              avg_abs_diff += a;
              ++n;
            }
          }
        }
        std::copy_n(ns_cb, size, solutions[cb].data());
        continue;
      }
      if (np_ == 4) {
        // four solution polarizations (SolverBase.cc:269-283): the entries of
        // (new - old) * old^-1, skipped when old is singular
        for (size_t i = 0; i != size; i += 4) {
          const Mat2<double> s{{solutions[cb][i], solutions[cb][i + 1], solutions[cb][i + 2],
                                solutions[cb][i + 3]}};
          Mat2<double> s_inv = s;
          if (Invert(s_inv)) {
            Mat2<double> ns{{ns_cb[i] - s.m[0], ns_cb[i + 1] - s.m[1], ns_cb[i + 2] - s.m[2],
                             ns_cb[i + 3] - s.m[3]}};
            ns = Mul(ns, s_inv);
            const double sumabs =
                std::abs(ns.m[0]) + std::abs(ns.m[1]) + std::abs(ns.m[2]) + std::abs(ns.m[3]);
            if (std::isfinite(sumabs)) {
              avg_abs_diff += sumabs;
              n += 4;
            }
          }
        }
        std::copy_n(ns_cb, size, solutions[cb].data());
        continue;
      }
      for (size_t i = 0; i != size; i += 2) {
        const DComplex s[2] = {solutions[cb][i], solutions[cb][i + 1]};
        const DComplex s_inv[2] = {1.0 / s[0], 1.0 / s[1]};
        if (IsFinite(s_inv[0]) && IsFinite(s_inv[1])) {
          DComplex ns[2] = {ns_cb[i], ns_cb[i + 1]};
          ns[0] = (ns[0] - s[0]) * s_inv[0];
          ns[1] = (ns[1] - s[1]) * s_inv[1];
          const double sumabs = std::abs(ns[0]) + std::abs(ns[1]);
          if (std::isfinite(sumabs)) {
            avg_abs_diff += sumabs;
            n += 2;
          }
        }
      }
      std::copy_n(ns_cb, size, solutions[cb].data());
    }
    const double step_magnitude =
        (n == 0 ? 0 : avg_abs_diff / s_.step_size / n);
    step_magnitudes.push_back(step_magnitude);
    if (use_constraint_accuracy)
      return step_magnitude <= s_.constraint_accuracy;
    return step_magnitude <= s_.accuracy;
  }

  // SolverBase.cc:63-86
  bool DetectStall(size_t iteration, const std::vector<double>& step_mags) {
    if (iteration == 1) {
      stall_.n_var_count = 1;
      stall_.step_mean = step_mags[iteration - 1];
      stall_.step_var = 0.0;
    } else {
      stall_.n_var_count += 1;
      const double delta_1 = step_mags[iteration - 1] - stall_.step_mean;
      stall_.step_mean += delta_1 / (double)stall_.n_var_count;
      const double delta_2 = step_mags[iteration - 1] - stall_.step_mean;
      stall_.step_var += delta_1 * delta_2;
    }
    if (iteration < 30) return false;
    return stall_.step_mean <
           s_.step_diff_sigma *
               std::sqrt(stall_.step_var / (double)stall_.n_var_count);
  }

  // SolverBase.h:274-285
  bool ReachedStoppingCriterion(size_t iteration, bool has_converged,
                                bool constraints_satisfied,
                                const std::vector<double>& step_mags) {
    bool has_stalled = false;
    if (s_.detect_stalling && constraints_satisfied)
      has_stalled = DetectStall(iteration, step_mags);
    const bool is_ready = iteration >= size_t(s_.max_iterations) ||
                          (has_converged && constraints_satisfied) ||
                          has_stalled;
    return iteration >= size_t(s_.min_iterations) && is_ready;
  }

  // IterativeDiagonalSolver.cc:38-97 and SolverBase.cc:192-213.
  void Solve(const std::vector<ChannelBlock<V>>& cbs, double* sol_flat,
             orc_result* result, double* step_mags_out,
             const orc_smoothness* smoothness = nullptr) {
    const size_t size = SolSize();
    std::vector<std::vector<DComplex>> solutions(n_cb_);
    for (size_t cb = 0; cb != n_cb_; ++cb)
      solutions[cb].assign(reinterpret_cast<DComplex*>(sol_flat) + cb * size,
                           reinterpret_cast<DComplex*>(sol_flat) + (cb + 1) * size);
    std::vector<DComplex> next(n_cb_ * size);
    std::vector<std::vector<V>> residual(n_cb_);
    for (size_t cb = 0; cb != n_cb_; ++cb) residual[cb].resize(cbs[cb].n_vis);

    size_t iteration = 0;
    bool has_converged = false;
    bool has_previously_converged = false;
    bool constraints_satisfied = false;
    std::vector<double> step_magnitudes;
    step_magnitudes.reserve(s_.max_iterations + 1);
    do {
      for (std::vector<DComplex>& sv : solutions) {
        if (np_ == 1)
          MakeSolutionsFinite1Pol(sv);
        else if (np_ == 4)
          MakeSolutionsFinite4Pol(sv);
        else
          MakeSolutionsFinite2Pol(sv);
      }
      StaticFor(n_cb_, n_block_threads_, [&](size_t b, size_t e, size_t) {
        for (size_t cb = b; cb < e; ++cb)
          PerformIteration(cbs[cb], residual[cb], solutions[cb].data(),
                           &next[cb * size]);
      });
      Step(solutions, next);
      // ApplyConstraints (SolverBase.cc:174-190): the smoothness constraint
      // is always satisfied; without constraints the same holds.
      if (smoothness)
        orc_smooth(smoothness, int32_t(n_cb_), int32_t(n_ant_), int32_t(n_sub_), int32_t(np_),
                   reinterpret_cast<double*>(next.data()));
      constraints_satisfied = true;
      has_converged = AssignSolutions(solutions, next, !constraints_satisfied,
                                      step_magnitudes);
      iteration++;
      has_previously_converged = has_converged || has_previously_converged;
    } while (!ReachedStoppingCriterion(iteration, has_converged,
                                       constraints_satisfied, step_magnitudes));

    for (size_t cb = 0; cb != n_cb_; ++cb)
      std::copy(solutions[cb].begin(), solutions[cb].end(),
                reinterpret_cast<DComplex*>(sol_flat) + cb * size);
    result->iterations = (has_converged && constraints_satisfied)
                             ? int64_t(iteration)
                             : int64_t(iteration + 1);
    result->constraint_iterations =
        constraints_satisfied ? 0 : int64_t(iteration + 1);
    result->has_converged = has_converged ? 1 : 0;
    result->n_step_magnitudes = int32_t(step_magnitudes.size());
    if (step_mags_out)
      std::copy(step_magnitudes.begin(), step_magnitudes.end(), step_mags_out);
  }

  void IterateOnce(const std::vector<ChannelBlock<V>>& cbs,
                   const double* sol_flat, double* next_flat) const {
    const size_t size = SolSize();
    StaticFor(n_cb_, n_block_threads_, [&](size_t b, size_t e, size_t) {
      for (size_t cb = b; cb < e; ++cb) {
        std::vector<V> residual(cbs[cb].n_vis);
        PerformIteration(cbs[cb], residual,
                         reinterpret_cast<const DComplex*>(sol_flat) + cb * size,
                         reinterpret_cast<DComplex*>(next_flat) + cb * size);
      }
    });
  }

 private:
  orc_settings s_;
  size_t np_;  // solution polarizations: 2 (diagonal), 1 (scalar) or 4 (full-Jones)
  size_t n_ant_, n_dir_, n_cb_, n_sub_;
  std::vector<size_t> n_sol_dir_;
  std::vector<size_t> sol_start_;
  size_t n_sub_threads_, n_block_threads_;
  StallState stall_;
};

template <typename V>
std::vector<ChannelBlock<V>> MakeBlocks(const orc_settings& s,
                                        const int64_t* vis_offsets,
                                        const float* data, const float* model,
                                        const uint32_t* ant1,
                                        const uint32_t* ant2,
                                        const uint32_t* sol_map) {
  const size_t n_dir = s.n_directions;
  const size_t nc = V::kCorr;
  std::vector<ChannelBlock<V>> cbs(s.n_channel_blocks);
  for (size_t cb = 0; cb != cbs.size(); ++cb) {
    const size_t v0 = vis_offsets[cb];
    const size_t nv = vis_offsets[cb + 1] - v0;
    ChannelBlock<V>& b = cbs[cb];
    b.n_vis = nv;
    b.data.resize(nv);
    b.model.resize(n_dir * nv);
    for (size_t v = 0; v != nv; ++v) b.data[v] = LoadVis<V>(&data[(v0 + v) * nc * 2]);
    for (size_t i = 0; i != n_dir * nv; ++i)
      b.model[i] = LoadVis<V>(&model[(v0 * n_dir + i) * nc * 2]);
    b.ant1 = ant1 + v0;
    b.ant2 = ant2 + v0;
    b.sol_map = sol_map + v0 * n_dir;
  }
  return cbs;
}

template <typename V>
int SolveTyped(const orc_settings* s, const int64_t* vo, const float* data,
               const float* model, const uint32_t* a1, const uint32_t* a2,
               const uint32_t* sm, double* solutions, orc_result* result,
               double* step_mags, const orc_smoothness* smoothness = nullptr,
               int n_pol = 2) {
  const std::vector<ChannelBlock<V>> cbs =
      MakeBlocks<V>(*s, vo, data, model, a1, a2, sm);
  IterativeDiagonalOracle<V> solver(*s, n_pol);
  solver.Solve(cbs, solutions, result, step_mags, smoothness);
  return 0;
}

template <typename V>
int IterateTyped(const orc_settings* s, const int64_t* vo, const float* data,
                 const float* model, const uint32_t* a1, const uint32_t* a2,
                 const uint32_t* sm, const double* solutions, double* next) {
  const std::vector<ChannelBlock<V>> cbs =
      MakeBlocks<V>(*s, vo, data, model, a1, a2, sm);
  IterativeDiagonalOracle<V> solver(*s);
  solver.IterateOnce(cbs, solutions, next);
  return 0;
}

}  // namespace

extern "C" int orc_solve(const orc_settings* s, const int64_t* vo,
                         const float* data, const float* model,
                         const uint32_t* a1, const uint32_t* a2,
                         const uint32_t* sm, double* solutions,
                         orc_result* result, double* step_mags) {
  if (!s || s->n_directions <= 0 || s->n_channel_blocks <= 0 ||
      s->n_antennas <= 0)
    return -1;
  if (s->n_correlations == 2) {
    if (s->fp64)
      return SolveTyped<Duo<double>>(s, vo, data, model, a1, a2, sm, solutions,
                                     result, step_mags);
    return SolveTyped<Duo<float>>(s, vo, data, model, a1, a2, sm, solutions,
                                  result, step_mags);
  }
  if (s->n_correlations == 4) {
    if (s->fp64)
      return SolveTyped<Full<double>>(s, vo, data, model, a1, a2, sm, solutions,
                                      result, step_mags);
    return SolveTyped<Full<float>>(s, vo, data, model, a1, a2, sm, solutions,
                                   result, step_mags);
  }
  return -2;
}

extern "C" int orc_solve_constrained(const orc_settings* s, const orc_smoothness* c,
                                     const int64_t* vo, const float* data, const float* model,
                                     const uint32_t* a1, const uint32_t* a2, const uint32_t* sm,
                                     double* solutions, orc_result* result, double* step_mags) {
  if (!s || !c || !c->frequencies || !(c->bandwidth_hz > 0.0) || s->n_directions <= 0 ||
      s->n_channel_blocks <= 0 || s->n_antennas <= 0)
    return -1;
  if (s->n_correlations == 2)
    return s->fp64 ? SolveTyped<Duo<double>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                             step_mags, c)
                   : SolveTyped<Duo<float>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                            step_mags, c);
  if (s->n_correlations == 4)
    return s->fp64 ? SolveTyped<Full<double>>(s, vo, data, model, a1, a2, sm, solutions,
                                              result, step_mags, c)
                   : SolveTyped<Full<float>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                             step_mags, c);
  return -2;
}

extern "C" int orc_solve_scalar(const orc_settings* s, const orc_smoothness* c,
                                const int64_t* vo, const float* data, const float* model,
                                const uint32_t* a1, const uint32_t* a2, const uint32_t* sm,
                                double* solutions, orc_result* result, double* step_mags) {
  if (!s || s->n_directions <= 0 || s->n_channel_blocks <= 0 || s->n_antennas <= 0) return -1;
  if (c && (!c->frequencies || !(c->bandwidth_hz > 0.0))) return -1;
  if (s->n_correlations == 1)
    return s->fp64 ? SolveTyped<Uni<double>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                             step_mags, c, 1)
                   : SolveTyped<Uni<float>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                            step_mags, c, 1);
  if (s->n_correlations == 2)
    return s->fp64 ? SolveTyped<Duo<double>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                             step_mags, c, 1)
                   : SolveTyped<Duo<float>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                            step_mags, c, 1);
  if (s->n_correlations == 4)
    return s->fp64 ? SolveTyped<Full<double>>(s, vo, data, model, a1, a2, sm, solutions,
                                              result, step_mags, c, 1)
                   : SolveTyped<Full<float>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                             step_mags, c, 1);
  return -2;
}

extern "C" int orc_solve_full_jones(const orc_settings* s, const orc_smoothness* c,
                                    const int64_t* vo, const float* data, const float* model,
                                    const uint32_t* a1, const uint32_t* a2, const uint32_t* sm,
                                    double* solutions, orc_result* result, double* step_mags) {
  if (!s || s->n_directions <= 0 || s->n_channel_blocks <= 0 || s->n_antennas <= 0) return -1;
  if (c && (!c->frequencies || !(c->bandwidth_hz > 0.0))) return -1;
  // the full-Jones solver takes FullSolveData only (IterativeFullJonesSolver.h)
  if (s->n_correlations != 4) return -2;
  return s->fp64 ? SolveTyped<Full<double>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                            step_mags, c, 4)
                 : SolveTyped<Full<float>>(s, vo, data, model, a1, a2, sm, solutions, result,
                                           step_mags, c, 4);
}

extern "C" int orc_iterate_once(const orc_settings* s, const int64_t* vo,
                                const float* data, const float* model,
                                const uint32_t* a1, const uint32_t* a2,
                                const uint32_t* sm, const double* solutions,
                                double* next) {
  if (s->n_correlations == 2)
    return s->fp64 ? IterateTyped<Duo<double>>(s, vo, data, model, a1, a2, sm,
                                                solutions, next)
                   : IterateTyped<Duo<float>>(s, vo, data, model, a1, a2, sm,
                                              solutions, next);
  if (s->n_correlations == 4)
    return s->fp64 ? IterateTyped<Full<double>>(s, vo, data, model, a1, a2, sm,
                                                 solutions, next)
                   : IterateTyped<Full<float>>(s, vo, data, model, a1, a2, sm,
                                               solutions, next);
  return -2;
}

// Build-defined chi^2 (the reference computes none): sum over blocks,
// visibilities and correlations of |D - sum_d Diag(g_a) M_d Diag(g_b)^H|^2,
// everything in double.
extern "C" double orc_chi2(int32_t n_ant, int32_t n_dir, int32_t n_cb,
                           int32_t n_corr, int32_t n_sub, const int64_t* vo,
                           const float* data, const float* model,
                           const uint32_t* ant1, const uint32_t* ant2,
                           const uint32_t* sol_map, const double* solutions) {
  (void)n_ant;
  const DComplex* sol = reinterpret_cast<const DComplex*>(solutions);
  const size_t size = size_t(n_ant) * n_sub * 2;
  double chi2 = 0.0;
  for (int32_t cb = 0; cb != n_cb; ++cb) {
    const size_t v0 = vo[cb];
    const size_t nv = vo[cb + 1] - v0;
    const DComplex* s = sol + cb * size;
    for (size_t v = 0; v != nv; ++v) {
      DComplex r[4];
      for (int k = 0; k != n_corr; ++k)
        r[k] = DComplex(data[((v0 + v) * n_corr + k) * 2],
                        data[((v0 + v) * n_corr + k) * 2 + 1]);
      const uint32_t a1 = ant1[v0 + v], a2 = ant2[v0 + v];
      for (int32_t d = 0; d != n_dir; ++d) {
        const uint32_t si = sol_map[v0 * n_dir + size_t(d) * nv + v];
        const DComplex* ga = &s[(a1 * size_t(n_sub) + si) * 2];
        const DComplex* gb = &s[(a2 * size_t(n_sub) + si) * 2];
        const float* m = &model[(v0 * n_dir + size_t(d) * nv + v) * n_corr * 2];
        for (int k = 0; k != n_corr; ++k) {
          const int i = (n_corr == 2) ? k : k / 2;
          const int j = (n_corr == 2) ? k : k % 2;
          r[k] -= ga[i] * DComplex(m[2 * k], m[2 * k + 1]) * std::conj(gb[j]);
        }
      }
      for (int k = 0; k != n_corr; ++k) chi2 += std::norm(r[k]);
    }
  }
  return chi2;
}

// orc_chi2 for full-Jones solutions ([cb][ant][sub][4]):
// |D - sum_d J_a M_d J_b^H|^2, everything in double.
extern "C" double orc_chi2_full_jones(int32_t n_ant, int32_t n_dir, int32_t n_cb,
                                      int32_t n_corr, int32_t n_sub, const int64_t* vo,
                                      const float* data, const float* model,
                                      const uint32_t* ant1, const uint32_t* ant2,
                                      const uint32_t* sol_map, const double* solutions) {
  if (n_corr != 4) return std::numeric_limits<double>::quiet_NaN();
  const DComplex* sol = reinterpret_cast<const DComplex*>(solutions);
  const size_t size = size_t(n_ant) * n_sub * 4;
  double chi2 = 0.0;
  for (int32_t cb = 0; cb != n_cb; ++cb) {
    const size_t v0 = vo[cb];
    const size_t nv = vo[cb + 1] - v0;
    const DComplex* s = sol + cb * size;
    for (size_t v = 0; v != nv; ++v) {
      Mat2<double> r;
      for (int k = 0; k != 4; ++k)
        r.m[k] = DComplex(data[((v0 + v) * 4 + k) * 2], data[((v0 + v) * 4 + k) * 2 + 1]);
      const uint32_t a1 = ant1[v0 + v], a2 = ant2[v0 + v];
      for (int32_t d = 0; d != n_dir; ++d) {
        const uint32_t si = sol_map[v0 * n_dir + size_t(d) * nv + v];
        const DComplex* ja = &s[(a1 * size_t(n_sub) + si) * 4];
        const DComplex* jb = &s[(a2 * size_t(n_sub) + si) * 4];
        const float* m = &model[(v0 * n_dir + size_t(d) * nv + v) * 8];
        const Mat2<double> mm{{DComplex(m[0], m[1]), DComplex(m[2], m[3]), DComplex(m[4], m[5]),
                               DComplex(m[6], m[7])}};
        const Mat2<double> term = Mul(Mat2<double>{{ja[0], ja[1], ja[2], ja[3]}},
                                      MultiplyHerm(mm, Mat2<double>{{jb[0], jb[1], jb[2], jb[3]}}));
        for (int k = 0; k != 4; ++k) r.m[k] -= term.m[k];
      }
      for (int k = 0; k != 4; ++k) chi2 += std::norm(r.m[k]);
    }
  }
  return chi2;
}
