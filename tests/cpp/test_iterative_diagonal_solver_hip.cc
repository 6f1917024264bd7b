// C++ tests of dp3amd::IterativeDiagonalSolverHip through its SolverBase
// interface, following DP3's own solver tests (ddecal/test/unit/tSolvers.cc:
// 235-310, iterative_[duo_]diagonal_dd_intervals; tSolvers.cc:512-529 for
// min_iterations) with the SolverTester fixture (SolverTester.cc:95-202,
// 293-311, 409-439: 50 antennas, 3 directions, 16 channels in 4 blocks, 50
// time steps, accuracy 1e-8, step 0.2, at most 100 iterations).
//
// The fixture data and the CPU restatement of the reference solver come from
// the oracle library (oracle/liboracle.so, test infrastructure only). Needs a
// GPU. Exit status 0 when every check passes.
#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../dp3_amd/cpp/iterative_diagonal_solver_hip.h"
#include "../../oracle/oracle.h"

using dp3amd::DuoSolveData;
using dp3amd::FullSolveData;
using dp3amd::IterativeDiagonalSolverHip;
using dp3amd::IterativeFullJonesSolverHip;
using dp3amd::IterativeScalarSolverHip;
using dp3amd::SolverBase;
using dp3amd::UniSolveData;
using DComplex = std::complex<double>;

namespace {

int g_failures = 0;
int g_checks = 0;

#define CHECK(cond)                                                           \
  do {                                                                        \
    ++g_checks;                                                               \
    if (!(cond)) {                                                            \
      ++g_failures;                                                           \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
    }                                                                         \
  } while (0)

// BOOST_CHECK_CLOSE: |a - b| within tolerance percent of both values.
bool Close(double a, double b, double tolerance_percent) {
  const double d = std::abs(a - b);
  return d <= tolerance_percent / 100.0 * std::abs(a) &&
         d <= tolerance_percent / 100.0 * std::abs(b);
}

struct Fixture {
  static constexpr size_t kMaxIterations = 100;
  static constexpr double kAccuracy = 1e-8;
  static constexpr double kStepSize = 0.2;
  int n_ant, n_dir, n_ch, n_t, n_cb;
  std::vector<int32_t> n_sol;
  std::vector<int64_t> offsets;
  std::vector<float> data, model;
  std::vector<uint32_t> ant1, ant2, sol_map;
  std::vector<double> input;  // [n_ant][n_sub][2]

  int n_pol = 2;  // solution polarizations of the solver under test

  // n_solution_pols: 2 (diagonal), 1 (scalar fixture) or 4 (full-Jones
  // solver on the diagonal fixture)
  Fixture(std::vector<int32_t> n_solutions, int n_corr, int n_solution_pols = 2)
      : n_sol(std::move(n_solutions)), n_pol(n_solution_pols) {
    const bool scalar = n_pol == 1;
    orc_solvertester_shape(&n_ant, &n_dir, &n_ch, &n_t, &n_cb);
    const int n_bl = n_ant * (n_ant - 1) / 2;
    int n_sub = 0;
    for (int s : n_sol) n_sub += s;
    std::vector<float> raw(size_t(n_t) * n_bl * n_ch * 8);
    std::vector<float> raw_model(raw.size() * n_dir);
    input.resize(size_t(n_ant) * n_sub * 4);
    // SetScalarSolutions or SetDiagonalSolutions (SolverTester.cc:331-363)
    if (scalar)
      orc_solvertester_fill_scalar(n_sol.data(), raw.data(), raw_model.data(), input.data());
    else
      orc_solvertester_fill(n_sol.data(), raw.data(), raw_model.data(), input.data());
    std::vector<int32_t> a1, a2;
    for (int a = 0; a < n_ant; ++a)
      for (int b = a + 1; b < n_ant; ++b) {
        a1.push_back(a);
        a2.push_back(b);
      }
    offsets.resize(n_cb + 1);
    orc_pack_solve_data(n_t, n_bl, n_ch, n_dir, n_cb, n_corr, n_ant, n_sol.data(), a1.data(),
                        a2.data(), nullptr, nullptr, offsets.data(), nullptr, nullptr, nullptr,
                        nullptr, nullptr);
    const size_t nv = size_t(offsets[n_cb]);
    data.resize(nv * n_corr * 2);
    model.resize(nv * n_dir * n_corr * 2);
    ant1.resize(nv);
    ant2.resize(nv);
    sol_map.resize(nv * n_dir);
    orc_pack_solve_data(n_t, n_bl, n_ch, n_dir, n_cb, n_corr, n_ant, n_sol.data(), a1.data(),
                        a2.data(), raw.data(), raw_model.data(), offsets.data(), data.data(),
                        model.data(), ant1.data(), ant2.data(), sol_map.data());
  }

  int NSub() const {
    int n = 0;
    for (int s : n_sol) n += s;
    return n;
  }

  template <typename SolveDataType>
  SolveDataType Make(int n_corr) const {
    SolveDataType out(n_cb);
    for (int cb = 0; cb < n_cb; ++cb) {
      auto& block = out.ChannelBlock(cb);
      const size_t v0 = size_t(offsets[cb]), nv = size_t(offsets[cb + 1] - offsets[cb]);
      block.Resize(nv, n_dir);
      for (int d = 0; d < n_dir; ++d) block.SetNSolutions(d, uint32_t(n_sol[d]));
      for (size_t i = 0; i < nv; ++i) {
        for (int k = 0; k < n_corr; ++k)
          block.Visibility(i).values[k] = {data[((v0 + i) * n_corr + k) * 2],
                                           data[((v0 + i) * n_corr + k) * 2 + 1]};
        block.SetAntennas(i, ant1[v0 + i], ant2[v0 + i]);
        for (int d = 0; d < n_dir; ++d) {
          const size_t m = (v0 * n_dir + d * nv + i) * n_corr;
          for (int k = 0; k < n_corr; ++k)
            block.ModelVisibility(d, i).values[k] = {model[(m + k) * 2], model[(m + k) * 2 + 1]};
          block.SetSolutionIndex(d, i, sol_map[v0 * n_dir + d * nv + i]);
        }
      }
    }
    return out;
  }

  // Unit gains; identity matrices for the full-Jones solver (tSolvers.cc:374-381).
  std::vector<std::vector<DComplex>> UnitSolutions() const {
    std::vector<std::vector<DComplex>> out(
        n_cb, std::vector<DComplex>(size_t(n_ant) * NSub() * n_pol, 1.0));
    if (n_pol == 4)
      for (auto& v : out)
        for (size_t i = 0; i < v.size(); i += 4) v[i + 1] = v[i + 2] = 0.0;
    return out;
  }

  void InitializeSolver(SolverBase& solver) const {
    solver.SetAccuracy(kAccuracy);
    solver.SetStepSize(kStepSize);
    solver.SetPhaseOnly(false);
    solver.SetMaxIterations(kMaxIterations);
    solver.Initialize(n_ant, std::vector<size_t>(n_sol.begin(), n_sol.end()), n_cb);
  }

  // SolverTester::CheckDualResults (SolverTester.cc:409-439).
  void CheckDualResults(const std::vector<std::vector<DComplex>>& sol, double tolerance) const {
    const int n_sub = NSub();
    const DComplex* in = reinterpret_cast<const DComplex*>(input.data());
    for (int ch = 0; ch < n_cb; ++ch)
      for (int ant = 0; ant < n_ant; ++ant)
        for (int s = 0; s < n_sub; ++s) {
          const DComplex sx0 = sol[ch][s * 2], sy0 = sol[ch][s * 2 + 1];
          const DComplex ix0 = in[s * 2], iy0 = in[s * 2 + 1];
          const DComplex sx = sol[ch][(s + ant * n_sub) * 2], sy = sol[ch][(s + ant * n_sub) * 2 + 1];
          const DComplex ix = in[(s + ant * n_sub) * 2], iy = in[(s + ant * n_sub) * 2 + 1];
          CHECK(Close(std::norm(sx), std::norm(ix), tolerance));
          CHECK(Close(std::norm(sy), std::norm(iy), tolerance));
          CHECK(Close((sx * std::conj(sx0)).real(), (ix * std::conj(ix0)).real(), tolerance));
          CHECK(Close((sy * std::conj(sy0)).real(), (iy * std::conj(iy0)).real(), tolerance));
        }
  }

  // SolverTester::CheckScalarResults (SolverTester.cc:382-407).
  void CheckScalarResults(const std::vector<std::vector<DComplex>>& sol, double tolerance) const {
    const int n_sub = NSub();
    const DComplex* in = reinterpret_cast<const DComplex*>(input.data());
    for (int ch = 0; ch < n_cb; ++ch)
      for (int ant = 0; ant < n_ant; ++ant)
        for (int s = 0; s < n_sub; ++s) {
          const DComplex sol0 = sol[ch][s], inp0 = in[s * 2];
          const DComplex g = sol[ch][s + ant * n_sub], i = in[(s + ant * n_sub) * 2];
          CHECK(Close(std::norm(g), std::norm(i), tolerance));
          CHECK(std::abs((g * std::conj(sol0)).real() - (i * std::conj(inp0)).real()) < tolerance);
          CHECK(std::abs((g * std::conj(sol0)).imag() - (i * std::conj(inp0)).imag()) < tolerance);
        }
  }

  // The CPU restatement of the reference solver on the same data.
  std::vector<std::vector<DComplex>> OracleSolve(int n_corr, orc_result* result,
                                                 const orc_smoothness* smoothness = nullptr) const {
    orc_settings s{};
    s.n_antennas = n_ant;
    s.n_directions = n_dir;
    s.n_channel_blocks = n_cb;
    s.n_correlations = n_corr;
    s.n_solutions_per_direction = n_sol.data();
    s.max_iterations = int(kMaxIterations);
    s.accuracy = kAccuracy;
    s.constraint_accuracy = 1e-4;
    s.step_size = kStepSize;
    s.detect_stalling = 1;
    s.step_diff_sigma = 0.1;
    s.n_threads = 4;
    std::vector<double> flat(size_t(n_cb) * n_ant * NSub() * n_pol * 2);
    for (size_t i = 0; i < flat.size(); i += 2) flat[i] = 1.0;
    if (n_pol == 4) {
      for (size_t i = 0; i < flat.size(); i += 8) flat[i + 2] = flat[i + 4] = 0.0;
      orc_solve_full_jones(&s, smoothness, offsets.data(), data.data(), model.data(), ant1.data(),
                           ant2.data(), sol_map.data(), flat.data(), result, nullptr);
    } else if (n_pol == 1)
      orc_solve_scalar(&s, smoothness, offsets.data(), data.data(), model.data(), ant1.data(),
                       ant2.data(), sol_map.data(), flat.data(), result, nullptr);
    else if (smoothness)
      orc_solve_constrained(&s, smoothness, offsets.data(), data.data(), model.data(),
                            ant1.data(), ant2.data(), sol_map.data(), flat.data(), result,
                            nullptr);
    else
      orc_solve(&s, offsets.data(), data.data(), model.data(), ant1.data(), ant2.data(),
                sol_map.data(), flat.data(), result, nullptr);
    std::vector<std::vector<DComplex>> out = UnitSolutions();
    for (int cb = 0; cb < n_cb; ++cb)
      for (size_t i = 0; i < out[cb].size(); ++i)
        out[cb][i] = {flat[(cb * out[cb].size() + i) * 2], flat[(cb * out[cb].size() + i) * 2 + 1]};
    return out;
  }
};

// Largest |a - b| / |b| after referencing both to antenna 0 per (block, sub, pol).
double PhaseReferencedError(const std::vector<std::vector<DComplex>>& a,
                            const std::vector<std::vector<DComplex>>& b, int n_ant, int n_sub) {
  double worst = 0.0;
  for (size_t cb = 0; cb < a.size(); ++cb)
    for (int ant = 0; ant < n_ant; ++ant)
      for (int s = 0; s < n_sub; ++s)
        for (int p = 0; p < 2; ++p) {
          const size_t i = (size_t(ant) * n_sub + s) * 2 + p, r = size_t(s) * 2 + p;
          const DComplex ra = a[cb][i] * std::conj(a[cb][r]) / std::abs(a[cb][r]);
          const DComplex rb = b[cb][i] * std::conj(b[cb][r]) / std::abs(b[cb][r]);
          worst = std::max(worst, std::abs(ra - rb) / std::abs(rb));
        }
  return worst;
}

template <typename SolveDataType>
void DdIntervals(int n_corr) {
  Fixture f({1, 2, 5}, n_corr);
  IterativeDiagonalSolverHip solver;
  f.InitializeSolver(solver);
  CHECK(solver.NSolutionPolarizations() == 2u);
  CHECK(solver.SupportsDdSolutionIntervals());
  const SolveDataType data = f.Make<SolveDataType>(n_corr);
  auto solutions = f.UnitSolutions();
  const SolverBase::SolveResult result = solver.Solve(data, solutions, 0.0);
  f.CheckDualResults(solutions, 1.0e-2);
  CHECK(result.iterations <= Fixture::kMaxIterations + 1);
  orc_result oracle_result{};
  const auto reference = f.OracleSolve(n_corr, &oracle_result);
  const double err = PhaseReferencedError(solutions, reference, f.n_ant, f.NSub());
  std::printf("dd intervals n_corr=%d: iterations %zu (oracle %lld), gain error vs oracle %.2e\n",
              n_corr, result.iterations, static_cast<long long>(oracle_result.iterations), err);
  CHECK(err <= 1e-4);
  // The factored per-baseline sums with an FP64 antenna reduction reach the
  // accuracy no later than the reference's FP32 per-visibility sums.
  CHECK(int64_t(result.iterations) <= oracle_result.iterations);
}

// The smoothness constraint through the SolverBase-compatible interface
// (SetSmoothness, SetDdConstraintWeights), against the oracle's constrained
// solve on the same fixture. The fixture's gains do not vary over the
// channel blocks, so the constraint leaves them in place: the checks are
// parity with the oracle and CheckDualResults.
void SmoothnessConstraint() {
  Fixture f({1, 2, 5}, 2);
  const DuoSolveData data = f.Make<DuoSolveData>(2);
  IterativeDiagonalSolverHip solver;
  f.InitializeSolver(solver);
  IterativeDiagonalSolverHip::Smoothness m;
  for (int cb = 0; cb < f.n_cb; ++cb) m.channel_block_frequencies.push_back(120e6 + 7.5e6 * cb);
  m.bandwidth_hz = 10e6;
  m.spectral_exponent = -1.0;
  for (int a = 0; a < f.n_ant; ++a) m.antenna_factors.push_back(0.8 + 0.01 * a);
  for (int sub = 0; sub < f.NSub(); ++sub) m.dd_factors.push_back(1.0 + 0.1 * sub);
  solver.SetSmoothness(m);
  std::vector<std::vector<double>> weights(f.NSub());
  for (int sub = 0; sub < f.NSub(); ++sub)
    for (int i = 0; i < f.n_ant * f.n_cb; ++i) weights[sub].push_back(0.5 + 0.01 * ((i + sub) % 37));
  solver.SetDdConstraintWeights(weights);
  auto solutions = f.UnitSolutions();
  const SolverBase::SolveResult result = solver.Solve(data, solutions, 0.0);
  f.CheckDualResults(solutions, 1.0e-2);
  std::vector<double> flat_weights;
  for (const auto& w : weights) flat_weights.insert(flat_weights.end(), w.begin(), w.end());
  orc_smoothness c{};
  c.frequencies = m.channel_block_frequencies.data();
  c.bandwidth_hz = m.bandwidth_hz;
  c.bandwidth_ref_frequency_hz = m.bandwidth_ref_frequency_hz;
  c.spectral_exponent = m.spectral_exponent;
  c.kernel_truncation = 1;
  c.sub_solution_weights = flat_weights.data();
  c.antenna_factors = m.antenna_factors.data();
  c.dd_factors = m.dd_factors.data();
  orc_result oracle_result{};
  const auto reference = f.OracleSolve(2, &oracle_result, &c);
  const double err = PhaseReferencedError(solutions, reference, f.n_ant, f.NSub());
  std::printf("smoothness: iterations %zu (oracle %lld), gain error vs oracle %.2e\n",
              result.iterations, static_cast<long long>(oracle_result.iterations), err);
  CHECK(err <= 1e-4);
  CHECK(int64_t(result.iterations) <= oracle_result.iterations);
  // a wrong dd factor count is a settings error (SmoothnessConstraint.cc:36-45)
  m.dd_factors.pop_back();
  bool thrown = false;
  try {
    solver.SetSmoothness(m);
  } catch (const std::runtime_error&) {
    thrown = true;
  }
  CHECK(thrown);
}

// tSolvers.cc:155-200 (iterative_scalar_dd_intervals,
// iterative_duo_scalar_dd_intervals): SetScalarSolutions, CheckScalarResults
// at 1e-3; and parity with the oracle's scalar solve.
template <typename SolveDataType>
void ScalarDdIntervals(int n_corr) {
  Fixture f({1, 2, 5}, n_corr, 1);
  IterativeScalarSolverHip solver;
  f.InitializeSolver(solver);
  CHECK(solver.NSolutionPolarizations() == 1u);
  const SolveDataType data = f.Make<SolveDataType>(n_corr);
  auto solutions = f.UnitSolutions();
  const SolverBase::SolveResult result = solver.Solve(data, solutions, 0.0);
  f.CheckScalarResults(solutions, 1.0e-3);
  orc_result oracle_result{};
  const auto reference = f.OracleSolve(n_corr, &oracle_result);
  double worst = 0.0;
  for (size_t cb = 0; cb < solutions.size(); ++cb)
    for (size_t i = 0; i < solutions[cb].size(); ++i) {
      const DComplex r0 = reference[cb][i % f.NSub()], s0 = solutions[cb][i % f.NSub()];
      const DComplex ra = solutions[cb][i] * std::conj(s0) / std::abs(s0);
      const DComplex rb = reference[cb][i] * std::conj(r0) / std::abs(r0);
      worst = std::max(worst, std::abs(ra - rb) / std::abs(rb));
    }
  std::printf("scalar dd intervals n_corr=%d: iterations %zu (oracle %lld), gain error vs oracle "
              "%.2e\n",
              n_corr, result.iterations, static_cast<long long>(oracle_result.iterations), worst);
  CHECK(worst <= 1e-4);
  CHECK(int64_t(result.iterations) <= oracle_result.iterations);
}

// tSolvers.cc:409-450 (iterative_full_jones_dd_intervals): SetDiagonalSolutions,
// unit matrices, the diagonals through CheckDualResults at 1e-2 %; parity with
// the oracle's full-Jones solve (per matrix, after removing antenna 0's J_00
// phase); Duo data raises std::logic_error.
void FullJonesDdIntervals() {
  Fixture f({1, 2, 5}, 4, 4);
  IterativeFullJonesSolverHip solver;
  f.InitializeSolver(solver);
  CHECK(solver.NSolutionPolarizations() == 4u);
  const FullSolveData data = f.Make<FullSolveData>(4);
  auto solutions = f.UnitSolutions();
  const SolverBase::SolveResult result = solver.Solve(data, solutions, 0.0);
  std::vector<std::vector<DComplex>> diagonals(solutions.size());
  for (size_t cb = 0; cb < solutions.size(); ++cb)
    for (size_t i = 0; i < solutions[cb].size(); i += 4) {
      diagonals[cb].push_back(solutions[cb][i]);
      diagonals[cb].push_back(solutions[cb][i + 3]);
    }
  f.CheckDualResults(diagonals, 1.0e-2);
  orc_result oracle_result{};
  const auto reference = f.OracleSolve(4, &oracle_result);
  const size_t n_sub = size_t(f.NSub());
  double worst = 0.0;
  for (size_t cb = 0; cb < solutions.size(); ++cb)
    for (size_t m = 0; m < solutions[cb].size() / 4; ++m) {
      const size_t r = m % n_sub;  // antenna 0, same sub-solution
      const DComplex pa = std::conj(solutions[cb][r * 4]) / std::abs(solutions[cb][r * 4]);
      const DComplex pb = std::conj(reference[cb][r * 4]) / std::abs(reference[cb][r * 4]);
      double num = 0.0, den = 0.0;
      for (int k = 0; k < 4; ++k) {
        num += std::norm(solutions[cb][m * 4 + k] * pa - reference[cb][m * 4 + k] * pb);
        den += std::norm(reference[cb][m * 4 + k] * pb);
      }
      worst = std::max(worst, std::sqrt(num / den));
    }
  std::printf("full-Jones dd intervals: iterations %zu (oracle %lld), gain error vs oracle %.2e\n",
              result.iterations, static_cast<long long>(oracle_result.iterations), worst);
  CHECK(worst <= 1e-4);
  CHECK(int64_t(result.iterations) <= oracle_result.iterations);
  bool threw = false;
  try {
    const DuoSolveData duo = f.Make<DuoSolveData>(2);
    auto duo_solutions = f.UnitSolutions();
    solver.Solve(duo, duo_solutions, 0.0);
  } catch (const std::logic_error&) {
    threw = true;
  }
  CHECK(threw);
}

void IterationSemantics() {
  Fixture f({1, 1, 1}, 2);
  const DuoSolveData data = f.Make<DuoSolveData>(2);
  IterativeDiagonalSolverHip solver;
  f.InitializeSolver(solver);
  // min_iterations honoured (tSolvers.cc:512-529)
  solver.SetMinIterations(10);
  solver.SetAccuracy(1e8);
  auto solutions = f.UnitSolutions();
  CHECK(solver.Solve(data, solutions, 0.0).iterations == 10);
  // max + 1 when not converged (SolverBase.cc:197-200)
  solver.SetMinIterations(0);
  solver.SetMaxIterations(7);
  solver.SetAccuracy(1e-12);
  solutions = f.UnitSolutions();
  CHECK(solver.Solve(data, solutions, 0.0).iterations == 8);
  CHECK(!solver.LastSolveConverged());
}

void Errors() {
  Fixture f({1, 1, 1}, 2);
  const DuoSolveData data = f.Make<DuoSolveData>(2);
  IterativeDiagonalSolverHip uninitialized;
  auto solutions = f.UnitSolutions();
  bool thrown = false;
  try {
    uninitialized.Solve(data, solutions, 0.0);
  } catch (const std::runtime_error&) {
    thrown = true;
  }
  CHECK(thrown);
  IterativeDiagonalSolverHip solver;
  f.InitializeSolver(solver);
  solutions[0].pop_back();
  thrown = false;
  try {
    solver.Solve(data, solutions, 0.0);
  } catch (const std::runtime_error&) {
    thrown = true;
  }
  CHECK(thrown);
  thrown = false;
  try {
    IterativeDiagonalSolverHip bad(7);
  } catch (const std::runtime_error&) {
    thrown = true;
  }
  CHECK(thrown);
  // a solver kind without Full support keeps the base class behaviour
  struct DuoOnly : SolverBase {
    size_t NSolutionPolarizations() const override { return 2; }
  } duo_only;
  FullSolveData full(1);
  std::vector<std::vector<DComplex>> none;
  thrown = false;
  try {
    duo_only.Solve(full, none, 0.0);
  } catch (const std::logic_error&) {
    thrown = true;
  }
  CHECK(thrown);
}

}  // namespace

int main() {
  const std::vector<std::pair<const char*, std::function<void()>>> cases = {
      {"iterative_duo_diagonal_dd_intervals", [] { DdIntervals<DuoSolveData>(2); }},
      {"iterative_diagonal_dd_intervals", [] { DdIntervals<FullSolveData>(4); }},
      {"iterative_uni_scalar_dd_intervals", [] { ScalarDdIntervals<UniSolveData>(1); }},
      {"iterative_duo_scalar_dd_intervals", [] { ScalarDdIntervals<DuoSolveData>(2); }},
      {"iterative_scalar_dd_intervals", [] { ScalarDdIntervals<FullSolveData>(4); }},
      {"iterative_full_jones_dd_intervals", FullJonesDdIntervals},
      {"iteration_semantics", IterationSemantics},
      {"smoothness_constraint", SmoothnessConstraint},
      {"errors", Errors},
  };
  for (const auto& c : cases) {
    std::printf("case: %s\n", c.first);
    std::fflush(stdout);
    const int before = g_failures;
    try {
      c.second();
    } catch (const std::exception& e) {
      ++g_failures;
      std::fprintf(stderr, "%s: unexpected exception: %s\n", c.first, e.what());
    }
    std::printf("%s: %s\n", c.first, g_failures == before ? "passed" : "FAILED");
  }
  std::printf("%d checks, %d failures\n", g_checks, g_failures);
  return g_failures == 0 ? 0 : 1;
}
