// The solver interface of DP3's SolverBase (ddecal/gain_solvers/SolverBase.h:
// 24-361), reduced to what the iterative diagonal solver uses: the settings
// with their reference defaults (SolverBase.cc:27-39), Initialize
// (SolverBase.cc:41-61), the Solve overloads per data kind (SolverBase.h:
// 205-227, std::logic_error for unsupported kinds) and SolveResult.
// Of DP3's constraints, the smoothness constraint is applied on the device
// after every step (IterativeDiagonalSolverHip::SetSmoothness); the others
// would need a per-iteration host round trip and are not offered here.
#ifndef DP3AMD_SOLVER_BASE_H
#define DP3AMD_SOLVER_BASE_H

#include <complex>
#include <cstddef>
#include <stdexcept>
#include <vector>

#include "solve_data.h"

namespace dp3amd {

class SolverBase {
 public:
  typedef std::complex<double> DComplex;

  struct SolveResult {
    size_t iterations = 0;
    size_t constraint_iterations = 0;
  };

  virtual ~SolverBase() = default;

  virtual void Initialize(size_t n_antennas, const std::vector<size_t>& n_solutions_per_direction,
                          size_t n_channel_blocks) {
    if (n_antennas == 0 || n_channel_blocks == 0 || n_solutions_per_direction.empty())
      throw std::runtime_error("Solver initialized with an empty dimension");
    for (size_t n : n_solutions_per_direction)
      if (n == 0) throw std::runtime_error("A direction needs at least one solution");
    n_antennas_ = n_antennas;
    n_solutions_per_direction_ = n_solutions_per_direction;
    n_channel_blocks_ = n_channel_blocks;
  }

  virtual size_t NSolutionPolarizations() const = 0;
  virtual bool SupportsDdSolutionIntervals() const { return false; }

  bool GetPhaseOnly() const { return phase_only_; }
  void SetPhaseOnly(bool phase_only) { phase_only_ = phase_only; }
  size_t GetMaxIterations() const { return max_iterations_; }
  void SetMaxIterations(size_t max_iterations) { max_iterations_ = max_iterations; }
  size_t GetMinIterations() const { return min_iterations_; }
  void SetMinIterations(size_t min_iterations) { min_iterations_ = min_iterations; }
  void SetAccuracy(double accuracy) { accuracy_ = accuracy; }
  double GetAccuracy() const { return accuracy_; }
  void SetConstraintAccuracy(double a) { constraint_accuracy_ = a; }
  double GetConstraintAccuracy() const { return constraint_accuracy_; }
  void SetStepSize(double step_size) { step_size_ = step_size; }
  double GetStepSize() const { return step_size_; }
  void SetDetectStalling(bool detect_stalling, double step_diff_sigma) {
    detect_stalling_ = detect_stalling;
    step_diff_sigma_ = step_diff_sigma;
  }
  bool GetDetectStalling() const { return detect_stalling_; }
  double GetStepDiffSigma() const { return step_diff_sigma_; }

  size_t NAntennas() const { return n_antennas_; }
  size_t NChannelBlocks() const { return n_channel_blocks_; }
  const std::vector<size_t>& NSolutionsPerDirection() const { return n_solutions_per_direction_; }
  size_t NSubSolutions() const {
    size_t n = 0;
    for (size_t s : n_solutions_per_direction_) n += s;
    return n;
  }

  virtual SolveResult Solve(const FullSolveData&, std::vector<std::vector<DComplex>>&, double) {
    throw std::logic_error(
        "Full-visibility solver called for a solver that does not support full-visibility "
        "solving");
  }
  virtual SolveResult Solve(const DuoSolveData&, std::vector<std::vector<DComplex>>&, double) {
    throw std::logic_error(
        "Duo-visibility (xx/yy) solver called for a solver that does not support "
        "duo-visibility solving");
  }
  virtual SolveResult Solve(const UniSolveData&, std::vector<std::vector<DComplex>>&, double) {
    throw std::logic_error(
        "Single-visibility (xx/yy) solver called for a solver that does not support "
        "single-visibility solving");
  }

 private:
  size_t n_antennas_ = 0;
  size_t n_channel_blocks_ = 0;
  std::vector<size_t> n_solutions_per_direction_;
  // SolverBase.cc:27-39
  size_t max_iterations_ = 100;
  size_t min_iterations_ = 0;
  double accuracy_ = 1e-5;
  double constraint_accuracy_ = 1e-4;
  double step_size_ = 0.2;
  bool detect_stalling_ = true;
  double step_diff_sigma_ = 0.1;
  bool phase_only_ = false;
};

}  // namespace dp3amd

#endif
