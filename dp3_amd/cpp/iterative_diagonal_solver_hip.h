// IterativeDiagonalSolverHip, IterativeScalarSolverHip and
// IterativeFullJonesSolverHip: DP3's IterativeDiagonalSolver
// (ddecal/gain_solvers/IterativeDiagonalSolver.h:13-41),
// IterativeScalarSolver (IterativeScalarSolver.h) and IterativeFullJonesSolver
// (IterativeFullJonesSolver.h) behind the same SolverBase interface, solving
// on an MI355X through libddecal_hip's C-ABI (include/ddecal_hip.h). Duo
// (MC2x2FDiag) and Full (MC2x2F) data (Full only for the full-Jones solver),
// two, one or four solution polarizations, direction-dependent solution
// intervals.
//
// Behaviour as the reference: the iteration count is max+1 when the solve
// does not converge (SolverBase.cc:197-200), NaN solutions are data, settings
// errors raise std::runtime_error and unsupported data kinds
// std::logic_error. The device solver is created on the first Solve after a
// settings change and cached, as the reference caches its buffers
// (keep_host_buffers, Settings.cc:212). The one constraint supported is the
// smoothness constraint (SetSmoothness), which the device loop applies.
#ifndef DP3AMD_ITERATIVE_DIAGONAL_SOLVER_HIP_H
#define DP3AMD_ITERATIVE_DIAGONAL_SOLVER_HIP_H

#include <cstdint>
#include <memory>
#include <vector>

#include "solver_base.h"

struct ddh_solver;

namespace dp3amd {

// The shared implementation; n_solution_polarizations is 2 (diagonal), 1
// (scalar) or 4 (full-Jones).
class IterativeSolverHip : public SolverBase {
 public:
  ~IterativeSolverHip() override;

  size_t NSolutionPolarizations() const override { return n_solution_polarizations_; }
  bool SupportsDdSolutionIntervals() const override { return true; }

  void Initialize(size_t n_antennas, const std::vector<size_t>& n_solutions_per_direction,
                  size_t n_channel_blocks) override;

  SolveResult Solve(const DuoSolveData& data, std::vector<std::vector<DComplex>>& solutions,
                    double time) override;
  SolveResult Solve(const FullSolveData& data, std::vector<std::vector<DComplex>>& solutions,
                    double time) override;
  // Single-correlation data: the scalar solver only (as the reference, which
  // instantiates IterativeScalarSolver<std::complex<float>> alone).
  SolveResult Solve(const UniSolveData& data, std::vector<std::vector<DComplex>>& solutions,
                    double time) override;

  // Whether the last Solve met its accuracy (the reference keeps this inside
  // MakeResult).
  bool LastSolveConverged() const { return last_converged_; }

  // The smoothness constraint, applied on the device after every step: the
  // parameters of DP3's SmoothnessConstraint (SmoothnessConstraint.h:23-75,
  // Gaussian kernel), its Initialize frequencies (channel-block frequencies,
  // ascending) and its SetAntennaFactors / SetDdSmoothingFactors (empty:
  // ones). Empty frequencies remove the constraint.
  struct Smoothness {
    std::vector<double> channel_block_frequencies;
    double bandwidth_hz = 0.0;
    double bandwidth_ref_frequency_hz = 0.0;
    double spectral_exponent = 0.0;
    bool kernel_truncation = true;
    std::vector<double> antenna_factors;
    std::vector<double> dd_factors;
  };
  void SetSmoothness(const Smoothness& smoothness);
  // Constraint::SetWeights: [n_antennas][n_channel_blocks] (DDECal.cc:615-621).
  void SetConstraintWeights(const std::vector<double>& weights);
  // SolverBase::SetDdConstraintWeights (SolverBase.cc:436-442): per
  // sub-solution [n_antennas][n_channel_blocks] weights.
  void SetDdConstraintWeights(const std::vector<std::vector<double>>& weights);

 protected:
  // model_storage: 0 keeps complex64 model visibilities on the device (the
  // reference precision), 1 stores them as FP16 with a per (block,
  // direction) scale.
  IterativeSolverHip(int model_storage, int n_solution_polarizations);

 private:
  template <typename SolveDataType>
  SolveResult SolveImpl(const SolveDataType& data, std::vector<std::vector<DComplex>>& solutions,
                        int n_correlations);
  ddh_solver* DeviceSolver(int n_correlations);
  void ApplyConstraint(ddh_solver* solver) const;
  void CheckSmoothness(const Smoothness& smoothness) const;

  int model_storage_;
  int n_solution_polarizations_;
  ddh_solver* solver_ = nullptr;
  std::vector<int64_t> key_;  // settings the cached device solver was made with
  bool last_converged_ = false;
  Smoothness smoothness_;
  std::vector<double> constraint_weights_;       // [n_ant][n_cb], empty: ones
  std::vector<double> sub_solution_weights_;     // [n_sub][n_ant][n_cb], empty: none
};

// IterativeDiagonalSolver: two solution polarizations.
class IterativeDiagonalSolverHip final : public IterativeSolverHip {
 public:
  explicit IterativeDiagonalSolverHip(int model_storage = 0) : IterativeSolverHip(model_storage, 2) {}
};

// IterativeScalarSolver: one gain per antenna and solution for both
// correlations (NSolutionPolarizations() == 1).
class IterativeScalarSolverHip final : public IterativeSolverHip {
 public:
  explicit IterativeScalarSolverHip(int model_storage = 0) : IterativeSolverHip(model_storage, 1) {}
};

// IterativeFullJonesSolver: a 2x2 gain matrix per antenna and solution
// (NSolutionPolarizations() == 4, solutions [ant][sub][4] row-major per
// channel block); Full data only, as the reference (Duo data raises
// std::logic_error, SolverBase.h:213-219).
class IterativeFullJonesSolverHip final : public IterativeSolverHip {
 public:
  explicit IterativeFullJonesSolverHip(int model_storage = 0)
      : IterativeSolverHip(model_storage, 4) {}
};

}  // namespace dp3amd

#endif
