// IterativeDiagonalSolverHip: DP3's IterativeDiagonalSolver
// (ddecal/gain_solvers/IterativeDiagonalSolver.h:13-41) behind the same
// SolverBase interface, solving on an MI355X through libddecal_hip's C-ABI
// (include/ddecal_hip.h). Duo (MC2x2FDiag) and Full (MC2x2F) data, two
// solution polarizations, direction-dependent solution intervals.
//
// Behaviour as the reference: the iteration count is max+1 when the solve
// does not converge (SolverBase.cc:197-200), NaN solutions are data, settings
// errors raise std::runtime_error and unsupported data kinds
// std::logic_error. The device solver is created on the first Solve after a
// settings change and cached, as the reference caches its buffers
// (keep_host_buffers, Settings.cc:212).
#ifndef DP3AMD_ITERATIVE_DIAGONAL_SOLVER_HIP_H
#define DP3AMD_ITERATIVE_DIAGONAL_SOLVER_HIP_H

#include <cstdint>
#include <memory>
#include <vector>

#include "solver_base.h"

struct ddh_solver;

namespace dp3amd {

class IterativeDiagonalSolverHip final : public SolverBase {
 public:
  // model_storage: 0 keeps complex64 model visibilities on the device (the
  // reference precision), 1 stores them as FP16 with a per (block,
  // direction) scale.
  explicit IterativeDiagonalSolverHip(int model_storage = 0);
  ~IterativeDiagonalSolverHip() override;

  size_t NSolutionPolarizations() const override { return 2; }
  bool SupportsDdSolutionIntervals() const override { return true; }

  void Initialize(size_t n_antennas, const std::vector<size_t>& n_solutions_per_direction,
                  size_t n_channel_blocks) override;

  SolveResult Solve(const DuoSolveData& data, std::vector<std::vector<DComplex>>& solutions,
                    double time) override;
  SolveResult Solve(const FullSolveData& data, std::vector<std::vector<DComplex>>& solutions,
                    double time) override;

  // Whether the last Solve met its accuracy (the reference keeps this inside
  // MakeResult).
  bool LastSolveConverged() const { return last_converged_; }

 private:
  template <typename SolveDataType>
  SolveResult SolveImpl(const SolveDataType& data, std::vector<std::vector<DComplex>>& solutions,
                        int n_correlations);
  ddh_solver* DeviceSolver(int n_correlations);

  int model_storage_;
  ddh_solver* solver_ = nullptr;
  std::vector<int64_t> key_;  // settings the cached device solver was made with
  bool last_converged_ = false;
};

}  // namespace dp3amd

#endif
