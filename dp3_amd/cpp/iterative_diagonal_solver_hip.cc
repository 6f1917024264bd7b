#include "iterative_diagonal_solver_hip.h"

#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/ddecal_hip.h"

namespace dp3amd {

namespace {

[[noreturn]] void ThrowLastError(const char* what) {
  throw std::runtime_error(std::string(what) + ": " + ddh_last_error());
}

}  // namespace

IterativeSolverHip::IterativeSolverHip(int model_storage, int n_solution_polarizations)
    : model_storage_(model_storage), n_solution_polarizations_(n_solution_polarizations) {
  if (model_storage != DDH_MODEL_C64 && model_storage != DDH_MODEL_F16)
    throw std::runtime_error("Unknown model storage for the HIP iterative solver");
}

IterativeSolverHip::~IterativeSolverHip() {
  if (solver_) ddh_solver_destroy(solver_);
}

void IterativeSolverHip::Initialize(size_t n_antennas,
                                            const std::vector<size_t>& n_solutions_per_direction,
                                            size_t n_channel_blocks) {
  SolverBase::Initialize(n_antennas, n_solutions_per_direction, n_channel_blocks);
  if (solver_) ddh_solver_destroy(solver_);
  solver_ = nullptr;
  key_.clear();
}

ddh_solver* IterativeSolverHip::DeviceSolver(int n_correlations) {
  if (NAntennas() == 0) throw std::runtime_error("Solve called before Initialize");
  double acc = GetAccuracy(), cacc = GetConstraintAccuracy(), step = GetStepSize(),
         sigma = GetStepDiffSigma();
  std::vector<int64_t> key = {n_correlations, int64_t(GetMaxIterations()),
                              int64_t(GetMinIterations()), GetDetectStalling(), GetPhaseOnly(),
                              model_storage_};
  for (double v : {acc, cacc, step, sigma}) {
    int64_t bits;
    std::memcpy(&bits, &v, sizeof bits);
    key.push_back(bits);
  }
  if (solver_ && key == key_) return solver_;
  if (solver_) ddh_solver_destroy(solver_);
  solver_ = nullptr;
  std::vector<int32_t> n_sol(NSolutionsPerDirection().begin(), NSolutionsPerDirection().end());
  ddh_config c;
  ddh_config_defaults(&c);
  c.n_antennas = int32_t(NAntennas());
  c.n_directions = int32_t(n_sol.size());
  c.n_channel_blocks = int32_t(NChannelBlocks());
  c.n_correlations = n_correlations;
  c.n_solutions_per_direction = n_sol.data();
  c.max_iterations = int32_t(GetMaxIterations());
  c.min_iterations = int32_t(GetMinIterations());
  c.accuracy = acc;
  c.constraint_accuracy = cacc;
  c.step_size = step;
  c.detect_stalling = GetDetectStalling() ? 1 : 0;
  c.step_diff_sigma = sigma;
  c.phase_only = GetPhaseOnly() ? 1 : 0;
  c.model_storage = model_storage_;
  c.n_solution_polarizations = n_solution_polarizations_;
  solver_ = ddh_solver_create(&c);
  if (!solver_) ThrowLastError("Could not create the HIP solver");
  key_ = key;
  ApplyConstraint(solver_);
  return solver_;
}

void IterativeSolverHip::CheckSmoothness(const Smoothness& m) const {
  if (m.channel_block_frequencies.empty() || NAntennas() == 0) return;  // checked at Solve
  if (m.channel_block_frequencies.size() != NChannelBlocks())
    throw std::runtime_error("Smoothness constraint: one frequency per channel block is needed");
  if (!m.antenna_factors.empty() && m.antenna_factors.size() != NAntennas())
    throw std::runtime_error("Smoothness constraint: one factor per antenna is needed");
  if (!m.dd_factors.empty() && m.dd_factors.size() != NSubSolutions())
    throw std::runtime_error("Invalid setting for the dd factors option of smoothness constraint");
}

void IterativeSolverHip::ApplyConstraint(ddh_solver* solver) const {
  const Smoothness& m = smoothness_;
  if (m.channel_block_frequencies.empty()) {
    if (ddh_solver_set_smoothness(solver, nullptr, 0.0, 0.0, 0.0, 0, nullptr, nullptr) != DDH_OK)
      ThrowLastError("Could not remove the smoothness constraint");
    return;
  }
  CheckSmoothness(m);
  if (ddh_solver_set_smoothness(solver, m.channel_block_frequencies.data(), m.bandwidth_hz,
                                m.bandwidth_ref_frequency_hz, m.spectral_exponent,
                                m.kernel_truncation ? 1 : 0,
                                m.antenna_factors.empty() ? nullptr : m.antenna_factors.data(),
                                m.dd_factors.empty() ? nullptr : m.dd_factors.data()) != DDH_OK)
    ThrowLastError("Could not set the smoothness constraint");
  if (!sub_solution_weights_.empty()) {
    if (ddh_solver_set_sub_solution_weights(solver, sub_solution_weights_.data()) != DDH_OK)
      ThrowLastError("Could not set the constraint weights");
  } else if (!constraint_weights_.empty()) {
    if (ddh_solver_set_constraint_weights(solver, constraint_weights_.data()) != DDH_OK)
      ThrowLastError("Could not set the constraint weights");
  }
}

void IterativeSolverHip::SetSmoothness(const Smoothness& smoothness) {
  CheckSmoothness(smoothness);
  smoothness_ = smoothness;
  if (solver_) ApplyConstraint(solver_);
}

void IterativeSolverHip::SetConstraintWeights(const std::vector<double>& weights) {
  if (weights.size() != NAntennas() * NChannelBlocks())
    throw std::runtime_error("Constraint weights: n_antennas x n_channel_blocks values needed");
  constraint_weights_ = weights;
  sub_solution_weights_.clear();
  if (solver_ && !smoothness_.channel_block_frequencies.empty())
    if (ddh_solver_set_constraint_weights(solver_, constraint_weights_.data()) != DDH_OK)
      ThrowLastError("Could not set the constraint weights");
}

void IterativeSolverHip::SetDdConstraintWeights(
    const std::vector<std::vector<double>>& weights) {
  if (weights.size() != NSubSolutions())
    throw std::runtime_error("Dd constraint weights: one vector per sub-solution needed");
  sub_solution_weights_.clear();
  for (const std::vector<double>& w : weights) {
    if (w.size() != NAntennas() * NChannelBlocks())
      throw std::runtime_error("Dd constraint weights: n_antennas x n_channel_blocks values needed");
    sub_solution_weights_.insert(sub_solution_weights_.end(), w.begin(), w.end());
  }
  if (solver_ && !smoothness_.channel_block_frequencies.empty())
    if (ddh_solver_set_sub_solution_weights(solver_, sub_solution_weights_.data()) != DDH_OK)
      ThrowLastError("Could not set the constraint weights");
}

template <typename SolveDataType>
SolverBase::SolveResult IterativeSolverHip::SolveImpl(
    const SolveDataType& data, std::vector<std::vector<DComplex>>& solutions,
    int n_correlations) {
  if (data.NChannelBlocks() != NChannelBlocks() || solutions.size() != NChannelBlocks())
    throw std::runtime_error("Solve data or solutions do not match the initialized channel blocks");
  const size_t per_cb = NAntennas() * NSubSolutions() * NSolutionPolarizations();
  std::vector<ddh_cb_view> views(NChannelBlocks());
  std::vector<double*> sol_ptrs(NChannelBlocks());
  for (size_t cb = 0; cb < NChannelBlocks(); ++cb) {
    const auto& block = data.ChannelBlock(cb);
    if (solutions[cb].size() != per_cb)
      throw std::runtime_error("Solution vector of a channel block has the wrong size");
    if (block.NDirections() != NSolutionsPerDirection().size())
      throw std::runtime_error("Solve data has a different number of directions");
    ddh_cb_view& v = views[cb];
    v.n_visibilities = int64_t(block.NVisibilities());
    const bool empty = block.NVisibilities() == 0;
    v.data = empty ? nullptr : reinterpret_cast<const float*>(&block.Visibility(0));
    v.model = empty ? nullptr : reinterpret_cast<const float*>(&block.ModelVisibility(0, 0));
    v.antenna_pairs = empty ? nullptr : reinterpret_cast<const uint32_t*>(block.AntennaIndices());
    v.solution_map = block.SolutionMapData();
    sol_ptrs[cb] = reinterpret_cast<double*>(solutions[cb].data());
  }
  ddh_result r{};
  if (ddh_solve(DeviceSolver(n_correlations), views.data(), int32_t(views.size()),
                sol_ptrs.data(), 0.0, &r) != DDH_OK)
    ThrowLastError("HIP solve failed");
  last_converged_ = r.has_converged != 0;
  SolveResult out;
  out.iterations = size_t(r.iterations);
  out.constraint_iterations = size_t(r.constraint_iterations);
  return out;
}

SolverBase::SolveResult IterativeSolverHip::Solve(
    const DuoSolveData& data, std::vector<std::vector<DComplex>>& solutions, double time) {
  if (n_solution_polarizations_ == 4) return SolverBase::Solve(data, solutions, time);
  return SolveImpl(data, solutions, 2);
}

SolverBase::SolveResult IterativeSolverHip::Solve(
    const FullSolveData& data, std::vector<std::vector<DComplex>>& solutions, double) {
  return SolveImpl(data, solutions, 4);
}

SolverBase::SolveResult IterativeSolverHip::Solve(
    const UniSolveData& data, std::vector<std::vector<DComplex>>& solutions, double time) {
  if (n_solution_polarizations_ != 1) return SolverBase::Solve(data, solutions, time);
  return SolveImpl(data, solutions, 1);
}

}  // namespace dp3amd
