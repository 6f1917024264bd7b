// Host-side solve data with the accessors of DP3's SolveData<MatrixType>
// (ddecal/gain_solvers/SolveData.h:31-133): per channel block the weighted
// visibilities, the model visibilities of every direction, the antenna pair
// and the per-direction solution index of every visibility. Visibilities are
// ordered time -> cross baseline -> channel in block, as SolveData's regular
// constructor packs them (SolveData.cc:129-181).
//
// The reference builds this from DPBuffers (SolveData.cc:77-188); here the
// caller fills it through the writer methods. The memory layout of a
// visibility matches the reference matrix types: MC2x2FDiag holds the two
// diagonal correlations, MC2x2F all four, each a complex<float>.
#ifndef DP3AMD_SOLVE_DATA_H
#define DP3AMD_SOLVE_DATA_H

#include <complex>
#include <cstddef>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <utility>
#include <vector>

namespace dp3amd {

struct UniMatrix {  // std::complex<float>: 0.5 (XX + YY) (SolveData.cc:29-31)
  std::complex<float> values[1];
};
struct DuoMatrix {  // aocommon::MC2x2FDiag: (XX, YY)
  std::complex<float> values[2];
};
struct FullMatrix {  // aocommon::MC2x2F: (XX, XY, YX, YY)
  std::complex<float> values[4];
};

template <typename MatrixType>
class SolveData {
 public:
  class ChannelBlockData {
   public:
    void Resize(size_t n_visibilities, size_t n_directions) {
      data_.assign(n_visibilities, MatrixType{});
      model_data_.assign(n_visibilities * n_directions, MatrixType{});
      antenna_indices_.assign(n_visibilities, {0u, 0u});
      n_solutions_.assign(n_directions, 1u);
      solution_map_.assign(n_visibilities * n_directions, 0u);
      n_visibilities_ = n_visibilities;
    }
    size_t NDirections() const { return n_solutions_.size(); }
    size_t NVisibilities() const { return n_visibilities_; }
    uint32_t Antenna1Index(size_t i) const { return antenna_indices_[i].first; }
    uint32_t Antenna2Index(size_t i) const { return antenna_indices_[i].second; }
    uint32_t SolutionIndex(size_t d, size_t i) const {
      return solution_map_[d * n_visibilities_ + i];
    }
    const uint32_t* SolutionMapData() const { return solution_map_.data(); }
    const MatrixType& Visibility(size_t i) const { return data_[i]; }
    const MatrixType& ModelVisibility(size_t d, size_t i) const {
      return model_data_[d * n_visibilities_ + i];
    }
    uint32_t NSolutionsForDirection(size_t d) const { return n_solutions_[d]; }
    uint32_t NSubSolutions() const {
      return std::accumulate(n_solutions_.begin(), n_solutions_.end(), uint32_t(0));
    }
    // [n_vis] (antenna1, antenna2) pairs, as the reference's antenna_indices_.
    const std::pair<uint32_t, uint32_t>* AntennaIndices() const {
      return antenna_indices_.data();
    }

    // Writers (the reference's SolveData constructor fills these members).
    MatrixType& Visibility(size_t i) { return data_[i]; }
    MatrixType& ModelVisibility(size_t d, size_t i) { return model_data_[d * n_visibilities_ + i]; }
    void SetAntennas(size_t i, uint32_t a1, uint32_t a2) { antenna_indices_[i] = {a1, a2}; }
    void SetSolutionIndex(size_t d, size_t i, uint32_t s) {
      solution_map_[d * n_visibilities_ + i] = s;
    }
    void SetNSolutions(size_t d, uint32_t n) { n_solutions_[d] = n; }

   private:
    size_t n_visibilities_ = 0;
    std::vector<MatrixType> data_;
    std::vector<MatrixType> model_data_;  // [n_dir][n_vis]
    std::vector<std::pair<uint32_t, uint32_t>> antenna_indices_;
    std::vector<uint32_t> n_solutions_;
    std::vector<uint32_t> solution_map_;  // [n_dir][n_vis]
  };

  explicit SolveData(size_t n_channel_blocks) : channel_blocks_(n_channel_blocks) {}
  size_t NChannelBlocks() const { return channel_blocks_.size(); }
  const ChannelBlockData& ChannelBlock(size_t i) const { return channel_blocks_[i]; }
  ChannelBlockData& ChannelBlock(size_t i) { return channel_blocks_[i]; }

 private:
  std::vector<ChannelBlockData> channel_blocks_;
};

using UniSolveData = SolveData<UniMatrix>;
using DuoSolveData = SolveData<DuoMatrix>;
using FullSolveData = SolveData<FullMatrix>;

}  // namespace dp3amd

#endif
