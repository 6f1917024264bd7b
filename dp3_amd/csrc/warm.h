// Device-side pieces of the learned warm start, shared by the network
// (mlp.hip) and the batched warm solve (solver_host.hip). Internal to the
// library; the C-ABI is include/ddecal_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

struct ddh_mlp;

namespace ddh {

constexpr int kWarmFeatures = 16;
constexpr int kWarmOutputs = 4;

// Throws unless the network maps kWarmFeatures inputs to kWarmOutputs outputs.
void CheckWarmNetwork(const ddh_mlp& net);
// Waits for work queued on the network's own stream (training).
void WaitForNetwork(ddh_mlp& net);
// Forward pass over the device rows x [n][n_in] into y [n][n_out] on
// `stream`. Batches may share a network from different host threads (the
// weights are read only; their packing and the layered fallback are
// serialized), as long as no training runs meanwhile.
void MlpForwardRows(ddh_mlp& net, const float* x, int64_t n, float* y, hipStream_t stream);
// One interval's solutions in the warm start's concatenated buffers: its
// first solution (ne complex values each: 2 for a diagonal pair, 4 for a
// 2x2 matrix) and its shape (n_cb * n_ant * n_sub solutions). The features
// and the start use polarizations 0 and 1: the pair's values, or the
// matrix's diagonal entries. scalar: the pair holds one gain twice (the
// scalar solver), and the start keeps it so.
struct WarmSpan {
  int64_t pair0;
  int32_t n_cb, n_ant, n_sub, ne, scalar;
};
// Features [pairs][kWarmFeatures] of every span (device array of n_spans;
// max_pairs: the largest span) in one launch.
void LaunchWarmFeatures(const double2* g0, const double2* next1, const double2* ga,
                        const double2* gb, const double* den, const WarmSpan* spans, int n_spans,
                        int64_t max_pairs, float* features, hipStream_t stream);
// init = g0 * exp(pred) per polarization (for a 2x2 matrix: row p of g0
// times exp(pred_p); for a scalar span both values times the factor of the
// two outputs' mean), every span in one launch.
void LaunchWarmApply(const double2* g0, const float* pred, const WarmSpan* spans, int n_spans,
                     int64_t max_pairs, double2* init, hipStream_t stream);

}  // namespace ddh
