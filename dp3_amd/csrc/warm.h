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
// Device input rows [n][n_in] of the next MlpForwardDevice; valid until the
// network is used again.
float* MlpInputBuffer(ddh_mlp& net, int64_t n);
// Forward pass over the n rows of the input buffer on `stream`; returns the
// device outputs [n][n_out].
const float* MlpForwardDevice(ddh_mlp& net, int64_t n, hipStream_t stream);
// One interval's solutions in the warm start's concatenated buffers: its
// first solution pair (two complex values per pair) and its shape
// (n_cb * n_ant * n_sub pairs).
struct WarmSpan {
  int64_t pair0;
  int32_t n_cb, n_ant, n_sub, pad;
};
// Features [pairs][kWarmFeatures] of every span (device array of n_spans;
// max_pairs: the largest span) in one launch.
void LaunchWarmFeatures(const double2* g0, const double2* next1, const double2* ga,
                        const double2* gb, const double* den, const WarmSpan* spans, int n_spans,
                        int64_t max_pairs, float* features, hipStream_t stream);
// init = g0 * exp(pred) per polarization, every span in one launch.
void LaunchWarmApply(const double2* g0, const float* pred, const WarmSpan* spans, int n_spans,
                     int64_t max_pairs, double2* init, hipStream_t stream);

}  // namespace ddh
