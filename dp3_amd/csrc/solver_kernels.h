// Device kernels of the iterative diagonal solver.
//
// One solver iteration (IterativeDiagonalSolver.cc:67-94) of the streaming
// iteration is three launches:
//
//   correlate      streams the data plane and every model plane of a block
//                  once. Per visibility it forms the full residual
//                  R = D - sum_d G_a M_d G_b^H (SolveData.h:204-244) and, per
//                  direction d and correlation k, accumulates the per-baseline
//                  sums U_dk = sum R_k conj(M_dk) and S_dk = sum |M_dk|^2.
//                  All directions are solved against residuals built from the
//                  same old solutions (IterativeDiagonalSolver.cc:115-126), so
//                  one pass serves every direction. The gains are constant
//                  over a baseline's visibilities within a segment, so the
//                  reference's per-visibility numerator/denominator terms
//                  (:172-197) factor into these sums (see antenna_solve).
//   antenna_solve  per (block, direction): gathers each antenna's baselines,
//                  forms numerator and denominator (IterativeDiagonalSolver.cc
//                  :134-138) from U, S and the gains in FP64, divides
//                  (:206-221), applies the step (SolverBase.cc:88-158) and the
//                  per-solution convergence term (SolverBase.cc:254-268), with
//                  per-workgroup partial sums.
//   decide         one workgroup: FP64 convergence test (SolverBase.cc:296-306),
//                  stall detection (:63-86), stopping criterion
//                  (SolverBase.h:274-285) and the NaN-repair values of the next
//                  iteration (SolverBase.cc:342-371), which consumers apply
//                  when they read a solution.
//
// The Gram iteration replaces the correlate by an apply of per-baseline cross
// products made once per solve; its fused form (at most 64 antennas) runs the
// apply, the antenna solve, the finishing step and the decision of an
// iteration as one launch (GramSolveBlock).
//
// Solutions ping-pong between two buffers by iteration parity, so no copy
// pass is needed.
//
// Every launch first reads the device-side `done` flag and returns when the
// solve has already stopped, so the host may enqueue iterations ahead of the
// convergence decision.
#ifndef DP3_AMD_SOLVER_KERNELS_H
#define DP3_AMD_SOLVER_KERNELS_H

#include <cstdint>
#include <vector>

#include "common.h"

struct ddh_problem;

namespace ddh {

struct CbInfo {
  int32_t n_ch;       // channels in this block
  int32_t pad;
  int64_t data_off;   // visibility offset of the data plane
  int64_t model_off;  // visibility offset of direction 0's model plane
};

struct DeviceState {
  int32_t iteration;  // iterations decided (the current solutions are sol[iteration & 1])
  int32_t done;
  int32_t has_converged;
  int32_t pending;    // merged fused iteration: 1 when iteration `iteration` has been
                      // finished (its solutions stored) but not yet decided
  double last_step;
  double n_var_count;
  double step_mean;
  double step_var;
};

// Everything one iteration needs. Passed by value to each kernel.
struct IterArgs {
  // problem (device layout)
  const void* data;          // Duo: float4 per vis; Full: 2 float4
  const void* model;         // complex64 or FP16 planes
  const float* model_scale;  // [n_cb][n_dir] (FP16 only)
  const int32_t* ant1;       // [bl_stride]
  const int32_t* ant2;
  const int32_t* csr_off;    // [n_ant+1]
  const int32_t* csr;        // per entry: (baseline << 1) | side, other antenna
  const CbInfo* cbinfo;      // [n_cb]
  const int32_t* seg_t0;     // [n_seg+1]
  const int32_t* seg_sol;    // [n_seg][n_dir]
  // solver
  const int32_t* n_sol_dir;  // [n_dir]
  const int32_t* sol_start;  // [n_dir]
  double2* sol0;             // [n_cb][n_ant][n_sub][2], current at even iterations
  double2* sol1;             // current at odd iterations
  double* mean;              // [n_cb] repair value of the current solutions
  double* wg_conv;           // [n_cb][n_dir] per-workgroup convergence sums (the merged
                             // fused iteration uses two such halves by iteration parity,
                             // likewise for the three arrays below)
  int32_t* wg_conv_n;
  double* wg_mean;           // [n_cb][n_dir] per-workgroup sums of |pol 0|
  int32_t* wg_mean_n;
  float* partials;           // [n_cb][n_seg][n_dir][n_corr][3][bl_stride]
  DeviceState* state;
  double* step_mags;         // [max_iterations + 1]
  int32_t n_t, n_bl, bl_stride, n_cb, n_dir, n_ant, n_sub, n_seg;
  int32_t max_iterations, min_iterations;
  double accuracy, step_size, step_diff_sigma;
  int32_t detect_stalling, phase_only;
  double* den_out;           // optional [n_cb][n_ant][n_sub][2] denominators
  int32_t split;             // threads per (antenna, solution) in antenna_solve
  int32_t solve_chunk;       // baselines of partials per LDS stage in antenna_solve
                             // (0: all of a block's; AntennaSolveChunk)
  int32_t max_nsd;           // largest n_sol_dir entry
  int32_t scalar;            // 1: one gain for both polarizations (scalar solver)
  int32_t full_jones;        // 1: 2x2 solutions [n_cb][n_ant][n_sub][4] (full-Jones
                             // solver; Full data, streaming iteration, partials
                             // [n_cb][n_seg][n_dir][kFullJonesTerms][bl_stride])
  float* fj_products;        // full-Jones model products, once per solve
                             // [n_cb][n_seg][n_dir][kFullJonesModelTerms][bl_stride]
  // Smoothness constraint on the stepped solutions (null sm_freq: none).
  const double* sm_freq;        // [n_cb] channel-block frequencies (Hz)
  const double* sm_corr;        // [n_cb] frequency correction of the kernel size per block
  const int32_t* sm_window;     // [n_cb][2] kernel window [first, last) per block
  const double* sm_weights;     // [n_ant][n_cb], or [n_sub][n_ant][n_cb] (sm_sub_weights)
  const double* sm_ant_factor;  // [n_ant]
  const double* sm_dd_factor;   // [n_sub]
  double2* sm_out;              // [n_cb][n_ant][n_sub][2] smoothed solutions
  double sm_bandwidth;
  int32_t sm_truncate, sm_sub_weights;
  // Gram iteration (Duo, n_dir <= kMaxGramDirections): data/model cross
  // products C_d and G_e,d per (block, segment, correlation, baseline), made
  // once per solve; [n_cb][n_seg][2][n_dir + n_dir(n_dir+1)/2][bl_stride].
  float2* gram;
  // Fused Gram iteration (GramSolve: apply and antenna solve in one kernel;
  // Duo data, at most 64 antennas, one solution per direction in one
  // segment): the cross products in walk order
  // [cb][k][walk_rows][GramWalkTerms(n_dir)][n_ant] (a row's terms of every
  // lane contiguous), and per walk slot (step row * 64 + lane) the lane's role
  // for its baseline (0: none, 1: the lane's antenna is the first, 2: the
  // second) and the baseline (-1: none). walk_rows = 0: no fused iteration.
  float2* gram_walk;
  const int32_t* walk_code;
  const int32_t* walk_bl;
  int32_t walk_rows;
  int32_t gs_groups;  // row groups per polarization of the fused iteration (1..8)
  int32_t* arrive;    // merged fused iteration: workgroups of the current launch that
                      // have read the state (the last one stores the next state)
  double* gs_part;    // per-workgroup sums by iteration parity: two halves of
                      // n_cb x 2 x kGramSolveMaxGroups x n_dir x 3 x 64 doubles, each
                      // [n_cb][2 x gs_groups][n_dir][3][64] from its start
};

// Largest row-group count of the fused Gram iteration.
constexpr int kGramSolveMaxGroups = 8;

// Per-baseline partials of the full-Jones correlate: X = sum R_all J_b M^H
// (8 floats) and Y = sum R_all^H J_a M (8); and the model products made once
// per solve, Q[p][q] = sum M_p conj(M_q) over the four entries of M
// (Hermitian: 4 real diagonal terms and 6 complex upper terms).
constexpr int kFullJonesTerms = 16;
constexpr int kFullJonesModelTerms = 16;
// Directions whose X and Y one full-Jones correlate workgroup accumulates;
// more directions take several workgroups (direction groups) per tile.
constexpr int kFullJonesGroupSize = 8;

// Largest direction count of the Gram iteration (its per-lane accumulators
// grow with n_dir^2).
constexpr int kMaxGramDirections = 8;
// Full-Jones Gram iteration: cross products per (block, segment, baseline),
// C_d and G_e,d (e < d), 16 complex terms each; the cross-products kernel
// deals out three blocks of 16 terms per wave.
constexpr int kFjGramBlocksPerWave = 3;
__host__ __device__ inline int FullJonesGramTerms(int n_dir) {
  return 16 * (n_dir + n_dir * (n_dir - 1) / 2);
}
__host__ __device__ inline int FjGramWaves(int n_dir) {
  return (n_dir * (n_dir + 1) / 2 + kFjGramBlocksPerWave - 1) / kFjGramBlocksPerWave;
}
__host__ __device__ inline int GramTerms(int n_dir) { return n_dir + n_dir * (n_dir + 1) / 2; }
// Terms per walk slot of the fused iteration: C_d (n_dir), the off-diagonal
// G_e,d (e < d), and the real diagonals G_d,d two per slot.
__host__ __device__ inline int GramWalkTerms(int n_dir) {
  return n_dir + n_dir * (n_dir - 1) / 2 + (n_dir + 1) / 2;
}

// One unit of batched work: correlate tiles use all four fields; antenna
// solve items use (interval, cb) with the direction in `seg`; prologue items
// use (interval, cb).
struct BatchTile {
  int32_t interval;
  int32_t cb;
  int32_t seg;
  int32_t tile;
};

// Host-side launchers (defined in solver_kernels.hip).
void LaunchCorrelate(const IterArgs& a, int n_corr, bool f16, int waves, int rows,
                     hipStream_t stream);
// threads: the antenna solve's block size (0: the default for a.split).
void LaunchAntennaSolve(const IterArgs& a, int n_corr, bool f16, int threads, hipStream_t stream);
void LaunchDecide(const IterArgs& a, hipStream_t stream);
void LaunchPrologue(const IterArgs& a, hipStream_t stream);
// Full-Jones solver: the model products of the problem, once per solve.
void LaunchFullJonesModelProducts(const IterArgs& a, bool f16, int waves, hipStream_t stream);
// Full-Jones Gram iteration: the cross products once per solve (after or
// before the model products, which the apply also reads), and per iteration
// the apply, which writes the correlate's X and Y partials.
void LaunchFullJonesGram(const IterArgs& a, bool f16, hipStream_t stream);
void LaunchFullJonesApplyGram(const IterArgs& a, hipStream_t stream);
void LaunchBatchFullJonesGram(const IterArgs* d_args, const BatchTile* d_tiles, int n_tiles,
                              int max_dir, bool f16, hipStream_t stream);
// FP64 chi^2 of the residual with solutions `sol` (current buffer of `a`):
// per-workgroup partial sums into `partial`, then their fixed-order sum,
// times `factor`, into `out`.
void LaunchChi2(const IterArgs& a, int n_corr, bool f16, const double2* sol, double* partial,
                int n_partial, double* out, hipStream_t stream, double factor = 1.0);
int Chi2Partials(const IterArgs& a);
size_t CorrelateSharedBytes(int n_dir, int n_corr, int waves);
size_t MaxSharedBytes();
size_t AntennaSolveSharedBytes(const IterArgs& a, int n_corr);
// Baselines per LDS stage of the antenna solve's partials: 0 (all) where a
// block's fit the workgroup's LDS with the gains, else the largest multiple
// of 64 that does (no limit on the station count beyond the gains' LDS).
int AntennaSolveChunk(const IterArgs& a);
// Block size of antenna_solve: about four threads per (antenna, solution).
int AntennaSolveThreads(int n_ant, int max_nsd);
// Antenna solve schedule: `split` threads per (antenna, solution) (a power of
// two, combined by a shuffle tree) in blocks of SolveThreads(a, requested)
// threads (requested 0: the default, at least four threads per item). The
// default split is the largest power of two up to 8 that fits twice into the
// default block. The full-Jones solve keeps its default block.
constexpr int kSolveMaxThreads = 1024;
int DefaultSolveSplit(int n_ant, int max_nsd);
int SolveThreads(const IterArgs& a, int requested);
void PrepareKernelAttributes();
// Tiles of one direction count: entries [first, first + count) of a tile list.
struct TileGroup {
  int32_t n_dir;
  int32_t first;
  int32_t count;
};
// Smoothness constraint launches of a batch (n_intervals == 0: none).
struct SmoothLaunch {
  int32_t n_intervals;
  int64_t max_elements;   // largest n_cb * n_ant * n_sub * 2 of the batch
  int32_t finish_threads; // block size of the per-(direction, block) pass
};
// Smoothness constraint of one solve: the smoothed series, then per
// (direction, block) the smoothed solutions written back with their
// convergence partials (replacing antenna_solve's).
void LaunchSmoothing(const IterArgs& a, hipStream_t stream);
int SmoothFinishThreads(int n_ant, int max_nsd);
// One batched iteration: the correlate once per tile group, then the antenna
// solve over n_items items and the decision for n_intervals intervals (zero
// counts skip a kernel).
void LaunchBatch(const IterArgs* d_args, const BatchTile* d_tiles, const TileGroup* groups,
                 int n_groups, const BatchTile* d_items, int n_items, int n_intervals,
                 size_t solve_lds, int solve_threads, int n_corr, bool f16, int waves, int rows,
                 const SmoothLaunch& smooth, hipStream_t stream, bool solve_chunked = false);
// Gram iteration: the cross products once per solve, then per iteration the
// correlate partials from them and the current gains.
void LaunchGram(const IterArgs& a, bool f16, int waves, hipStream_t stream);
void LaunchApplyGram(const IterArgs& a, bool f16, hipStream_t stream);
// Fused Gram iteration: the cross products into walk order once per solve,
// then per iteration workgroups per block for apply + antenna solve and a
// finishing kernel (the decision follows as before). Merged form: the sums
// of the first iteration alone, then per iteration one kernel that finishes
// the previous iteration and forms the sums of the next, and the decision.
enum GramSolveMode {
  kGramSolveAndFinish = 0,   // the state's iteration: sums, then the finishing kernel
  kGramSolveOnly = 1,        // the state's iteration: its sums only
  kGramFinishThenSolve = 2,  // the end of the state's iteration, then the next one's sums
};
void LaunchGramPermute(const IterArgs& a, hipStream_t stream);
void LaunchGramSolve(const IterArgs& a, bool f16, hipStream_t stream,
                     GramSolveMode mode = kGramSolveAndFinish);
// Batched: items (interval, cb); groups split them by direction count.
void LaunchBatchGramPermute(const IterArgs* d_args, const BatchTile* d_items, int n_items,
                            int max_dir, hipStream_t stream);
// d_items: (interval, cb, part) workgroups; d_blocks: (interval, cb) for the
// finishing kernel (n_blocks 0: the sums only). finish_first: each workgroup
// finishes the state's iteration and forms the next one's sums (merged form;
// n_blocks 0).
void LaunchBatchGramSolve(const IterArgs* d_args, const BatchTile* d_items,
                          const TileGroup* groups, int n_groups, bool f16,
                          const BatchTile* d_blocks, int n_blocks, hipStream_t stream,
                          bool finish_first = false);
// LDS of the Gram apply: the segment's gains of every antenna and direction.
size_t ApplyGramSharedBytes(int n_ant, int n_dir);
constexpr size_t kApplyGramMaxShared = 64 * 1024;
size_t GramSharedBytes(int n_dir, int waves);
void LaunchBatchGram(const IterArgs* d_args, const BatchTile* d_tiles, const TileGroup* groups,
                     int n_groups, bool f16, int waves, hipStream_t stream);
// One batched Gram iteration: apply per tile group, then antenna solve and decision.
// apply_ant: the largest antenna count of the batch (sizes the apply's LDS).
void LaunchBatchApply(const IterArgs* d_args, const BatchTile* d_tiles, const TileGroup* groups,
                      int n_groups, int apply_ant, const BatchTile* d_items, int n_items,
                      int n_intervals, size_t solve_lds, int solve_threads, bool f16,
                      const SmoothLaunch& smooth, hipStream_t stream, bool solve_chunked = false);
// Solution values of every batched interval between its solver buffers and
// one concatenated buffer `buf` (interval i at d_off[i] .. d_off[i + 1],
// complex values; max_count the largest), one launch:
//   kMoveToSol0        buf -> sol0
//   kMoveFromSol1      sol1 -> buf
//   kMoveFromCurrent   the current solutions (by iteration parity) -> buf
enum BatchMove { kMoveToSol0 = 0, kMoveFromSol1 = 1, kMoveFromCurrent = 2 };
void LaunchBatchMove(const IterArgs* d_args, const int64_t* d_off, int n_intervals,
                     int64_t max_count, BatchMove mode, double2* buf, hipStream_t stream);
void LaunchBatchPrologue(const IterArgs* d_args, const BatchTile* d_items, int n_items,
                         hipStream_t stream);
void LaunchBatchTail(const IterArgs* d_args, const BatchTile* d_items, int n_items,
                     int n_intervals, size_t solve_lds, int solve_threads, int n_corr, bool f16,
                     const SmoothLaunch& smooth, hipStream_t stream, bool solve_chunked = false);
// The smoothing passes (when any interval has the constraint) and the decision.
void LaunchBatchSmoothDecide(const IterArgs* d_args, const BatchTile* d_items, int n_items,
                             int n_intervals, const SmoothLaunch& smooth, hipStream_t stream);
// One batched full-Jones iteration: the correlate per direction-count group
// (tile entries hold tile + group * bl_stride / 64), the antenna solve over
// n_items (interval, block, direction) items, smoothing and the decision.
// Gram iteration: d_apply_tiles (interval, block, segment, tile) for the
// apply over at most apply_dir directions, in place of the correlate groups.
void LaunchBatchFullJones(const IterArgs* d_args, const BatchTile* d_tiles, const TileGroup* groups,
                          int n_groups, const BatchTile* d_items, int n_items, int n_intervals,
                          int solve_threads, bool f16, int waves, const SmoothLaunch& smooth,
                          hipStream_t stream, const BatchTile* d_apply_tiles = nullptr,
                          int n_apply_tiles = 0, int apply_dir = 0);
// The full-Jones model products of every interval, once per solve: one
// workgroup per (interval, block, segment, tile) entry.
void LaunchBatchFullJonesModelProducts(const IterArgs* d_args, const BatchTile* d_tiles,
                                       int n_tiles, bool f16, int waves, hipStream_t stream);

// Layout conversion: reference order [plane][t][bl][ch][n_corr] complex64 ->
// device order [plane][t][ch][bl_stride] (complex64 or FP16/scale). With
// `uni` the source holds one complex value per visibility and the Duo
// (n_corr 2) destination gets it in both correlations.
void LaunchRepack(const float* src, void* dst, int n_planes, int n_t, int n_bl,
                  int n_ch, int bl_stride, int n_corr, bool f16,
                  const float* plane_scale, hipStream_t stream, bool uni = false);
// Per-plane max |re|,|im| of reference-order planes (FP16 scale), written as
// float bits into max_bits[n_planes] (which must be zeroed first).
void LaunchPlaneMaxAbs(const float* src, int64_t plane_floats, int n_planes,
                       uint32_t* max_bits, hipStream_t stream);
void LaunchScaleFromMax(const uint32_t* max_bits, float* scale, int n,
                        hipStream_t stream);

// Weighting and packing of raw interval buffers into a device problem
// (AssignAndWeight + SolveData, defined in ingest.hip).
struct IngestArgs {
  const float* data;       // [n_t][n_bl_all][n_ch][4] complex64
  const uint8_t* flags;    // [n_t][n_bl_all][n_ch][4] or null
  const float* weights;    // [n_t][n_bl_all][n_ch][4] or null (unit weights)
  const float* model;      // [n_dir][n_t][n_bl_all][n_ch][4] complex64
  const int32_t* cross;    // [n_bl_all] solve-data baseline index, -1 for autos
  const int32_t* ch_start; // [n_cb + 1] first channel of each block
  const int32_t* ch_cb;    // [n_ch] block of each channel
  const CbInfo* cbinfo;    // [n_cb] plane offsets of the problem
  void* out_data;          // problem data planes
  void* out_model;         // problem model planes
  const float* scale;      // [n_cb][n_dir] FP16 scale (written by the first pass)
  uint32_t* max_bits;      // [n_cb][n_dir] workspace of the FP16 maxima
  int32_t n_t, n_bl_all, n_ch, n_cb, n_dir, bl_stride, linear;
  int32_t uni;             // single-correlation problem: both Duo slots 0.5 (XX + YY)
};
void LaunchAssignAndWeigh(const IngestArgs& g, int n_corr, bool f16, hipStream_t stream);

// Internals of a device problem, shared by the solver host code and the
// predict (solver_host.hip owns the struct behind ddh_problem).
struct ProblemView {
  void* model;
  void* data;
  float* scale;
  const int32_t* ant1;
  const int32_t* ant2;
  const CbInfo* cbinfo;
  int n_t, n_bl, bl_stride, n_cb, n_dir, n_corr, n_ant;
  bool f16;
  bool uni;  // single-correlation problem kept as Duo pairs of equal values
  std::vector<int32_t> ch_per_cb;
  std::vector<CbInfo> cb_host;
  hipStream_t stream;
};
ProblemView GetProblemView(ddh_problem* p);

}  // namespace ddh

#endif
