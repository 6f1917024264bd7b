// C-ABI of the MI355X-native DDECal iterative-diagonal solver and
// point/Gaussian predict (library: dp3_amd/libddecal_hip.so).
//
// Every entry point uses plain C types. Functions returning int return 0 on
// success and a negative status on error; ddh_last_error() then holds the
// message (thread-local). NaN solutions are data, not errors.
//
// Reference interfaces replaced (lofar-astron/DP3 6.6.0):
//   ddh_solver_create   <- SolverFactory CreateDiagonalSolver
//                          (ddecal/SolverFactory.cc:86-146) + SolverBase
//                          setters (SolverBase.h:109-177) + Initialize
//                          (SolverBase.cc:41-61)
//   ddh_solve           <- IterativeDiagonalSolver<VisMatrix>::Solve
//                          (ddecal/gain_solvers/IterativeDiagonalSolver.cc:38-97;
//                          SolverBase.h:205-219), taking SolveData's channel-
//                          block arrays (SolveData.h:86-92) and the
//                          solutions vector<vector<complex<double>>>.
//   ddh_problem_*       <- SolveData regular layout (SolveData.cc:77-188), kept
//                          device resident in a baseline-fastest layout.
//   ddh_solve_problem   <- same Solve, on a device-resident problem.
//   ddh_batch_*         <- DDECal::doSolve's loop over solution intervals
//                          (steps/DDECal.cc:595-708), all intervals at once.
//   ddh_predict         <- Simulator::visit(PointSource|GaussianSource)
//                          (base/Simulator.cc:110-468) + OnePredict's Stokes-I
//                          copy (steps/OnePredict.cc:437-457).
#ifndef DDECAL_HIP_H
#define DDECAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDH_OK 0
#define DDH_ERR_ARGUMENT -1
#define DDH_ERR_LAYOUT -2
#define DDH_ERR_DEVICE -3
#define DDH_ERR_UNSUPPORTED -4

// Model visibility storage on the device.
#define DDH_MODEL_C64 0   // complex64 (reference precision)
#define DDH_MODEL_F16 1   // complex FP16 with a per (block, direction) scale

// How an iteration correlates data and model (the results agree to rounding).
#define DDH_ITERATION_STREAM 0  // stream data and model planes every iteration
#define DDH_ITERATION_GRAM 1    // per-baseline data/model cross products made once
                                // per solve, applied to the gains every
                                // iteration (up to 8 directions; Duo data for
                                // the diagonal and scalar solvers, Full data
                                // for the full-Jones solver)

typedef struct ddh_solver ddh_solver;
typedef struct ddh_problem ddh_problem;

typedef struct {
  int32_t n_antennas;
  int32_t n_directions;
  int32_t n_channel_blocks;
  int32_t n_correlations;                    // 2 (Duo, XX/YY), 4 (Full) or 1 (single:
                                             // 0.5 (XX + YY), SolveData.cc:29-31; the
                                             // scalar solver only)
  const int32_t* n_solutions_per_direction;  // [n_directions]
  int32_t max_iterations;                    // SolverBase default 100
  int32_t min_iterations;                    // default 0
  double accuracy;                           // default 1e-5
  double constraint_accuracy;                // default 1e-4
  double step_size;                          // default 0.2
  int32_t detect_stalling;                   // default 1
  double step_diff_sigma;                    // default 0.1
  int32_t phase_only;                        // default 0
  int32_t model_storage;                     // DDH_MODEL_*
  int32_t iteration;                         // DDH_ITERATION_*
  // 2: IterativeDiagonalSolver (default); 1: IterativeScalarSolver
  // (IterativeScalarSolver.cc:30-226), one gain per antenna and solution for
  // both correlations, solutions [n_ant][n_sub] per channel block; 4:
  // IterativeFullJonesSolver (IterativeFullJonesSolver.cc:18-224), a 2x2
  // matrix per antenna and solution, solutions [n_ant][n_sub][4] (row-major
  // 00, 01, 10, 11) per channel block; Full data and the streaming iteration
  // only; single and batched solves (a batch holds full-Jones intervals only),
  // no learned warm start, no one-step
  int32_t n_solution_polarizations;
} ddh_config;

// Fills the SolverBase defaults (SolverBase.cc:27-39).
void ddh_config_defaults(ddh_config* config);

typedef struct {
  int64_t iterations;             // SolveResult::iterations (max+1 if not converged)
  int64_t constraint_iterations;  // SolveResult::constraint_iterations
  int32_t has_converged;
  int32_t n_step_magnitudes;
  double last_step_magnitude;
} ddh_result;

// One channel block of a DuoSolveData / FullSolveData (host memory,
// SolveData::ChannelBlockData): visibilities ordered time -> cross
// baseline -> channel in block (SolveData.cc:129-181).
typedef struct {
  int64_t n_visibilities;
  const float* data;             // [n_vis][n_corr] complex64 (&Visibility(0))
  const float* model;            // [n_dir][n_vis][n_corr] complex64 (&ModelVisibility(0,0))
  const uint32_t* antenna_pairs; // [n_vis][2] (antenna_indices_)
  const uint32_t* solution_map;  // [n_dir][n_vis] (SolutionMapData())
} ddh_cb_view;

const char* ddh_last_error(void);
int ddh_device_count(void);

ddh_solver* ddh_solver_create(const ddh_config* config);
void ddh_solver_destroy(ddh_solver* solver);

// Drop-in solve from host SolveData arrays. solutions[cb] points at the
// channel block's [n_ant][n_sub][2] complex128 vector: initial guess in,
// result out. `time` is passed to constraints in the reference; the only
// constraint applied here is the smoothness constraint
// (ddh_solver_set_smoothness), which does not depend on it, so it is unused.
int ddh_solve(ddh_solver* solver, const ddh_cb_view* channel_blocks,
              int32_t n_channel_blocks, double* const* solutions, double time,
              ddh_result* result);

// Device-resident problem in the library's own layout: per channel block,
// planes of [time][channel][baseline_stride] visibilities, baseline fastest.
// sol_index[d * n_times + t] is the absolute solution index of direction d at
// time step t (SolveData.cc:167-168).
ddh_problem* ddh_problem_create(const ddh_solver* solver, int32_t n_times,
                                int32_t n_baselines, const int32_t* antenna1,
                                const int32_t* antenna2,
                                const int32_t* channels_per_block,
                                const uint32_t* sol_index);
void ddh_problem_destroy(ddh_problem* problem);
// Copies host SolveData channel blocks (layout as in ddh_cb_view, with the
// problem's structure) into the device layout.
int ddh_problem_upload(ddh_problem* problem, const ddh_cb_view* channel_blocks);
// Baselines per (time, channel) row of a device plane: n_baselines rounded
// up to a multiple of 64, the padding reading as zeros.
int ddh_problem_baseline_stride(const ddh_problem* problem);
// Device pointers of the data plane / a model plane of block cb (for device
// producers such as the predict). Model planes of FP16 problems hold
// half2 pairs; their scale is set by ddh_problem_set_model_scale.
void* ddh_problem_data_ptr(ddh_problem* problem, int32_t cb);
void* ddh_problem_model_ptr(ddh_problem* problem, int32_t cb, int32_t direction);
int ddh_problem_set_model_scale(ddh_problem* problem, const float* scale /*[n_cb][n_dir]*/);
// Reads back a device plane into host memory in the device layout.
int ddh_problem_download_plane(ddh_problem* problem, int32_t cb, int32_t plane,
                               void* host);

// Fills every data and model plane of `problem` from the raw buffers of its
// interval: AssignAndWeight (ddecal/gain_solvers/SolverTools.cc:81-167,
// linear_weighting_mode as linear_weighting; Weigh :48-79) followed by the
// regular SolveData constructor (ddecal/gain_solvers/SolveData.cc:77-188), in
// one pass on the device. Buffers in DPBuffer order over the interval's time
// steps, every baseline of the buffer (auto-correlations included):
//   data    [n_times][n_baselines][n_channels][4] complex64
//   flags   [n_times][n_baselines][n_channels][4] bytes, or NULL (none set)
//   weights [n_times][n_baselines][n_channels][4] float, or NULL (unit)
//   model   [n_directions][n_times][n_baselines][n_channels][4] complex64
// with n_channels the sum of the problem's channels per block, and
// antenna1/antenna2 [n_baselines] host arrays whose cross-correlations are
// the problem's baselines in order. A visibility with a set flag or a
// non-finite value in any correlation of the data or of a model is zeroed in
// the data and every model. FP16 problems get their per-(block, direction)
// model scale from the weighted models (an extra pass over the buffers).
// The _device form takes device pointers (data, model, weights 16-byte
// aligned, flags 4-byte aligned) and enqueues on `stream` (a hipStream_t;
// NULL: the problem's stream, which is its solver's); the host form copies
// the buffers in and reports the kernels' device time.
int ddh_problem_assign_and_weight_device(ddh_problem* problem, int32_t n_baselines,
                                         const int32_t* antenna1, const int32_t* antenna2,
                                         const float* data, const uint8_t* flags,
                                         const float* weights, const float* model,
                                         int32_t linear_weighting, void* stream);
int ddh_problem_assign_and_weight(ddh_problem* problem, int32_t n_baselines,
                                  const int32_t* antenna1, const int32_t* antenna2,
                                  const float* data, const uint8_t* flags,
                                  const float* weights, const float* model,
                                  int32_t linear_weighting, double* kernel_ms);

// Solve a device-resident problem. solutions as in ddh_solve.
int ddh_solve_problem(ddh_solver* solver, ddh_problem* problem,
                      double* const* solutions, ddh_result* result);

// One sky-model component (PointSource / GaussianSource, base/PointSource.h,
// base/GaussianSource.h) of direction `direction`.
typedef struct {
  double ra, dec;                  // radians
  double I, Q, U, V;               // Stokes at the reference frequency (Jy)
  double ref_freq;                 // Hz, for spectral terms
  int32_t n_spectral_terms;        // 0..8 (0 = flat spectrum)
  int32_t log_si;                  // 1: logarithmic spectral index
  double spectral_terms[8];
  int32_t has_rm;                  // rotation measure (pol_fraction, pol_angle, rm)
  int32_t is_gaussian;
  double pol_fraction, pol_angle, rm;
  double major_axis, minor_axis;   // FWHM, radians
  double position_angle;           // radians, north through east
  int32_t pa_is_absolute;
  int32_t direction;               // solver direction (patch) index
} ddh_component;

// Predicts every model plane of `problem` (all channel blocks and
// directions): station_uvw [n_times][n_stations][3] metres, freq and
// chan_width [total channels] Hz. Antenna indices of the problem's baselines
// index the stations. stokes_i_only writes XX = YY = I (XY = YX = 0).
// correct_freq_smearing multiplies by sinc(0.5 (phi_q - phi_p) width)
// (Simulator.cc:190-199). scaled_ncp_uvw (3 doubles, the north celestial
// pole in UVW scaled by the earth's angular speed over one integration, as
// FastPredict.cc:447-462 derives it; element 0 taken as zero, as
// Simulator.cc:493-503 does) turns on the time-smearing factor
// sinc(0.5 (rot_q - rot_p) f) of Simulator.cc:177-188 (the constructor of
// Simulator.h:142-150); NULL leaves it off (the constructor of :117-122).
// FP64 throughout, each direction's sum rounded once.
// FP16 problems get a per-direction scale bound from the spectra.
int ddh_predict(ddh_problem* problem, double ref_ra, double ref_dec,
                const double* station_uvw, int32_t n_stations, const double* freq,
                const double* chan_width, const ddh_component* components,
                int32_t n_components, int32_t stokes_i_only,
                int32_t correct_freq_smearing, const double* scaled_ncp_uvw);

// Writes the data planes of `problem` as sum_d Diag(g_a) M_d Diag(g_b)^H plus
// complex Gaussian noise (noise_sigma on each of re and im), with
// gains[cb] -> [n_ant][n_sub][2] complex128 and sol_index[d][t] as in
// ddh_problem_create. Builds synthetic solution intervals on the device.
int ddh_problem_synthesize_data(ddh_problem* problem, double* const* gains,
                                int32_t n_sub, const uint32_t* sol_index,
                                float noise_sigma, uint32_t seed);

// Smoothness constraint (ddecal/constraints/SmoothnessConstraint.h:23-75 with
// the Gaussian kernel of KernelSmoother.h), applied on the device to the
// next solutions after every step, before the convergence test
// (SolverBase::ApplyConstraints, SolverBase.cc:174-190; the constraint is
// always satisfied). Replaces the constructor, Initialize (the channel-block
// frequencies, ascending, Hz), SetAntennaFactors ([n_antennas], NULL: ones)
// and SetDdSmoothingFactors ([n_sub], NULL: ones) of SmoothnessConstraint.cc:
// 11-46 and SolverFactory.cc:427-435. NULL frequencies switch the constraint
// off. Weights default to ones until set.
int ddh_solver_set_smoothness(ddh_solver* solver, const double* channel_block_frequencies,
                              double bandwidth_hz, double bandwidth_ref_frequency_hz,
                              double spectral_exponent, int32_t kernel_truncation,
                              const double* antenna_factors, const double* dd_factors);
// Constraint::SetWeights (DDECal.cc:615-621): [n_antennas][n_channel_blocks].
int ddh_solver_set_constraint_weights(ddh_solver* solver, const double* weights);
// Constraint::SetSubSolutionWeights (SolverBase.cc:436-442):
// [n_sub][n_antennas][n_channel_blocks], used instead of the plain weights.
int ddh_solver_set_sub_solution_weights(ddh_solver* solver, const double* weights);

// Kernel-level timing of the last solve: total device time of the
// correlate kernel and number of its launches that did work.
int ddh_solver_last_timing(const ddh_solver* solver, double* correlate_ms,
                           int64_t* correlate_launches, double* solve_ms);

// Schedule of the correlate kernel: waves per 64-baseline tile (1..4) and
// visibility rows loaded per loop trip (1, or 2 for Duo data); hipGraph
// replay of the iteration body; iterations enqueued between convergence
// polls. Results do not depend on the schedule beyond float rounding.
int ddh_set_schedule(ddh_solver* solver, int32_t waves_per_tile,
                     int32_t rows_per_trip, int32_t use_graph,
                     int32_t poll_interval);
// Fused Gram iteration (on by default): for Gram solves of Duo data with at
// most 64 antennas, each antenna pair at most once and one solution per
// direction in one time segment, one kernel per channel block forms the
// correlate sums and each antenna's numerator and denominator on chip,
// without per-baseline partials in device memory. Off: the apply and antenna
// solve kernels. Results agree to FP32 summation order.
// enable: 0 off; 1 on, each iteration's finishing step (divide, step,
// convergence terms) done inside the next iteration's kernel where no
// smoothness constraint is set; 2 on, with a finishing kernel per iteration.
// 1 and 2 give the same values.
int ddh_set_gram_fusion(ddh_solver* solver, int32_t enable);
// Schedule of the fused Gram iteration: each block's walk over the antenna
// pairs is split, per polarization, into `groups` (1..8) workgroups whose
// per-antenna sums a finishing kernel adds up (default 2; set by the
// autotuners).
int ddh_set_gram_solve_groups(ddh_solver* solver, int32_t groups);
// Schedule of the antenna solve (streaming and unfused Gram iterations of
// the scalar and diagonal solvers; IterativeDiagonalSolver.cc:130-222 per
// (block, direction)): `split` threads (a power of two up to 64) share each
// (antenna, solution)'s baseline list, in workgroups of `threads` (a multiple
// of 64 up to 1024, at least split x antennas x solutions of the largest
// direction). 0 selects the default for either (about four threads per
// (antenna, solution), split 2 for LOFAR-sized arrays); the autotuners pick
// them. The split changes the FP64 summation order of each antenna's
// numerator and denominator only.
int ddh_set_antenna_solve_schedule(ddh_solver* solver, int32_t split, int32_t threads);
// The split and block size the solver's antenna solve uses.
int ddh_antenna_solve_schedule(const ddh_solver* solver, int32_t* split, int32_t* threads);
// Records HIP events around every correlate launch of the next solves
// (ddh_solver_last_timing then reports their summed duration).
int ddh_set_profiling(ddh_solver* solver, int32_t enable);

// One undamped solver update from `solutions_in` (the solution of the
// Jacobi step, IterativeDiagonalSolver.cc:206-221, without the damping of
// SolverBase::Step), with its denominators: next_out[cb] like the solutions,
// den_out[cb] -> [n_ant][n_sub][2] doubles (may be NULL; for the scalar
// solver both entries hold the sum over the polarizations that its update
// divides by; for the full-Jones solver the diagonal of each 2x2
// denominator, the solution of
// IterativeFullJonesSolver.cc:178-189 without the damping). Used as the
// data/model correlation summary of the warm start.
int ddh_one_step(ddh_solver* solver, ddh_problem* problem, double* const* solutions_in,
                 double* const* next_out, double* const* den_out);

// FP64 chi^2 of the residual D - sum_d G_a M_d G_b^H of a device problem for
// the given solutions (no counterpart in the reference, which computes none).
int ddh_chi2(ddh_solver* solver, ddh_problem* problem, double* const* solutions,
             double* chi2);

// ddh_chi2 without allocation, for callers that cannot allocate device
// memory (code under graph capture, tight serving loops); no counterpart in
// the reference. Enqueues on `stream` (a hipStream_t; NULL: the null stream)
// and returns without waiting. solutions: device memory, [n_cb][n_ant][n_sub]
// pairs (diagonal solver) or row-major 2x2 matrices (full-Jones solver) of
// complex128, the layout ddh_solve takes per block, blocks consecutive;
// chi2: one double of device memory. The per-workgroup partial sums go to a
// fixed device array of 65536 doubles (512 KiB, held by each GPU until the
// process ends), so the problem may have at most 65536 64-baseline tiles x
// channel blocks; larger problems and the scalar solver are rejected. Calls
// on one GPU share that array: calls that overlap in time on one GPU race.
int ddh_chi2_device(ddh_solver* solver, ddh_problem* problem, const double* solutions,
                    double* chi2, void* stream);

// Batched solves: many solution intervals (each its own solver
// configuration and device problem; all with the same correlation count and
// model storage) advance together, one launch per kernel per iteration over
// all intervals' tiles, until every interval met its stopping criterion.
// Replaces DDECal's serial loop over buffered solution intervals
// (steps/DDECal.cc:595-708). Solvers must outlive the batch; a problem must
// live while the batch is bound to it (a rebind to other problems releases
// it: the batch keeps copies of the shapes it checks, not the problems).
typedef struct ddh_batch ddh_batch;
ddh_batch* ddh_batch_create(ddh_solver* const* solvers, ddh_problem* const* problems,
                            int32_t n_intervals);
void ddh_batch_destroy(ddh_batch* batch);
int ddh_batch_set_schedule(ddh_batch* batch, int32_t waves_per_tile, int32_t rows_per_trip,
                           int32_t use_graph, int32_t poll_interval);
// Row groups of the batch's fused Gram iteration (as ddh_set_gram_solve_groups)
// and the current value.
int ddh_batch_set_gram_solve_groups(ddh_batch* batch, int32_t groups);
int ddh_batch_gram_solve_groups(const ddh_batch* batch);
// Antenna solve schedule of every interval of the batch (as
// ddh_set_antenna_solve_schedule; the block covers the largest interval) and
// the values in use.
int ddh_batch_set_antenna_solve_schedule(ddh_batch* batch, int32_t split, int32_t threads);
int ddh_batch_antenna_solve_schedule(const ddh_batch* batch, int32_t* split, int32_t* threads);
// Points the batch at other problems of the same shapes (the next intervals
// of the same sequences), keeping buffers, schedule and captured graph.
int ddh_batch_rebind(ddh_batch* batch, ddh_problem* const* problems);
// solutions[i][cb] -> interval i, channel block cb ([n_ant][n_sub][2] complex128).
int ddh_batch_solve(ddh_batch* batch, double* const* const* solutions, ddh_result* results);
int ddh_batch_last_time(const ddh_batch* batch, double* ms);
// Learned warm start for every interval of the batch, all on the device:
// solutions[i] hold the previous interval's solutions g0 on entry and the
// final solutions on return. From g0 the batch runs one undamped update
// (with its denominators) and two stretches of `stretch` damped iterations,
// turns them into features, runs the network (ddh_mlp, 16 -> 4) and starts
// the solve from gb * exp(output), gb the iterate after the second stretch
// (the network predicts the remaining correction, so the stretches' progress
// is kept). Same arithmetic as the per-interval
// sequence ddh_one_step, 2 x ddh_solve_problem(max `stretch`, accuracy 0,
// no stall detection), ddh_warmstart_features, ddh_mlp_forward and
// ddh_warmstart_apply (the _2x2 variants for a full-Jones batch). The
// network must not be in use on another stream.
// No counterpart in the reference (DDECal.cc:487-499 offers unit or
// previous solutions only).
typedef struct ddh_mlp ddh_mlp;
int ddh_batch_solve_warm(ddh_batch* batch, ddh_mlp* net, int32_t stretch,
                         double* const* const* solutions, ddh_result* results);
// Device time of the warm-start phase of the last ddh_batch_solve_warm
// (ddh_batch_last_time covers the solve from the learned start).
int ddh_batch_last_warm_time(const ddh_batch* batch, double* ms);
// HIP events around every batched correlate launch (Gram iteration: apply)
// and every antenna solve + decision of the next solves (graph replay off
// while enabled); ddh_batch_last_timing reports their summed durations and
// the number of iterations that did work.
int ddh_batch_set_profiling(ddh_batch* batch, int32_t enable);
int ddh_batch_last_timing(const ddh_batch* batch, double* correlate_ms, double* solve_ms,
                          int64_t* launches);

// Schedule autotuner: times the correlate kernel under every candidate
// schedule (waves per tile x rows per trip; Gram iteration: the cross
// products) on the actual shapes and keeps the fastest; for a fused Gram
// iteration it then times the fused kernel under 1, 2, 4 and 8 row groups and
// keeps the fastest (ddh_batch_gram_solve_groups reads it back); otherwise
// (scalar and diagonal solvers) it times the antenna solve under every split
// of 1..16 threads per (antenna, solution) with one, two and four times the
// threads that split needs, and the default, and keeps the fastest
// (ddh_antenna_solve_schedule / ddh_batch_antenna_solve_schedule read it
// back). Leaves the solve state reset (unit gains).
int ddh_autotune(ddh_solver* solver, ddh_problem* problem, int32_t* waves_per_tile,
                 int32_t* rows_per_trip, double* best_ms);
int ddh_batch_autotune(ddh_batch* batch, int32_t* waves_per_tile, int32_t* rows_per_trip,
                       double* best_ms);

// Applies the solutions of an interval to the model data of every direction
// and subtracts the sum from the data: DDECal::CorrectAndSubtractModels
// (steps/DDECal.cc:1074-1124, ApplySolutionAndSubtract :1031-1072, the 2x2
// products of :992-1020 in FP64). Buffers in DPBuffer order over the time
// steps of the interval, unweighted:
//   data  [n_times][n_baselines][n_channels][4] complex64 (data minus the
//         corrected models when subtract != 0)
//   model [n_directions][n_times][n_baselines][n_channels][4] complex64
//         (replaced by the corrected models when keep_model != 0)
//   solutions [n_channel_blocks][n_antennas][n_directions][n_pol] complex128,
//         n_pol 1 (scalar), 2 (diagonal) or 4 (full Jones)
//   channel_block_start [n_channel_blocks + 1], antenna1/2 [n_baselines];
//         starts from 0 to n_channels, increasing (every block holds a
//         channel, as DDECal's blocks do; the host form rejects others);
//         at most 1024 channel blocks
// The _device form takes device pointers for everything and enqueues on
// `stream` (a hipStream_t, NULL for the default stream); the host form copies
// host arrays in and out and reports the kernel's device time. Its device
// copies are kept per host thread and reused by later calls of the same or a
// smaller size (freed when the thread ends).
int ddh_correct_and_subtract_device(int32_t n_times, int32_t n_baselines, int32_t n_channels,
                                    int32_t n_directions, int32_t n_antennas,
                                    int32_t n_channel_blocks, int32_t n_pol,
                                    const int32_t* antenna1, const int32_t* antenna2,
                                    const int32_t* channel_block_start, const double* solutions,
                                    float* data, float* model, int32_t subtract,
                                    int32_t keep_model, void* stream);
int ddh_correct_and_subtract(int32_t n_times, int32_t n_baselines, int32_t n_channels,
                             int32_t n_directions, int32_t n_antennas, int32_t n_channel_blocks,
                             int32_t n_pol, const int32_t* antenna1, const int32_t* antenna2,
                             const int32_t* channel_block_start, const double* solutions,
                             float* data, float* model, int32_t subtract, int32_t keep_model,
                             double* kernel_ms);

// Warm-start network (no counterpart in the reference): a 3-layer
// perceptron with bf16 weights on the GPU (bf16 MFMA GEMMs, FP32
// accumulation, FP32 master weights, Adam). Parameters are one flat FP32
// array: W1 [n_in x h], b1 [h], W2 [h x h], b2 [h], W3 [h x n_out], b3 [n_out].
ddh_mlp* ddh_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint32_t seed);
void ddh_mlp_destroy(ddh_mlp* net);
int64_t ddh_mlp_num_params(const ddh_mlp* net);
int ddh_mlp_get_params(const ddh_mlp* net, float* params);
int ddh_mlp_set_params(ddh_mlp* net, const float* params);
int ddh_mlp_forward(ddh_mlp* net, const float* x, int64_t n, float* y);
// Forward passes (ddh_mlp_forward and the batched warm start) of a
// 16 -> 32k -> 32k -> 4 network (k = 1..8) run as one kernel by default
// (enable = 1): the same MFMA steps over the same bf16 operands as the
// layer-by-layer GEMMs (enable = 0), hence the same outputs, without the
// hidden activations' HBM traffic. Other shapes always run layer by layer.
int ddh_mlp_set_fused_forward(ddh_mlp* net, int32_t enable);
// Mean squared error over n rows and its gradient w.r.t. the parameters.
int ddh_mlp_loss_and_grads(ddh_mlp* net, const float* x, const float* t, int64_t n,
                           float* loss, float* grads);
int ddh_mlp_train(ddh_mlp* net, const float* x, const float* t, int64_t n, int32_t epochs,
                  int32_t batch, float lr, uint32_t seed, float* last_epoch_loss);
// C [M x N] = op(A) op(B) through the bf16 MFMA GEMM (operands rounded to
// bf16, FP32 accumulation); op transposes when the flag is set (A given as
// [K x M], B as [N x K]).
int ddh_gemm_bf16(const float* a, const float* b, float* c, int32_t M, int32_t N, int32_t K,
                  int32_t trans_a, int32_t trans_b);
// Features per solution pair (16 floats) on the current interval: the
// undamped one-step update of the previous solutions g0 (next1, with its
// denominators den) and two later iterates ga, gb of the damped solver, each
// relative to g0; and the initial solutions prev * exp(pred) from the
// network's 4 outputs per pair (the batched warm start passes prev = gb).
// Scalar solutions go through these as pairs holding each gain twice; the
// batched warm start then scales both by the mean of the two polarizations'
// outputs, formed in float (0.5f * (a + b)), so the pair stays one gain.
int ddh_warmstart_features(const double* g0, const double* next1, const double* ga,
                           const double* gb, const double* den, int32_t n_cb, int32_t n_ant,
                           int32_t n_sub, float* features);
int ddh_warmstart_apply(const double* prev, const float* pred, int64_t n_pairs, double* init);
// The same for 2x2 solutions ([n_cb][n_ant][n_sub][4] complex128, the
// full-Jones layout; den two values per solution, the diagonal of the
// one-step update's 2x2 denominator): the features from the diagonal entries
// J_00, J_11, and the start with row p of prev times exp(pred_p) (a matrix
// with a non-finite entry starts from the identity). No counterpart in the
// reference.
int ddh_warmstart_features_2x2(const double* g0, const double* next1, const double* ga,
                               const double* gb, const double* den, int32_t n_cb, int32_t n_ant,
                               int32_t n_sub, float* features);
int ddh_warmstart_apply_2x2(const double* prev, const float* pred, int64_t n_solutions,
                            double* init);

// Bandwidth calibration: a plain copy and a plain streaming read over
// `bytes` (each repeated `repeats` times). copy_gbs counts bytes read plus
// written. No counterpart in the reference; used to put the solver's HBM
// figures next to what the device achieves on the simplest streaming kernels.
int ddh_calibrate_bandwidth(int64_t bytes, int32_t repeats, double* copy_gbs,
                            double* read_gbs);

#ifdef __cplusplus
}
#endif

#endif
