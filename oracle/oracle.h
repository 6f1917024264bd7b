// CPU oracle for the DDECal iterative-diagonal solve path and the
// point/Gaussian predict that feeds it.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke
// check and the bench's CPU-baseline leg may load it. The product path
// (dp3_amd/) never links, loads or calls it.
//
// It restates, in plain C++, the reference algorithm of lofar-astron/DP3 6.6.0:
//   ddecal/gain_solvers/IterativeDiagonalSolver.cc:20-222   (solve loop)
//   ddecal/gain_solvers/IterativeScalarSolver.cc:30-226     (scalar solve)
//   ddecal/gain_solvers/SolverBase.cc:63-371, SolverBase.h:274-285
//                                                   (step, convergence, stall,
//                                                    NaN repair, result)
//   ddecal/gain_solvers/SolveData.h:204-244            (AddOrSubtractDirection)
//   ddecal/gain_solvers/SolveData.cc:18-31,77-188      (regular packing)
//   ddecal/gain_solvers/SolverTools.cc:48-167          (Weigh, AssignAndWeight)
//   base/Simulator.h:35-60, base/Simulator.cc:110-549  (predict arithmetic)
//   base/PointSource.cc:51-99                          (Stokes spectrum)
//   steps/OnePredict.cc:437-457                        (copy to DPBuffer)
//   ddecal/test/unit/SolverTester.cc:95-202,348-363    (test fixture data)
//   steps/DDECal.cc:992-1124                    (correct and subtract models)
//   ddecal/constraints/SmoothnessConstraint.cc:50-97, KernelSmoother.h
//                                                   (smoothness constraint)
//
// All entry points are extern "C" with plain pointers so that Python (ctypes)
// can drive them. Complex numbers are interleaved (re, im).
#ifndef DDECAL_ORACLE_H
#define DDECAL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n_antennas;
  int32_t n_directions;
  int32_t n_channel_blocks;
  int32_t n_correlations;  // 2 = Duo (MC2x2FDiag), 4 = Full (MC2x2F)
  const int32_t* n_solutions_per_direction;  // [n_directions]
  int32_t max_iterations;
  int32_t min_iterations;
  double accuracy;
  double constraint_accuracy;
  double step_size;
  int32_t detect_stalling;
  double step_diff_sigma;
  int32_t phase_only;
  int32_t fp64;       // 0: reference precision mix, 1: FP64 everywhere
  int32_t n_threads;  // 1: deterministic single thread
} orc_settings;

typedef struct {
  int64_t iterations;             // SolveResult::iterations
  int64_t constraint_iterations;  // SolveResult::constraint_iterations
  int32_t has_converged;
  int32_t n_step_magnitudes;
} orc_result;

// Smoothness constraint (SmoothnessConstraint.h:23-75, Gaussian kernel).
typedef struct {
  const double* frequencies;           // [n_cb] channel-block frequencies, ascending (Hz)
  double bandwidth_hz;                 // kernel size (> 0)
  double bandwidth_ref_frequency_hz;   // 0: the same kernel size at every frequency
  double spectral_exponent;
  int32_t kernel_truncation;
  const double* weights;               // [n_ant][n_cb] (SetWeights), NULL: ones
  const double* sub_solution_weights;  // [n_sub][n_ant][n_cb] (SetSubSolutionWeights) or NULL
  const double* antenna_factors;       // [n_ant] (SetAntennaFactors), NULL: ones
  const double* dd_factors;            // [n_sub] (SetDdSmoothingFactors), NULL: ones
} orc_smoothness;

// SmoothnessConstraint::Apply on solutions [n_cb][n_ant][n_sub][n_pol]
// complex128, in place.
int orc_smooth(const orc_smoothness* constraint, int32_t n_cb, int32_t n_ant, int32_t n_sub,
               int32_t n_pol, double* solutions);

// Channel-block data in the reference SolveData layout, concatenated over
// channel blocks. vis_offsets[cb] .. vis_offsets[cb+1] are the visibilities of
// block cb.
//   data     : [n_vis_total][n_corr] complex64
//   model    : per block [n_dir][n_vis_cb][n_corr] complex64, blocks
//              concatenated (block cb starts at vis_offsets[cb]*n_dir*n_corr)
//   ant1/ant2: [n_vis_total]
//   sol_map  : per block [n_dir][n_vis_cb] (same concatenation as model)
// solutions  : [n_cb][n_ant][n_sub][2] complex128, in = initial, out = result
// step_magnitudes (optional, may be NULL): capacity max_iterations + 1.
int orc_solve(const orc_settings* settings, const int64_t* vis_offsets,
              const float* data, const float* model, const uint32_t* ant1,
              const uint32_t* ant2, const uint32_t* sol_map, double* solutions,
              orc_result* result, double* step_magnitudes);
// IterativeScalarSolver (IterativeScalarSolver.cc:30-226): one solution
// polarization, solutions [n_cb][n_ant][n_sub] complex128; with a smoothness
// constraint when `smoothness` is not NULL. n_correlations 1 (single, the
// std::complex<float> instantiation), 2 or 4.
int orc_solve_scalar(const orc_settings* settings, const orc_smoothness* smoothness,
                     const int64_t* vis_offsets, const float* data, const float* model,
                     const uint32_t* ant1, const uint32_t* ant2, const uint32_t* sol_map,
                     double* solutions, orc_result* result, double* step_magnitudes);
// IterativeFullJonesSolver (IterativeFullJonesSolver.cc:18-224): 2x2
// solutions [n_cb][n_ant][n_sub][4] complex128 (row-major 00, 01, 10, 11),
// Full data only (n_correlations 4, else -2); with a smoothness constraint
// when `smoothness` is not NULL.
int orc_solve_full_jones(const orc_settings* settings, const orc_smoothness* smoothness,
                         const int64_t* vis_offsets, const float* data, const float* model,
                         const uint32_t* ant1, const uint32_t* ant2, const uint32_t* sol_map,
                         double* solutions, orc_result* result, double* step_magnitudes);
// orc_solve with a smoothness constraint applied to the next solutions after
// every step (SolverBase::ApplyConstraints, SolverBase.cc:166-190; the
// constraint is always satisfied).
int orc_solve_constrained(const orc_settings* settings, const orc_smoothness* smoothness,
                          const int64_t* vis_offsets, const float* data, const float* model,
                          const uint32_t* ant1, const uint32_t* ant2, const uint32_t* sol_map,
                          double* solutions, orc_result* result, double* step_magnitudes);

// One PerformIteration + SolveDirection pass for all blocks without step:
// writes next_solutions [n_cb][n_ant][n_sub][2] complex128 from solutions.
int orc_iterate_once(const orc_settings* settings, const int64_t* vis_offsets,
                     const float* data, const float* model, const uint32_t* ant1,
                     const uint32_t* ant2, const uint32_t* sol_map,
                     const double* solutions, double* next_solutions);

// FP64 chi^2 of the residual D - sum_d G_a M_d G_b^H over all blocks (the
// weighted data the solver saw).
double orc_chi2(int32_t n_ant, int32_t n_dir, int32_t n_cb, int32_t n_corr,
                int32_t n_sub, const int64_t* vis_offsets, const float* data,
                const float* model, const uint32_t* ant1, const uint32_t* ant2,
                const uint32_t* sol_map, const double* solutions);
// The same for full-Jones solutions [n_cb][n_ant][n_sub][4] (residual
// D - sum_d J_a M_d J_b^H); n_corr must be 4 (NaN otherwise).
double orc_chi2_full_jones(int32_t n_ant, int32_t n_dir, int32_t n_cb, int32_t n_corr,
                           int32_t n_sub, const int64_t* vis_offsets, const float* data,
                           const float* model, const uint32_t* ant1, const uint32_t* ant2,
                           const uint32_t* sol_map, const double* solutions);

// Restatement of SolveData's regular constructor (SolveData.cc:77-188).
// Inputs per time step t: data [n_t][n_bl][n_ch][4] complex64 and
// model [n_dir][n_t][n_bl][n_ch][4]. Outputs use the orc_solve layout.
// vis_offsets must hold n_cb+1 entries; call with data==NULL to get offsets.
// n_corr 4 (Full), 2 (Duo: XX, YY) or 1 (single: 0.5 (XX + YY)).
int orc_pack_solve_data(int32_t n_t, int32_t n_bl, int32_t n_ch, int32_t n_dir,
                        int32_t n_cb, int32_t n_corr, int32_t n_ant,
                        const int32_t* n_solutions_per_direction,
                        const int32_t* antennas1, const int32_t* antennas2,
                        const float* data, const float* model,
                        int64_t* vis_offsets, float* out_data, float* out_model,
                        uint32_t* out_ant1, uint32_t* out_ant2,
                        uint32_t* out_sol_map);

// Restatement of AssignAndWeight (SolverTools.cc:81-167), in place, with
// linear_weighting_mode = false. data [n_t][n_bl][n_ch][4], flags (uint8)
// and weights with the same shape, model [n_dir][n_t][n_bl][n_ch][4].
void orc_assign_and_weight(int32_t n_t, int32_t n_bl, int32_t n_ch,
                           int32_t n_dir, float* data, const uint8_t* flags,
                           const float* weights, float* model);

// Restatement of radec2lmn (Simulator.h:35-60).
void orc_radec2lmn(double ref_ra, double ref_dec, double ra, double dec,
                   double* lmn);

// One model component, as the predict consumes it.
typedef struct {
  double ra, dec;
  double I, Q, U, V;              // Stokes at reference frequency
  double ref_freq;                // spectral reference frequency
  int32_t n_spectral_terms;       // 0 = flat
  const double* spectral_terms;   // polynomial coefficients
  int32_t log_si;                 // logarithmic spectral index
  int32_t has_rm;                 // rotation measure
  double pol_fraction, pol_angle, rm;
  int32_t is_gaussian;
  double major_axis, minor_axis;  // FWHM, radians
  double position_angle;          // radians
  int32_t pa_is_absolute;
} orc_component;

// Restatement of Simulator::visit + OnePredict's Stokes-I / full copy:
// accumulates every component into buffer [n_bl][n_ch][n_corr] complex128
// (n_corr = 1 when stokes_i_only, else 4). station_uvw [n_st][3] (metres).
void orc_simulate(double ref_ra, double ref_dec, int32_t n_st, int32_t n_bl,
                  const int32_t* bl_p, const int32_t* bl_q, int32_t n_ch,
                  const double* freq, const double* chan_width,
                  const double* station_uvw, const double* scaled_ncp_uvw,
                  int32_t correct_time_smearing, int32_t correct_freq_smearing,
                  int32_t stokes_i_only, int32_t n_comp,
                  const orc_component* components, double* buffer);

// PointSource::stokes (PointSource.cc:51-99) for one frequency: out I,Q,U,V.
void orc_stokes(const orc_component* c, double freq, double* iquv);

// SolverTester::FillDdIntervalData fixture (SolverTester.cc:95-202) with
// SetDiagonalSolutions (:348-363): 50 antennas, 3 directions, 16 channels,
// 50 time steps, n_solutions_per_direction given (e.g. {1,2,5} or {1,1,1}).
// Writes unweighted data [n_t][n_bl][n_ch][4], model [n_dir][n_t][n_bl][n_ch][4]
// and input_solutions [n_ant][n_sub][2] complex128. Sizes are fixed by the
// fixture; query them with orc_solvertester_shape.
void orc_solvertester_shape(int32_t* n_ant, int32_t* n_dir, int32_t* n_ch,
                            int32_t* n_t, int32_t* n_cb);
void orc_solvertester_fill(const int32_t* n_solutions_per_direction,
                           float* data, float* model, double* input_solutions);
// SolverTester::SetScalarSolutions (SolverTester.cc:331-346) instead of
// SetDiagonalSolutions, then the same data (input_solutions [ant][sub][2]
// with equal polarizations).
void orc_solvertester_fill_scalar(const int32_t* n_solutions_per_direction, float* data,
                                  float* model, double* input_solutions);
// SolverTester::SetLeakageSolutions (SolverTester.cc:365-379: off-diagonal
// leakage terms per (antenna, solution)), then FillDdIntervalData with
// leakage_only (:159-172: J = [[1, l_0], [l_1, 1]] on both sides);
// input_solutions [ant][sub][2] holds (l_0, l_1).
void orc_solvertester_fill_leakage(const int32_t* n_solutions_per_direction, float* data,
                                   float* model, double* input_solutions);

// Restatement of DDECal::CorrectAndSubtractModels (steps/DDECal.cc:992-1124)
// over one solution interval: data [n_t][n_bl][n_ch][4] and model
// [n_dir][n_t][n_bl][n_ch][4] complex64 (unweighted), solutions
// [n_cb][n_ant][n_dir][n_pol] complex128 (n_pol 1 scalar, 2 diagonal,
// 4 full Jones), channel_block_start [n_cb + 1]. With subtract the corrected
// models are subtracted from the data; with keep_model they replace the
// model data.
void orc_correct_and_subtract(int32_t n_t, int32_t n_bl, int32_t n_ch, int32_t n_dir,
                              int32_t n_ant, int32_t n_cb, int32_t n_pol,
                              const int32_t* antenna1, const int32_t* antenna2,
                              const int32_t* channel_block_start, const double* solutions,
                              float* data, float* model, int32_t subtract, int32_t keep_model);

#ifdef __cplusplus
}
#endif

#endif
