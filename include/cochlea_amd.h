/*
 * cochlea_amd.h -- C-ABI of libcochlea_amd.so, the MI355X (gfx950) drop-in for
 * the GPU boundary of YonatanKoral/CochlearModel.
 *
 * The reference's boundary is include/cochlea_common.h (extern "C" blocks at
 * :86-185 and :194-418), implemented in src/cochlea.cu.  Every function below
 * keeps the reference name and argument meaning; what changes:
 *   - plain C types only (no Log&, cudaEvent_t&, cudaDeviceProp, no C++
 *     exceptions across the boundary);
 *   - every call returns 0 on success and a negative status on failure
 *     (-1 runtime/HIP error, -3 invalid argument); cm_last_error() describes it;
 *   - only the Run_Fast_BM_Calculation=17 method is provided.
 * Device memory is owned by the library (grow-only holders, as in
 * src/cochlea.cu:63-631).  Single host thread; calls are not reentrant.
 * Entries prefixed cm_ have no reference counterpart (batched launches,
 * timing, the learned warm start).
 */
#ifndef COCHLEA_AMD_H
#define COCHLEA_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "../cochlearmodel_amd/csrc/cm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* device_jnd_params of include/cochlea_common.h:57-67 (same field order and
 * x86-64 layout; the reference's bool is an int8 here). */
typedef struct {
  double dA;
  double recipdA;
  double dBSPLSignal;
  double dBSPLNoise;
  double Wn;
  int8_t isReference;
  int8_t pad0[3];
  int32_t calculateIndexPosition;
  int32_t referenceIndex;
  int32_t pad1;
  double frequency;
} cm_jnd_profile;

/* ---- diagnostics and streams (no reference counterpart) ---------------- */
const char *cm_last_error(void);
int cm_set_stream(void *hip_stream);
int cm_device_synchronize(void);
/* replace viewGPUStatus and the cudaEvents* timers (include/cochlea_common.h:96-99,
 * 289; src/cochlea.cu:708-746): events are hipEvent_t handles passed as void*,
 * the prefix a C string; cudaEventsDestroy releases a pair */
int viewGPUStatus(int flags, const char *prefix);
int cudaEventsCreate(void **start, void **stop, int condition);
int cudaEventsStartTimer(void *start, void *stop, int condition);
int cudaEventsStopTimer(void *start, void *stop, int condition);
float cudaEventsStopQueryTimer(void *start, void *stop, int condition, const char *prefix);
int cudaEventsDestroy(void *start, void *stop);
/* Duration of the last timed BM launch (Show_Run_Time & 1, or timed=1). */
float cm_last_kernel_ms(void);
/* BM launch shape: waves (time blocks) per workgroup, persistent queue. */
int cm_set_bm_launch(int waves_per_block, int persistent);
/* Jacobi sweep loop unroll (1, 2 or 4; default 2) of the following bm_kernel
 * launches (no reference counterpart: the reference's sweep loop is
 * include/cochlea.cuh:13962-13982).  Results are identical for every value;
 * cm_bm_autotune sets the fastest. */
int cm_set_bm_jacobi_unroll(int unroll);
/* Sections per lane (1, 2 or 4; default 4) of the following fused BM launches
 * (the learned lockstep kernel keeps 4).  4: one wave64 per time block, 4
 * sections per lane; 2 or 1: a block over a workgroup of 2 or 4 waves, the
 * wave-boundary pressures and partial reductions exchanged through LDS (the
 * reference runs one section per thread in 256-thread blocks,
 * include/cochlea.cuh:13708-13736).  Results are identical for every value;
 * cm_bm_autotune sets the fastest. */
int cm_set_bm_sections_per_lane(int sections_per_lane);
/* Solver of the BMOHCNewKernel launches (no reference counterpart): jacobi_mode
 * 0 = the reference's fixed sweeps (default), 1 = Jacobi to relative residual
 * tol_rel checked every check_every sweeps, capped at max_sweeps; warm_start 0
 * previous pressure, 1 extrapolated, 2 learned (needs cm_ws_load*). */
int cm_set_bm_solver(int jacobi_mode, float tol_rel, int check_every, int max_sweeps, int warm_start);
/* Launch order of the next BM launches over an n-block table: wave w
 * integrates block h_order[w] (a permutation; results do not depend on it).
 * NULL clears it. */
int cm_set_bm_order(const int32_t *h_order, int n);
/* Times the launch configurations (4 sections per lane: Jacobi unroll x waves
 * per workgroup x natural or longest-first order; 2 and 1 sections per lane:
 * Jacobi unroll x order) on every trial_stride-th block of the table (taken in
 * longest-first order) and reports the fastest; installs its sections per
 * lane and unroll, and the longest-first order when it wins.  Results are
 * identical for every configuration.
 * h_items: the block table on the host; d_speeds receives the trial rows. */
int cm_bm_autotune(const cm_bm_scalars *sc, const float *d_input, int64_t input_len, const float *d_sec,
                   const float *d_gamma, const cm_bm_item *h_items, int n_items, float *d_speeds, int64_t rows,
                   int trial_stride, cm_bm_tuning *out);

/* ---- BM velocity ------------------------------------------------------- */
/* replaces BMOHCKernel_Init, include/cochlea_common.h:209-230 */
int BMOHCKernel_Init(const float *gamma, const float *mass, const float *M, const float *U, const float *L,
                     const float *R, const float *S, const float *Q, const float *S_ohc, const float *S_tm,
                     const float *R_tm, float num_frequencies, float dbA, size_t inputBufferNodes,
                     size_t resultBufferNodes, size_t lambdaBufferNodes, int first_time, int Show_Run_Time,
                     int Show_Device_Data);

/* replaces BMOHCNewKernel, include/cochlea_common.h:126-176 (the reference's
 * tres pointer, cudaEvent_t&, cudaDeviceProp and Log& arguments are dropped) */
int BMOHCNewKernel(const float *input_samples, int override_input_samples, float w_ohc, float time_step,
                   float time_step_out, float delta_x, float alpha_r, int enable_psi, int enable_OW, int base_index,
                   float Ts, float _ohc_alpha_l, float _ohc_alpha_s, float model_Gme, float model_a0, float model_a1,
                   float model_a2, float sigma_ow, float eta_1, float eta_2, int Time_Blocks, int samplesBufferLengthP1,
                   int overlap_nodes_for_block, long overlapTimeMicroSec, int show_transient,
                   float cuda_max_time_step, float cuda_min_time_step, int Decouple_Filter,
                   float Max_M1_SP_Error_Parameter, float Max_Tolerance_Parameter, int Relative_Error_Parameters,
                   int calculate_boundary_conditions, float M1_SP_Fix_Factor, float Tolerance_Fix_Factor,
                   float SPLREfVal, int Show_Calculated_Power, int Show_Device_Data, int Show_Run_Time,
                   int JACOBBY_Loops_Fast, int JACOBBY_Loops_Slow, int Cuda_Outern_Loops,
                   int Run_Fast_BM_Calculation, int BMOHC_Kernel_Configuration);

/* Batched BM launch over an explicit block table (device pointers).  Rows of
 * d_speeds are input-sample indices; any per-row / per-block output may be
 * NULL.  No reference counterpart: the reference runs one parameter set per
 * launch (src/solver.cpp:312).  The kernel reads input only inside
 * [0, input_len) (zero elsewhere) and writes rows only inside [0, rows); each
 * item's set must index a row of d_gamma (the table is device memory, which
 * the host does not read). */
int cm_bm_run(const cm_bm_scalars *sc, const float *d_input, int64_t input_len, const float *d_sec,
              const float *d_gamma, const cm_bm_item *d_items, int n_items, float *d_speeds, int64_t rows,
              int32_t *d_failed_nodes, float *d_conv_time, float *d_conv_loops, float *d_conv_jacobi,
              int32_t *d_failed_blocks, float *d_conv_time_blocks, float *d_conv_loops_blocks,
              float *d_conv_jacobi_blocks, cm_item_stats *d_stats, int waves_per_block, int persistent, int timed);
int cm_get_last_blocks(cm_bm_item *out, int max_blocks);
int cm_get_last_stats(cm_item_stats *out, int max_blocks);

/* replaces BMOHCKernel_Wait_Threads, include/cochlea_common.h:179
 * (src/cochlea.cu:819-822): device-wide synchronisation */
int BMOHCKernel_Wait_Threads(void);
/* replaces BMOHCKernel_Copy_Results (src/cochlea.cu:823-827) */
int BMOHCKernel_Copy_Results(float *target, size_t resultNodes, size_t offset);
/* replaces ReverseKernel_Copy_Results (src/cochlea.cu:828-833) */
int ReverseKernel_Copy_Results(const float *src, size_t size);
/* replaces extractConvergenceTimes (src/cochlea.cu:456-458) */
int extractConvergenceTimes(float *convergence_times, size_t nodes);
/* replace the getCuda* pointer getters (src/cochlea.cu:3645-3676) */
void *getCudaFailedTimeNodes(void);
void *getCudaFailedBlocks(void);
void *getCudaConvergeSpeedBlocks(void);
void *getCudaConvergedTimeNodes(void);
void *getCudaConvergedBlocks(void);
void *getCudaConvergedJacobyLoopsPerIteration(void);
void *getCudaConvergedJacobyLoopsPerIterationBlocks(void);
void *getCudaBMVelocity(void);
void *getCudaInputSamples(void);
/* replace GeneralKernel_Copy_Results[_Double] and the
 * GeneralKernel_Copy_Results_Template<T> family (src/cochlea.cu:3678-3693) */
int GeneralKernel_Copy_Results_Bytes(void *target, const void *src, size_t bytes, size_t offset_bytes);
int GeneralKernel_Copy_Results(float *target, const float *src, size_t size);
int GeneralKernel_Copy_Results_Double(double *target, const double *src, size_t size);
/* replaces setupToleranceProfile (src/cochlea.cu:4256-4287) */
int setupToleranceProfile(const cm_jnd_profile *profiles, int is_first_time_for_parameters_set,
                          float Max_M1_SP_Error_Parameter, float Max_Tolerance_Parameter,
                          int Relative_Error_Parameters, float M1_SP_Fix_Factor, float Tolerance_Fix_Factor,
                          int Blocks_Per_Interval, int from_profile_index, int calculatedProfiles);
/* replaces BMOHCKernel_Free (src/cochlea.cu:815-822) */
int BMOHCKernel_Free(void);

/* ---- IHC / lambda ------------------------------------------------------ */
/* replaces IHCNewKernel, include/cochlea_common.h:238-274 (src/cochlea.cu:2417-2598);
 * Nerves_Clusters = {spont[3], unused[3], W[3]}; lambda_offset must be 0 */
int IHCNewKernel(const double *IHC_Damage_Factor, const float *Nerves_Clusters, const double *dc_filter,
                 int order_of_dc_filter, const double *ac_b_filter, const double *ac_a_filter, int is_ac_iir_filter,
                 int order_of_ac_filter, int cochlea_sections, int time_blocks, double SPLRefVal,
                 const float *backup_speeds, int backup_speeds_length, int calcTime, int writeTime, int allocateTime,
                 int intervalTimeNodes, int max_backup_nodes_len, int lambda_offset, float Lambda_SAT, float eta_AC,
                 float eta_DC, int first_time, int first_time_for_param_set, int loadedFromHD,
                 int disable_advanced_memory_handling, int review_memory_handling, int asMemoryHandlingOnly,
                 float scaleBMVelocityForLambdaCalculation, int CalculateJNDOnGPU, int maxJNDIntervals,
                 int overlapNodes, int Decouple_Filter, int Show_Run_Time);
/* replace loadAihc / enableloadAihc (src/cochlea.cu:449-455): 3 x 256 floats */
int loadAihc(const float *Aihc);
int enableloadAihc(void);
/* replace allocateBuffer / releaseBuffer (src/cochlea.cu:2599-2628) */
int allocateBuffer(int size_in_nodes, int Show_Run_Time);
int releaseBuffer(int Show_Run_Time);
/* replaces RunIHCKernel, include/cochlea_common.h:276 (src/cochlea.cu:3573-3644);
 * the IHC stage is not written over the BM velocity buffer */
int RunIHCKernel(float *host_backup, int Show_Run_Time, int save_lambda, int backup_stage,
                 int Decouple_Unified_IHC_Factor);
/* replaces BMOHCKernel_Copy_Lambda (src/cochlea.cu:834-837); reads the saved-lambda
 * buffer when this run filled it, else the lambda planes */
int BMOHCKernel_Copy_Lambda(float *target, size_t lambdaNodes, int offset);
/* replaces ReverseKernel_Copy_Results_Template<JNDFloat> (src/cochlea.cu:3691-3693,
 * include/cochlea_common.h:188); cuda_target: getCudaLambda(), getCudaBuffer() or the
 * BM velocity buffer */
int ReverseKernel_Copy_Results_Float(const float *cpu_src, float *cuda_target, size_t start_time_node,
                                     size_t time_nodes, int sections);
/* replaces updateCUDALambdaArray<JNDFloat> (src/cochlea.cu:4512-4540,
 * include/cochlea_common.h:191): lambda-stage resume fix-up on getCudaLambda() */
int updateCUDALambdaArray(float *lambda_array, float *cuda_buffer, size_t calc_time_nodes, int sections,
                          int Show_Run_Time, int Show_Device_Data, int cuda_buffer_update);
/* batched lambda fix-up over device pointers [3][rows][256] (no reference counterpart) */
int cm_fix_lambda_run(float *d_lambda, float *d_buffer, int64_t rows, const float *spont3);
/* replace getCudaLambda / getCudaBuffer / getCudaMeanRate */
void *getCudaLambda(void);
void *getCudaBuffer(void);
void *getCudaMeanRate(void);
/* batched IHC/lambda over device pointers and an interval table (no
 * reference counterpart); d_scratch holds 2 x rows x 256 floats (AC, dS) for
 * the two-pass kernels.  d_scratch = NULL runs the chain in one pass with no
 * workspace (slower, half the HBM traffic): FIR AC filters only, every
 * interval with kbase 0, at most 97 DC taps.  With an IIR
 * AC filter (d_ac_a) h_iir_ivs lists the runs the recursion restarts on:
 * rows [row0, row0 + len), period len (CudaIIRFilterDecoupled,
 * src/cochlea.cu:2885-2908).  n_ivs = 0 is a no-op; negative counts and
 * intervals outside [0, rows) return -3 before any device call. */
int cm_ihc_lambda_run(const float *d_bm, int64_t rows, const cm_ihc_interval *h_ivs, int n_ivs,
                      const cm_ihc_interval *h_iir_ivs, int n_iir_ivs, const double *d_ac_b, const double *d_ac_a,
                      int n_ac, const double *d_dc, int n_dc, const double *d_damage, const float *d_A,
                      const float *spont3, float lambda_sat, float eta_ac, float eta_dc, float scale_bm,
                      float *d_scratch, float *d_lambda, float *d_stages);

/* cm_ihc_lambda_run without any device allocation (no reference counterpart;
 * for callers that cannot allocate: code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The same inputs and outputs with
 * no workspace: the chain runs in one pass, its interval table and padded tap
 * tables staged in a fixed device array of the library (133 KiB per GPU, held
 * until the process ends).  Limits: FIR AC filters, at most 4096 intervals
 * (kbase 0), 512 AC taps, 97 DC taps; larger requests return -3.  Every call on
 * one GPU shares that array: calls that overlap on a GPU (several streams or
 * host threads) race. */
int cm_ihc_lambda_run_noalloc(const float *d_bm, int64_t rows, const cm_ihc_interval *h_ivs, int n_ivs,
                              const double *d_ac_b, int n_ac, const double *d_dc, int n_dc, const double *d_damage,
                              const float *d_A, const float *spont3, float lambda_sat, float eta_ac, float eta_dc,
                              float scale_bm, float *d_lambda, float *d_stages);

/* ---- JND --------------------------------------------------------------- */
/* replaces CudaCalculateJND, include/cochlea_common.h:292-328 (src/cochlea.cu:3946-4065);
 * backup_stage 1..6 (AvgMeanRate, dMeanRate, dLambda, preFisherAI,
 * preFisherAITimeReduced, CRLB_RA; src/cochlea.cu:3860-3918) writes that
 * stage to getCudaBuffer(): [3][JNDIntervals] for 1, [3][JNDIntervals][256]
 * for 2, 5 and 6, and the interval bodies of the [3][rows][256] lambda layout
 * for 3 and 4 */
int CudaCalculateJND(int calculate_ai, int calculate_rms, int mean_size, int fisher_size, double SPLRefVal, double Fs,
                     double scaleBMVelocityForLambdaCalculation, const double *nIHC,
                     const int *JND_Calculated_Intervals, int numOFJNDCalculated, const int *JND_Refrence_Intervals,
                     int numOFJNDReferences, int handeledIntervalsJND, int JNDIntervalsFull, int JNDIntervals,
                     int JNDIntervalHeadNodes, int overlapNodes, int JNDIntervalNodes, int lengthOffset,
                     const int *JND_Serial_Intervals_Positions, const int *JND_Interval_To_Reference, float *F_RA,
                     float *FisherAISum, int writeMatrixSize, int calculatedMatrixSize,
                     double JND_Delta_Alpha_Length_Factor, int JND_USE_Spont_Base_Reference_Lambda,
                     int Show_Run_Time, int calcdA, int show_generated_input_params_cuda, int backup_stage);
/* batched JND over device pointers (no reference counterpart).  Refuses (-3,
 * before any device call) negative sizes, an empty window (length_nodes < 1),
 * Fs <= 0, and intervals or windows beyond the lambda planes; n_intervals = 0
 * is a no-op.  d_ref_index / d_out_index are device arrays whose values the
 * caller keeps within the intervals and the output columns. */
int cm_jnd_run(const float *d_lambda, int64_t rows, int n_intervals, int interval_nodes, int head_nodes,
               int length_nodes, const float *d_dA, const int32_t *d_ref_index, const int32_t *d_out_index,
               int out_stride, const double *d_nIHC, const float *W3, double Fs, double delta_alpha_factor,
               int calc_ai, int calc_rms, float *d_mean_rate, float *d_F_RA, float *d_FisherAISum);
/* replaces AudioGramCreator::calcJNDFinal (src/AudioGramCreator.cpp:454-508), host FP64 */
int cm_jnd_final(const float *F_RA, const float *FisherAISum, int JNDIntervalsFull, int JNDIntervals,
                 int handled, const int *refs, int nrefs, const int *serial, double SPLRefVal, double max_db,
                 int calc_ai, int calc_rms, double *AiJND, double *RateJND, float *FisherAI, float *JND_RA);
/* replaces the Failed_Converged_Signals sums of calcJNDFinal
 * (src/AudioGramCreator.cpp:459-467, 490-493): per calculated interval, the
 * failed-block flags of its Decouple_Filter blocks, added into
 * failed_signals[serial[g]] */
int cm_jnd_failed_signals(const int32_t *failed_blocks, int n_blocks, int JNDIntervals, int handled, const int *refs,
                          int nrefs, const int *serial, int decouple_filter, int32_t *failed_signals);

/* Post-BM stages as one HIP graph (no reference counterpart): IHC/lambda with a
 * FIR AC filter (arguments as cm_ihc_lambda_run) and the JND reductions
 * (arguments as cm_jnd_run) captured once; replays read the BM velocity at
 * d_bm and rewrite lambda, mean rate and the Fisher sums.  Every device pointer
 * must stay valid while the graph exists. */
int cm_post_bm_graph_create(const float *d_bm, int64_t rows, const cm_ihc_interval *h_ivs, int n_ivs,
                            const double *d_ac_b, int n_ac, const double *d_dc, int n_dc, const double *d_damage,
                            const float *d_A, const float *spont3, float lambda_sat, float eta_ac, float eta_dc,
                            float scale_bm, float *d_scratch, float *d_lambda, int n_intervals, int interval_nodes,
                            int head_nodes, int length_nodes, const float *d_dA, const int32_t *d_ref_index,
                            const int32_t *d_out_index, int out_stride, const double *d_nIHC, const float *W3,
                            double Fs, double delta_alpha_factor, int calc_ai, int calc_rms, float *d_mean_rate,
                            float *d_F_RA, float *d_FisherAISum, void **handle);
int cm_post_bm_graph_launch(void *handle);
int cm_post_bm_graph_destroy(void *handle);

/* Complex-JND aggregation (Sim_Type 4), ComplexJNDProfile::calculateMinValue /
 * calculateGradientMinMaxValue (src/ComplexJNDProfile.cpp:19-95) as driven by
 * AudioGramCreator::calcComplexJND (src/AudioGramCreator.cpp:548-571).
 * intervals: n_profiles x n_levels indexes into raw (profile = noise * n_freq +
 * freq, built by cm_complex_jnd_intervals as CParams::updateInputProfile,
 * src/params.cpp:693-735); failed: per raw entry (may be NULL); methods bit 0 =
 * minimum, bit 1 = gradient with threshold eps.  Outputs [method][profile]:
 * jnd_final (1111 when nothing qualifies), warnings, conv_warnings (the
 * jnd_final, jnd_final_warnings and jnd_final_signal_convergence_warning vectors
 * of src/solver.cpp:812-822). */
int cm_complex_jnd_intervals(int n_freq, int n_levels, int n_noises, int32_t *intervals);
int cm_complex_jnd(const double *raw, const int32_t *failed, int n_raw, const int32_t *intervals, int n_profiles,
                   int n_levels, int methods, float eps, float *jnd_final, int32_t *warnings, int32_t *conv_warnings);

/* AC filter design (no reference binary counterpart; replaces the firpm-based
 * equiripple branch of SubModel::analyzeACFilter, src/SubModel.cpp:153-360):
 * own Parks-McClellan low-pass (bands in cycles per sample) and the
 * reference's design procedure around it for a Function_Filter string */
int cm_remez_lowpass(int ntaps, double fpass, double fstop, double wpass, double wstop, int density, double *h,
                     double *delta, int *iterations);
int cm_design_ac_filter(double Fs, const char *function_filter, double *h, double *a, int max_taps, int *n_taps,
                        int *is_iir);

/* ---- input generation -------------------------------------------------- */
/* replace InputProfilesArrayInitialization / Termination (src/cochlea.cu:2391-2415) */
int InputProfilesArrayInitialization(int maxJNDIntervals, int wn_length, int signal_length, int signal_mode,
                                     int Show_Generated_Input);
int InputProfilesArrayTermination(void);
/* replaces uploadProfiles (src/cochlea.cu:4288-4297) */
int uploadProfiles(const cm_jnd_profile *profiles, int numOfProfiles, int profilesLoaded);
/* the host normalisation of src/cochlea.cu:4165-4254: out4 = {wn_dc, wn_norm, sig_dc, sig_norm} */
int cm_input_normalization(const float *WN, int wn_length, const float *Signal, int signal_length, int signal_mode,
                           int Normalize_Sigma_Type, int Normalize_Sigma_Type_Signal, double given_interval,
                           double remove_dc, int e0, int e1, int s0, int s1, float Fs, double *out4);
/* replaces generateInputFromProfile, include/cochlea_common.h:373-413 (src/cochlea.cu:4298-4510);
 * writes the BM input buffer */
int generateInputFromProfile(const cm_jnd_profile *profiles, const float *WN, int wn_length, const float *Signal,
                             int signal_length, int signal_mode, int Normalize_Sigma_Type,
                             int Normalize_Sigma_Type_Signal, double Normalize_Noise_Energy_To_Given_Interval,
                             double Remove_Generated_Noise_DC, int start_dc_expected_value_calculation,
                             int end_dc_expected_value_calculation, int start_dc_normalized_value_calculation,
                             int end_dc_normalized_value_calculation, int numOfProfiles, int profilesLoaded,
                             int from_profile_index, int calculatedProfiles, int overlapNodes, int IntervalLength,
                             int JND_Interval_Head, int JND_Interval_Actual_Length, float Fs,
                             int Show_Generated_Input, float *target_input, int Show_Generated_Configuration,
                             int is_first_time_for_parameters_set, float Max_M1_SP_Error_Parameter,
                             float Max_Tolerance_Parameter, int Relative_Error_Parameters, float M1_SP_Fix_Factor,
                             float Tolerance_Fix_Factor, int Blocks_Per_Interval, int Show_Run_Time,
                             const float *fir_transfer_function, int fir_transfer_function_length,
                             const float *iir_transfer_function, int iir_transfer_function_length);
/* batched input synthesis over device pointers (no reference counterpart):
 * d_out[p * interval_len + n] for profile start_profile + p, p < n_profiles.
 * Negative sizes or Fs <= 0 return -3 before any device call; an empty batch
 * is a no-op. */
int cm_generate_input_run(const cm_jnd_profile *d_profiles, const float *d_WN, const float *d_Signal, float *d_out,
                          double wn_dc, double wn_norm, int signal_mode, double sig_dc, double sig_norm,
                          int start_profile, int n_profiles, int interval_len, int head, float Fs);
/* replaces IHCKernel_Free (src/cochlea.cu:2629-2634) */
int IHCKernel_Free(void);

/* ---- host physical model ---------------------------------------------- */
/* replaces CModel::CModel vectors (src/model.cpp:72-86): 7 x 256 doubles
 * M, R, S, R_tm, S_tm, S_ohc, Q */
int cm_model_vectors(double *out7x256);
/* replaces CSolver::Init_Tri_Matrix (src/solver.cpp:155-176); any may be NULL */
int cm_tri_matrix(double *lower, double *mid, double *upper);
/* replaces the per-section packing of Run_Cuda (src/solver.cpp:357-376):
 * CM_SEC_COUNT x 256 floats in cm_types.h CM_SEC_* order */
int cm_pack_section_params(float *out);
/* scalar arguments Run_Cuda passes to BMOHCNewKernel (src/solver.cpp:629-659) */
int cm_default_bm_scalars(float Fs, cm_bm_scalars *out);

/* ---- learned warm start (no reference counterpart) ----------------------
 * Kernels of the 769 -> 512 -> 512 -> 256 GELU pressure-guess MLP.  Matrices
 * are row-major; bf16 buffers are passed as void* (uint16 bit patterns).
 * C[M x N] (fp32) = op(A) op(B) (+ bias[N]) (+ beta * C), op = transpose when
 * the flag is set; BF16 MFMA with FP32 accumulation. */
int cm_gemm_bf16(int transA, int transB, int M, int N, int K, const void *A, int lda, const void *B, int ldb,
                 float *C, int ldc, const float *bias, float beta);
/* G = gelu(Z) (tanh form) to bf16 and/or fp32 (either output may be NULL) */
int cm_gelu_fwd(const float *Z, void *Gbf, float *Gf, int64_t n);
/* dZ = dG * gelu'(Z) to bf16 and/or fp32 */
int cm_gelu_bwd(const float *dG, const float *Z, void *dZbf, float *dZf, int64_t n);
/* fp32 -> bf16, round to nearest even */
int cm_to_bf16(const float *x, void *y, int64_t n);
/* db[n] = sum over rows of dZ[M x N] */
int cm_colsum(const float *dZ, float *db, int M, int N);
/* dY = 2 (Y - T) / (M N) (fp32 and bf16); row_loss[m] = sum_n (Y - T)^2 */
int cm_mse_grad(const float *Y, const float *T, float *dY, void *dYbf, float *row_loss, int M, int N);
/* In-solver use (cm_bm_scalars.warm_start = 2, jacobi_mode = 1): lockstep
 * groups of time blocks evaluate the network at every predictor pass; the
 * guess p_lin + rms(r_lin) c_out S(y) (S = [1,2,1]/4) competes with the
 * previous pressure and its extrapolation p_lin by residual.
 * Upload a model: W1 769x512, W2 512x512, W3 512x256 row-major fp32 (packed to
 * bf16 MFMA fragments on the host), biases fp32, target scale c_out. */
int cm_ws_load(const float *W1, const float *b1, const float *W2, const float *b2, const float *W3, const float *b3,
               float c_out);
/* run the warm-start network on every k-th lockstep round only (k >= 1; 1 = every
 * step); other rounds take the better of the previous and extrapolated pressure
 * (no reference counterpart) */
int cm_ws_set_every(int k);
/* Lockstep group size of the learned kernel (no reference counterpart): 8, 12
 * or 16 blocks per workgroup, or 0 (default) to choose it from the launch
 * size.  Results do not depend on it. */
int cm_ws_set_group(int g);
/* Training records: with warm_start 1 or 2 (jacobi_mode 1) every cm_bm_run
 * also writes, for every stride-th committed step of item i (slot
 * i * per_item + step / stride), the network input (769 floats at 0),
 * rms(r_lin) (769), the step index (770), Y (256 at 772), p_lin (256 at 1028)
 * and, with warm_start 2, the network output (256 at 1288); 1544 floats per
 * record.  d_records = NULL switches capture off. */
int cm_ws_set_capture(float *d_records, int64_t n_records, int stride, int per_item);
/* cm_ws_load from a file of float32: W1, b1, W2, b2, W3, b3 (as above) and c_out */
int cm_ws_load_file(const char *path);
/* Exact training targets: (A^-1 Y - p_lin) / (rms(r_lin) c_out) per record, A
 * from the m row of d_sec, FP64 inverse.  d_targets: n x 256 floats. */
int cm_ws_targets(const float *d_sec, const float *d_records, int64_t n, float c_out, float *d_targets);
int cm_ws_release(void);
/* Adam step on fp32 master weights w (+ decoupled-free L2 term wd * w in the
 * gradient); refreshes the bf16 copy wbf when it is not NULL */
int cm_adam(float *w, void *wbf, const float *g, float *m1, float *m2, int64_t n, float lr, float b1, float b2,
            float eps, int step, float wd);

#ifdef __cplusplus
}
#endif
#endif
