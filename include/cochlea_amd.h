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
/* Duration of the last timed BM launch (Show_Run_Time & 1, or timed=1). */
float cm_last_kernel_ms(void);
/* BM launch shape: waves (time blocks) per workgroup, persistent queue. */
int cm_set_bm_launch(int waves_per_block, int persistent);

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
 * launch (src/solver.cpp:312). */
int cm_bm_run(const cm_bm_scalars *sc, const float *d_input, int64_t input_len, const float *d_sec,
              const float *d_gamma, const cm_bm_item *d_items, int n_items, float *d_speeds, int64_t rows,
              int32_t *d_failed_nodes, float *d_conv_time, float *d_conv_loops, float *d_conv_jacobi,
              int32_t *d_failed_blocks, float *d_conv_time_blocks, float *d_conv_loops_blocks,
              float *d_conv_jacobi_blocks, cm_item_stats *d_stats, int waves_per_block, int persistent, int timed);
int cm_get_last_blocks(cm_bm_item *out, int max_blocks);
int cm_get_last_stats(cm_item_stats *out, int max_blocks);

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

/* ---- host physical model ---------------------------------------------- */
/* replaces CModel::CModel vectors (src/model.cpp:72-86): 7 x 256 doubles
 * M, R, S, R_tm, S_tm, S_ohc, Q */
int cm_model_vectors(double *out7x256);
/* replaces CSolver::Init_Tri_Matrix (src/solver.cpp:155-176); any may be NULL */
int cm_tri_matrix(double *lower, double *mid, double *upper);
/* replaces the per-section packing of Run_Cuda (src/solver.cpp:357-376):
 * 8 x 256 floats in cm_types.h CM_SEC_* order */
int cm_pack_section_params(float *out8x256);
/* scalar arguments Run_Cuda passes to BMOHCNewKernel (src/solver.cpp:629-659) */
int cm_default_bm_scalars(float Fs, cm_bm_scalars *out);

#ifdef __cplusplus
}
#endif
#endif
