// Internal declarations shared by the HIP translation units of libcochlea_amd.
#ifndef CM_INTERNAL_H
#define CM_INTERNAL_H

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "cm_types.h"
#include "../../include/cochlea_amd.h"

namespace cm {

struct BmLaunch {
  cm_bm_scalars sc;
  const float *__restrict__ input;
  int64_t input_len;
  const float *__restrict__ sec;     // 8 x 256
  const float *__restrict__ gamma;   // sets x 256
  const cm_bm_item *__restrict__ items;
  const int32_t *__restrict__ order;  // optional launch order: wave w integrates items[order[w]]
  int32_t n_items;
  int32_t use_queue;
  int32_t *queue;                    // work counter (use_queue != 0)
  float *__restrict__ speeds;        // rows x 256
  int64_t rows;
  int32_t *failed_nodes;
  float *conv_time, *conv_loops, *conv_jacobi;
  int32_t *failed_blocks;
  float *conv_time_blocks, *conv_loops_blocks, *conv_jacobi_blocks;
  cm_item_stats *stats;
};

// Learned warm start of the predictor-pass pressure solve (warm_start = 2) and
// the capture of its training records (capture != 0, warm_start = 1 guesses).
struct WsArgs {
  // MFMA-packed bf16 weights: [k-tile pairs][n tiles][64 lanes] x 8 (lane l of
  // pair j, tile t: W[32 j + 4 (l/16) + q][16 t + l%16] then the same for k tile 2j+1)
  const uint4 *__restrict__ w1;      // [25][32][64] (K 769 padded to 800)
  const uint4 *__restrict__ w2;      // [16][32][64]
  const uint4 *__restrict__ w3;      // [16][16][64]
  const float *__restrict__ b1, *__restrict__ b2, *__restrict__ b3;
  float c_out;                       // target scale: p = p_lin + rms(r_lin) * c_out * S(y)
  int32_t use_net;                   // 1: evaluate the network and offer its guess
  int32_t capture;                   // 1: write records (training data; network output too when use_net)
  float *records;                    // capture: [items * per_item][CM_WS_RECORD] floats
  int32_t stride;                    // capture every stride-th committed step
  int32_t per_item;                  // capture slots per item
  int32_t net_every;                 // evaluate the network on every net_every-th round (>= 1)
  int32_t group;                     // lockstep group size 8, 12 or 16; 0: chosen from the launch size
};

// record layout (floats): features [0, 769), rms(r_lin) at 769, committed step
// index at 770, Y at [772, 1028), p_lin at [1028, 1284), network output at
// [1288, 1544) when the network ran
constexpr int CM_WS_IN = 769;
constexpr int CM_WS_RECORD = 1544;
constexpr int CM_WS_REC_Y = 772;
constexpr int CM_WS_REC_PLIN = 1028;
constexpr int CM_WS_REC_OUT = 1288;

__device__ __forceinline__ uint16_t cm_f32_to_bf16(float x) {
  // round to nearest even; NaN kept quiet
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// GELU, tanh form, shared by training and in-solver inference
__device__ __forceinline__ float cm_gelu(float x) {
  return 0.5f * x * (1.0f + tanhf(0.7978845608f * (x + 0.044715f * x * x * x)));
}

struct JndArgs {
  const float *lambda;      // 3 x rows x 256
  int64_t rows;             // plane stride
  int n_intervals;
  int interval_nodes;
  int head_nodes;
  int length_nodes;
  const float *dA;          // per local interval
  const int32_t *ref_index; // per local interval (itself for references)
  const int32_t *out_index; // per local interval: column in F_RA / FisherAISum
  int out_stride;           // columns per fiber group in F_RA / FisherAISum
  const double *nIHC;       // 256
  float W[3];
  double Fs;
  float T;                  // (float)(delta_alpha_factor / Fs)
  float Tlength;            // (float)(float(len) / Fs)
  int calc_ai, calc_rms;
  float *mean_rate;         // 3 x n_intervals x 256
  float *F_RA;
  float *FisherAISum;
  // GlobalCalculateJND's backup stages (src/cochlea.cu:3860-3918) written to
  // buffer when backup_stage is 1..6 (NULL / 0: none): 1 AvgMeanRate
  // [3][n_intervals]; 2 dMeanRate, 5 preFisherAITimeReduced, 6 CRLB_RA
  // [3][n_intervals][256]; 3 dLambda, 4 preFisherAI at the lambda nodes of
  // the interval bodies ([3][rows][256])
  int backup_stage;
  float *buffer;
};

void set_error(const char *fmt, ...);
int device_cus();
// learned warm start state (ws_abi.hip)
bool ws_wants_group_kernel(const cm_bm_scalars &sc);
int ws_args(const cm_bm_scalars &sc, int n_items, WsArgs *out);
int check(hipError_t e, const char *what);
hipStream_t stream();
float *bm_speeds_ptr();
size_t bm_result_nodes();
float *bm_input_ptr();
size_t bm_input_cap();
int bm_ensure_input(size_t n);

}  // namespace cm

extern "C" int cm_launch_bm(const cm::BmLaunch *a, int waves_per_block, int unroll, int sections_per_lane,
                            int grid_blocks, hipStream_t stream);
// a block over a workgroup of 2 or 4 waves (sections_per_lane 2 or 1)
extern "C" int cm_launch_bm_split(const cm::BmLaunch *a, int sections_per_lane, int unroll, hipStream_t stream);
// lockstep-group launch of the warm-start kernel; returns the group size used
extern "C" int cm_launch_bm_ws(const cm::BmLaunch *a, const cm::WsArgs *w, int n_cus, hipStream_t stream);
extern "C" int cm_launch_ihc_lambda(const float *d_bm, int64_t rows, const cm_ihc_interval *d_ivs,
                                    const cm_ihc_interval *h_ivs, int n_ivs, const cm_ihc_interval *d_iir_ivs,
                                    int n_iir_ivs, const double *d_ac_b, const double *d_ac_bpad,
                                    const double *d_ac_a, int n_ac, const double *d_dc, const double *d_dcpad, int n_dc, const double *d_damage, const float *d_A,
                                    const float *spont3, float lambda_sat, float eta_ac, float eta_dc, float scale_bm,
                                    float *d_scratch, float *d_lambda, float *d_stages, hipStream_t stream);
extern "C" int cm_launch_fix_lambda(float *d_lambda, float *d_buffer, int64_t rows, const float *spont3,
                                    hipStream_t stream);
extern "C" int cm_launch_jnd(const cm::JndArgs *a, hipStream_t stream);
extern "C" int cm_launch_calc_dA(const float *input, float *dA, int n, int interval_nodes, int head_nodes,
                                 int length_nodes, int offset_start, hipStream_t stream);
extern "C" int cm_launch_generate_input(const cm_jnd_profile *d_prof, const float *d_WN, const float *d_Signal,
                                        float *d_out, double wn_dc, double wn_norm, int signal_mode, double sig_dc,
                                        double sig_norm, int start_profile, int n_profiles, int interval_len,
                                        int head, float Fs, hipStream_t stream);
extern "C" int cm_launch_hearing_aid(const float *d_in, float *d_tmp, float *d_out, int n_profiles, int interval_len,
                                     const float *d_fir, int fir_n, const float *d_iir, int iir_n,
                                     hipStream_t stream);

#endif
