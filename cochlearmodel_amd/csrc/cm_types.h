// Plain-C types shared by the HIP kernels, the C-ABI layer and the C++ host.
// They carry the argument meaning of the reference's BM launch
// (include/cochlea.cuh:13574-13620, src/cochlea.cu:990-1184) in struct form.
#ifndef CM_TYPES_H
#define CM_TYPES_H

#include <stdint.h>

#define CM_SECTIONS 256
#define CM_LAMBDA_COUNT 3

#ifdef __cplusplus
extern "C" {
#endif

// Scalar constants of one BM launch.
typedef struct {
  float w_ohc;
  float time_step;      // initial dt
  float time_step_out;  // sample clock: dt0 * int(1/(Fs*dt_max))
  float delta_x;
  float alpha_r;
  float Ts;
  float alpha_l;
  float alpha_s;
  float Gme;
  float a0, a1, a2;
  float sigma_ow;
  float eta1, eta2;
  float dt_min, dt_max;
  int32_t j_fast;        // JACOBBY_Loops_Fast (passes >= 2)
  int32_t j_slow;        // JACOBBY_Loops_Slow (passes 0, 1)
  int32_t max_passes;    // Cuda_Outern_Loops
  int32_t decision_mode; // 0: decision from the full reduction; 1: zero-partial lane outcome
  int32_t jacobi_mode;   // 0: fixed sweeps (reference); 1: sweep until the relative residual <= tol_rel
  int32_t check_every;   // tolerance mode: sweeps between residual checks
  int32_t max_sweeps;    // tolerance mode: sweep cap per pass
  float tol_rel;         // tolerance mode: |Y - A p|_2 <= tol_rel |Y|_2
  int32_t warm_start;    // tolerance mode, predictor pass guess: 0 previous pressure; 1 the better (by
                         // residual, ties to the previous pressure) of it and the linear extrapolation
                         // p + (dt/dt_prev)(p - p_prev) from the last two steps' pressures
} cm_bm_scalars;

// One time block (or one whole signal in continuous mode).
typedef struct {
  int32_t start;        // first input sample integrated from zero state
  int32_t mosi;         // first output row (out_sample_index)
  int32_t mesi;         // end row (end_sample_index)
  int32_t write_start;  // first row this block owns in per-row outputs
  float tol;            // throw tolerance
  float m1tol;          // max m1 sp tolerance
  int32_t set;          // parameter set (gamma vector) index
  int32_t pad;
} cm_bm_item;

// Per-section packed parameters: CM_SEC_COUNT x 256 floats in this order.
// CM_SEC_M is the tridiagonal's diagonal as float (used by the residual test).
enum { CM_SEC_MASS = 0, CM_SEC_RSM, CM_SEC_R, CM_SEC_S, CM_SEC_Q, CM_SEC_SOHC, CM_SEC_STM, CM_SEC_RTM, CM_SEC_M,
       CM_SEC_COUNT };

// Per-block statistics written by the BM kernel.
typedef struct {
  int64_t steps;   // committed steps
  int64_t passes;  // predictor/corrector passes
  int64_t sweeps;  // Jacobi sweeps
} cm_item_stats;

// Launch configuration chosen by cm_bm_autotune (all candidates give identical
// results; only the time differs).
typedef struct {
  int32_t waves_per_block;   // 1, 2 or 4 time blocks per workgroup
  int32_t order_lpt;         // 1: items launched longest first
  int32_t n_trial_items;     // items in the timed sample
  int32_t jacobi_unroll;     // 1, 2 or 4 sweeps per trip of the Jacobi loop
  float ms[18];              // trial times, 4 sections per lane, index 6 u + 2 w + o: unroll (2, 4, 1) x
                             // wpb (1, 2, 4) x (natural, longest first)
  int32_t sections_per_lane; // 4 (one wave per block), 2 or 1 (a block over 2 or 4 waves)
  float ms_split[12];        // trial times of 2 and 1 sections per lane, index 6 v + 2 u + o:
                             // (2, 1) x unroll (2, 4, 1) x (natural, longest first)
} cm_bm_tuning;

// A run of rows of the IHC/lambda stage sharing one FIR window rule:
// k = kbase + (row - row0); taps used = min((k+1) mod per, taps) with
// per = L when k < frdp-1 and per = tla afterwards (src/cochlea.cu:2719-2728).
typedef struct {
  int64_t row0;   // first row in the buffer
  int32_t len;    // rows in the interval
  int32_t kbase;  // k of row0
  int32_t L;      // interval length before the final regular division
  int32_t tla;    // analysed length after it
  int32_t frdp;   // final regular division position
  int32_t set;    // IHC damage vector index
} cm_ihc_interval;

#ifdef __cplusplus
}
#endif
#endif
