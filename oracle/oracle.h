/*
 * CPU oracle for the cochlear hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is a plain CPU restatement of the reference's algorithm
 * (YonatanKoral/CochlearModel, kernel Run_Fast_BM_Calculation=17 and the
 * IHC / lambda / JND stages).  It exists so that tests, smoke() and the
 * bench's cpu_baseline leg can check the HIP product against it.  Nothing
 * in the product links, loads or calls it.
 *
 * Every function names the reference lines it follows.  Arithmetic is
 * IEEE single (or double for the FP64 rebuild) with no FMA contraction:
 * the build uses -ffp-contract=off and every fused multiply-add is written
 * as fmaf()/fma() explicitly.
 */
#ifndef COCHLEA_ORACLE_H
#define COCHLEA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SECTIONS 256
#define ORC_LAMBDA_COUNT 3

/* Scalar model constants of one BM launch (include/cochlea.cuh:13574-13620
 * argument list; values packed by src/solver.cpp:629-659). */
typedef struct {
  float w_ohc;
  float time_step;      /* initial dt */
  float time_step_out;  /* sample clock (Q9) */
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
  int32_t j_fast;       /* JACOBBY_Loops_Fast */
  int32_t j_slow;       /* JACOBBY_Loops_Slow */
  int32_t max_passes;   /* Cuda_Outern_Loops */
  int32_t decision_mode;/* 0 = lane-0 full reduction, 1 = zero-lane race outcome */
  int32_t jacobi_mode;  /* 0 = fixed sweeps (reference), 1 = to tolerance, 2 = exact FP64 Thomas solve (O2) */
  int32_t check_every;  /* tolerance mode: sweeps between residual checks */
  int32_t max_sweeps;   /* tolerance mode: sweep cap per pass */
  float tol_rel;        /* tolerance mode: stop when |Y - A p| <= tol_rel |Y| */
  int32_t warm_start;   /* tolerance mode, predictor pass: 0 = previous pressure; 1 = also try the
                           linear extrapolation p + (dt/dt_prev)(p - p_prev) and keep the guess
                           with the smaller residual (ties keep the previous pressure); 2 = also
                           the learned guess (orc_ws_set_net; FP32 run only) */
} orc_bm_scalars;

/* One work item = one reference "time block" (src/cochlea.cu:1148-1176). */
typedef struct {
  int32_t start;        /* first input sample integrated (zero state) */
  int32_t mosi;         /* gen_model_out_sample_index */
  int32_t mesi;         /* gen_model_end_sample_index */
  int32_t write_start;  /* first row this item owns in per-row outputs */
  float tol;            /* gen_model_throw_tolerance */
  float m1tol;          /* gen_model_max_m1_sp_tolerance */
  int32_t set;          /* parameter-set index (selects gamma) */
  int32_t pad;
} orc_bm_item;

/* Model vectors of CModel::CModel (src/model.cpp:72-86), FP64.
 * out layout: 7 x 256 doubles: M, R, S, R_tm, S_tm, S_ohc, Q. */
int orc_model_vectors(double *out7x256);
/* Scalar constants as the solver passes them (src/solver.cpp:629-659). */
int orc_default_scalars(float Fs, orc_bm_scalars *out);
/* Mid-diagonal of CSolver::Init_Tri_Matrix (src/solver.cpp:155-176). */
int orc_tri_mid_diag(double *m256);

/* Reference time-block decomposition (src/cochlea.cu:1148-1176) plus the
 * disjoint row ownership used for per-row outputs. */
int orc_block_decomposition(int time_blocks, int bufP1, int overlap, int decouple_filter,
                            int32_t *start, int32_t *mosi, int32_t *mesi, int32_t *write_start);

/* Kernel-17 restatement.  sec_f32 = 9 x 256 floats (mass, rsm, R, S, Q,
 * S_ohc, S_tm, R_tm, m = diagonal as float); gamma = n_sets x 256.  Outputs may be NULL. */
int orc_bm_run_f32(const orc_bm_scalars *sc, const float *sec_f32, const float *gamma,
                   const orc_bm_item *items, int n_items, const float *input, int64_t input_len,
                   float *speeds, int64_t rows,
                   int32_t *failed_nodes, float *conv_time, float *conv_loops, float *conv_jacobi,
                   int32_t *failed_blocks, float *conv_time_blocks, float *conv_loops_blocks,
                   float *conv_jacobi_blocks, int64_t *item_steps, int64_t *item_passes,
                   int64_t *item_sweeps, int n_threads);
/* FP64 rebuild of the same method (state and constants promoted from FP32). */
int orc_bm_run_f64(const orc_bm_scalars *sc, const float *sec_f32, const float *gamma,
                   const orc_bm_item *items, int n_items, const float *input, int64_t input_len,
                   double *speeds, int64_t rows, int64_t *item_steps, int64_t *item_passes,
                   int64_t *item_sweeps, int n_threads);

/* Step control of one pass (include/cochlea.cuh:14021-14088, lane-0 decision):
 * per-section vectors of 256 floats; dt in, then out5 = {dt after the pass,
 * L, e_a = max|a_ref - a|, e_v = sum|v - v_ref|, F = sum(|v - v_ref| - max(tol, v dt))}
 * and another_reset = {another_loop, reset_iteration}. */
int orc_step_control_f32(const float *aref, const float *na, const float *vref, const float *nv, float dt, float tol,
                         float m1tol, float dt_min, float dt_max, int decision_mode, float *out5,
                         int32_t *another_reset);

/* One record per predictor/corrector pass of a traced item. */
typedef struct {
  int32_t step;      /* committed steps before this pass's step */
  int32_t pass;      /* pass index within the step (0 = predictor) */
  int32_t in_s;      /* input sample at the pass */
  int32_t decided;   /* 1 when step control ran (every pass after the predictor) */
  float t;           /* local time at the pass */
  float dt_in, dt_out;
  float L, F, e_a, e_v;
  int32_t another, reset, capped;   /* after the decision and the pass cap */
  int32_t sweeps;    /* Jacobi sweeps of the pass (0 for the exact solve) */
} orc_pass_record;
/* Kernel-17 run of one item that records its passes (up to cap; the run stops
 * once the record is full). */
int orc_bm_trace_f32(const orc_bm_scalars *sc, const float *sec, const float *gamma, const orc_bm_item *item,
                     const float *input, int64_t input_len, orc_pass_record *rec, int cap, int *n_rec);

/* FP64 rebuild with the sample clock (t, time_step_out, Ts, the interpolation
 * weight) kept in float as the method has it: separates the clock's rounding
 * from the state's. */
int orc_bm_run_f64_clock32(const orc_bm_scalars *sc, const float *sec, const float *gamma, const orc_bm_item *items,
                           int n_items, const float *input, int64_t input_len, double *speeds, int64_t rows,
                           int64_t *item_steps, int64_t *item_passes, int64_t *item_sweeps, int n_threads);

/* Network of the learned predictor-pass guess (warm_start = 2 in the FP32
 * run; no counterpart in the reference): W1 769 x 512, W2 512 x 512, W3 512 x
 * 256 row-major float32 (rounded to bfloat16 here), biases 512, 512, 256,
 * output scale c_out.  Copied; process-wide. */
int orc_ws_set_net(const float *w1, const float *b1, const float *w2, const float *b2, const float *w3,
                   const float *b3, float c_out);

/* The network on n_rows feature rows (769 floats each, rounded to bfloat16
 * here) -> 256 outputs per row. */
int orc_ws_forward(const float *x, int n_rows, float *y);

/* Jacobi sweeps of the kernel convention (Q7) on one right-hand side. */
int orc_jacobi(const float *Y, const float *rsm, float *p, int sweeps);

/* tanh used by both the oracle and the GPU kernel (defined algorithm). */
float orc_tanhf(float x);
float orc_expf(float x);

/* IHC + lambda (src/cochlea.cu:2712-2790, 3200-3240, 3365-3408, 3431-3506).
 * bm: rows x 256 (row-major, section fastest).  lambda: 3 x rows x 256.
 * stages (optional, may be NULL): 5 x rows x 256 = AC, dS, DC, PRE_IHC, IHC. */
int orc_ihc_lambda(const float *bm, int64_t rows, int interval_len, int final_regular_division,
                   const double *ac_b, const double *ac_a /* NULL: FIR */, int n_ac, const double *dc_b, int n_dc,
                   const double *ihc_damage, const float *A3x256, const float *spont3,
                   float lambda_sat, float eta_ac, float eta_dc, float scale_bm,
                   float *lambda, float *stages, int n_threads);

/* JND Fisher reductions (src/cochlea.cu:3741-3944) and final dB
 * (src/AudioGramCreator.cpp:454-508).  backup_stage 1..6 with a buffer writes
 * that stage of GlobalCalculateJND (src/cochlea.cu:3815-3911): 1 AvgMeanRate
 * [3][n]; 2 dMeanRate, 5 preFisherAITimeReduced, 6 CRLB_RA [3][n][256];
 * 3 dLambda, 4 preFisherAI at the body nodes of [3][rows][256]. */
int orc_jnd(const float *lambda, int64_t rows, int n_intervals, int interval_nodes, int head_nodes,
            int length_nodes, const float *dA, const int32_t *ref_index, const double *nIHC,
            const float *W3, double Fs, double delta_alpha_factor, int calc_ai, int calc_rms,
            float *mean_rate, float *F_RA, float *FisherAISum, int backup_stage, float *buffer);

/* Profile input synthesis (src/cochlea.cu:4070-4103). */
int orc_generate_input(int n_profiles, const double *dA, const double *Wn, const double *freq,
                       const float *WN, double wn_dc, double wn_norm, double sig_norm,
                       int interval_len, int head, float Fs, float *out);

#ifdef __cplusplus
}
#endif
#endif
