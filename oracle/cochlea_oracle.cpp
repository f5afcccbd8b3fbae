/*
 * CPU oracle for the cochlear hot path -- TEST INFRASTRUCTURE ONLY.
 * Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker; the product never links or calls it.
 *
 * It restates, in plain scalar C++, what the reference computes:
 *   - CModel vectors          src/model.cpp:72-86, include/const.h:216-255
 *   - tridiagonal diagonal    src/solver.cpp:155-176 (kernel row convention Q7)
 *   - time-block layout       src/cochlea.cu:1148-1176
 *   - BM/OHC kernel #17       include/cochlea.cuh:13573-14245 (+ reductions :149-186)
 *   - IHC + lambda            src/cochlea.cu:2712-2790, 2980-2990, 3200-3240, 3365-3408, 3431-3506
 *   - JND Fisher sums         src/cochlea.cu:3741-3944
 *   - profile input synthesis src/cochlea.cu:4070-4103
 * and the learned predictor-pass guess this build adds to kernel 17 (no
 * counterpart in the reference): the same features, network, smoothing and
 * residual gate as the GPU's lockstep kernel, with the network on the host.
 *
 * Arithmetic contract shared with the GPU kernels (not with their code):
 *   * FP32 state, every operation IEEE-rounded, no FMA contraction
 *     (built with -ffp-contract=off); __fdividef of the reference is an
 *     IEEE division here.
 *   * tanhf is orc_tanhf below (a defined algorithm, checked against the
 *     FP64 libm tanh in the tests), so CPU and GPU agree bit for bit.
 *   * The three step-control reductions over 256 sections use one fixed
 *     tree: 4 consecutive sections per lane summed as (s0+s1)+(s2+s3),
 *     xor-butterfly over lanes with strides 1,2,4,8, then
 *     (r0+r16)+(r32+r48) over the four 16-lane rows.
 *   * Step-control race Q1 resolved as lane 0 (full reduction) unless
 *     decision_mode==1, which reproduces what a zero-partial lane decides.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NS = ORC_SECTIONS;

// ---------------------------------------------------------------- tanh ---
// exp by Cody-Waite reduction y = k*ln2 + r, |r| <= ln2/2, Taylor to r^8.
float exp_reduced(float y) {
  const float log2e = 1.44269502e+00f;
  const float ln2_hi = 6.93145752e-01f;   // high bits of ln2
  const float ln2_lo = 1.42860677e-06f;   // ln2 - ln2_hi
  float k = rintf(y * log2e);
  float r = fmaf(-k, ln2_hi, y);
  r = fmaf(-k, ln2_lo, r);
  float p = 2.48015873e-05f;              // 1/8!
  p = fmaf(p, r, 1.98412701e-04f);        // 1/7!
  p = fmaf(p, r, 1.38888892e-03f);        // 1/6!
  p = fmaf(p, r, 8.33333377e-03f);        // 1/5!
  p = fmaf(p, r, 4.16666679e-02f);        // 1/4!
  p = fmaf(p, r, 1.66666672e-01f);        // 1/3!
  p = fmaf(p, r, 5.00000000e-01f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

// tanh: odd Taylor series below 0.55, 1 - 2/(e^{2a}+1) up to 10, 1 above.
float tanh_defined(float x) {
  float a = fabsf(x);
  float r;
  if (a < 0.55f) {
    float z = a * a;
    float q = -2.39129114e-04f;           // -443861162/1856156927625
    q = fmaf(q, z, 5.90027441e-04f);      // 6404582/10854718875
    q = fmaf(q, z, -1.45583438e-03f);     // -929569/638512875
    q = fmaf(q, z, 3.59212812e-03f);      // 21844/6081075
    q = fmaf(q, z, -8.86323989e-03f);     // -1382/155925
    q = fmaf(q, z, 2.18694881e-02f);      // 62/2835
    q = fmaf(q, z, -5.39682540e-02f);     // -17/315
    q = fmaf(q, z, 1.33333340e-01f);      // 2/15
    q = fmaf(q, z, -3.33333343e-01f);     // -1/3
    r = fmaf(a * z, q, a);
  } else if (a < 10.0f) {
    float e = exp_reduced(a + a);
    r = 1.0f - 2.0f / (e + 1.0f);
  } else {
    r = 1.0f;
  }
  return copysignf(r, x);
}

// ------------------------------------------------------ math adapters ---
template <typename T> struct M;
template <> struct M<float> {
  static float tanh_(float x) { return tanh_defined(x); }
};
template <> struct M<double> {
  static double tanh_(double x) { return std::tanh(x); }
};

// Step-control reductions (fixed tree, see header comment).
template <typename T> T tree_sum(const T *x) {
  T lane[64], tmp[64];
  for (int l = 0; l < 64; ++l) lane[l] = (x[4 * l] + x[4 * l + 1]) + (x[4 * l + 2] + x[4 * l + 3]);
  for (int k = 1; k <= 8; k <<= 1) {
    for (int l = 0; l < 64; ++l) tmp[l] = lane[l] + lane[l ^ k];
    std::memcpy(lane, tmp, sizeof(lane));
  }
  return (lane[0] + lane[16]) + (lane[32] + lane[48]);
}
template <typename T> T tree_max(const T *x) {
  T m = x[0];
  for (int i = 1; i < NS; ++i) m = std::fmax(m, x[i]);
  return m;
}

// Exact solve of the kernel-convention tridiagonal (row i: low_i p_{i-1} +
// m_i p_i + p_{i+1} = Y_i, p_{-1} = p_256 = 0) by the Thomas algorithm in FP64.
template <typename T> void thomas_exact(const T *Y, const T *mf, const T *low, T *p) {
  double c[NS], d[NS];
  for (int i = 0; i < NS; ++i) {
    const double a = i > 0 ? (double)low[i] : 0.0;
    const double b = (double)mf[i];
    const double up = i < NS - 1 ? 1.0 : 0.0;
    const double den = i > 0 ? b - a * c[i - 1] : b;
    c[i] = up / den;
    d[i] = ((double)Y[i] - (i > 0 ? a * d[i - 1] : 0.0)) / den;
  }
  double x = d[NS - 1];
  p[NS - 1] = (T)x;
  for (int i = NS - 2; i >= 0; --i) {
    x = d[i] - c[i] * x;
    p[i] = (T)x;
  }
}

// ------------------------------------------------------ step control -----
// Mutants: -DORC_MUTANT=k compiles one plausible transcription mistake of the
// step-control ladder into this oracle, so that a test can show the ladder
// pins reject it.  The default build (ORC_MUTANT 0) is the restatement.
//   1 middle band 0.5 <= L <= 6, F <= 0 accepts instead of iterating
//   2 ode::Tester thresholds (L > 2 halves, L < 0.25 doubles)
//   3 F > 0 tested before L > 6
//   4 |v| instead of signed v inside fmax(tol, v dt)
//   5 F counts violating sections instead of summing their excess
//   6 a halving pass re-runs the step instead of committing it
//   7 a halving step advances local time by the old dt
//   8 e_v >= m1tol (not strict) enables L
//   9 e_a summed instead of maximised
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

// Outcome of one pass's step control.
template <typename T> struct Decision {
  T emax, esum, fsum, L;
  int another, reset, rerun;
};

// The Lipschitz / fault test of kernel #17 (include/cochlea.cuh:14021-14088)
// as lane 0 computes it (Q1): the per-section terms
//   ex = |a_ref - a|, ey = |v - v_ref|, ez = ey - max(tol, v dt) (v signed),
// reduced to e_a = max ex, e_v = sum ey, F = sum ez by the fixed tree; then
//   L = e_a dt / e_v if e_v > m1tol else 0, and the ladder
//   L > 6: dt <- max(dt/2, dt_min), exit (the step commits, Q2)
//   F > 0: iterate
//   L < 0.5: dt <- min(2 dt, dt_max), exit
//   else (0.5 <= L <= 6, F <= 0): iterate (no middle-band accept, Q3).
// decision_mode 1 replaces the three reductions by zeros (what a lane
// holding no partial sums decides).  dt is updated in place.
template <typename T>
Decision<T> step_control(const T *aref, const T *na, const T *vref, const T *nv, T &dt, T mtt, T mmspt, T dt_min,
                         T dt_max, int decision_mode) {
  T ex[NS], ey[NS], ez[NS];
  for (int i = 0; i < NS; ++i) {
    ex[i] = std::fabs(aref[i] - na[i]);
    ey[i] = std::fabs(nv[i] - vref[i]);
#if ORC_MUTANT == 4
    ez[i] = ey[i] - std::fmax(mtt, std::fabs(nv[i]) * dt);
#else
    ez[i] = ey[i] - std::fmax(mtt, nv[i] * dt);
#endif
#if ORC_MUTANT == 5
    ez[i] = ez[i] > 0 ? (T)1 : (T)0;
#endif
  }
  Decision<T> d;
#if ORC_MUTANT == 9
  d.emax = tree_sum(ex);
#else
  d.emax = tree_max(ex);
#endif
  d.esum = tree_sum(ey);
  d.fsum = tree_sum(ez);
  d.rerun = 0;
  if (decision_mode == 1) { d.emax = 0; d.esum = 0; d.fsum = 0; }
  T L;
#if ORC_MUTANT == 8
  if (d.esum >= mmspt) { L = d.emax * dt; L = L / d.esum; } else L = 0;
#else
  if (d.esum > mmspt) { L = d.emax * dt; L = L / d.esum; } else L = 0;
#endif
  d.L = L;
#if ORC_MUTANT == 2
  const T hi = (T)2, lo = (T)0.25f;
#else
  const T hi = (T)6, lo = (T)0.5f;
#endif
#if ORC_MUTANT == 3
  if (d.fsum > 0) {
    d.another = 1; d.reset = 0;
  } else if (L > hi) {
#else
  if (L > hi) {
#endif
    dt = dt * (T)0.5f;
    if (dt < dt_min) dt = dt_min;
    d.another = 0; d.reset = 1;
#if ORC_MUTANT == 6
    d.another = 1; d.rerun = 1;
#endif
#if ORC_MUTANT != 3
  } else if (d.fsum > 0) {
    d.another = 1; d.reset = 0;
#endif
  } else if (L < lo) {
    dt = dt * (T)2;
    if (dt > dt_max) dt = dt_max;
    d.another = 0; d.reset = 1;
  } else {
#if ORC_MUTANT == 1
    d.another = 0; d.reset = 1;
#else
    d.another = 1; d.reset = 0;
#endif
  }
  return d;
}

// ------------------------------------------------ learned warm start ----
// The predictor-pass guess network (no counterpart in the reference): x (769)
// -> GELU(x W1 + b1) (512) -> GELU(. W2 + b2) (512) -> . W3 + b3 (256), with
// every GEMM operand rounded to bfloat16 as the GPU rounds it, float
// accumulation and float bias adds.  The weights are kept transposed (one
// contiguous row per output) for the dot products.
constexpr int WS_IN = 769, WS_H = 512, WS_OUT = NS;
struct WsNet {
  std::vector<uint16_t> w1t, w2t, w3t;   // bfloat16 bit patterns, [out][in]
  std::vector<float> b1, b2, b3;
  float c_out = 0.f;
  bool loaded = false;
};
WsNet g_net;

float bf16_round(float x) {   // nearest even; NaN stays a (quiet) NaN
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    u = (u | 0x00400000u) & 0xffff0000u;
  } else {
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
  }
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

// GELU, tanh form written as x sigmoid(2u), u = sqrt(2/pi) (x + 0.044715 x^3):
// x / (1 + 2^z), z = -2 u log2(e).
float ws_gelu(float x) {
  const float t = x * x;
  const float z = (x * std::fma(0.044715f, t, 1.0f)) * -2.302208198f;
  return x / (1.0f + std::exp2(z));
}

// out[j] = sum_k x[k] w[j][k] + b[j] over K inputs: 16 interleaved float
// partial sums per output (two 8-wide vectors, multiply then add), combined
// pairwise, four outputs per pass over x (a fixed order, so runs are
// reproducible).  The weights are bfloat16 bit patterns (the network's
// precision), widened to float as they are read: half the bytes of float
// weights, which is what bounds this loop.
typedef float ws_v8 __attribute__((vector_size(32)));
typedef uint32_t ws_u8 __attribute__((vector_size(32)));
typedef uint16_t ws_h8 __attribute__((vector_size(16)));
inline ws_v8 ws_ld8(const float *p) {
  ws_v8 v;
  std::memcpy(&v, p, sizeof(v));
  return v;
}
inline ws_v8 ws_ldb8(const uint16_t *p) {
  ws_h8 h;
  std::memcpy(&h, p, sizeof(h));
  const ws_u8 u = __builtin_convertvector(h, ws_u8) << 16;
  ws_v8 v;
  std::memcpy(&v, &u, sizeof(v));
  return v;
}
inline float ws_b2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
void ws_gemv(const float *x, const uint16_t *wt, int K, int N, const float *b, float *out) {
  const int K16 = K & ~15;
  for (int j = 0; j < N; j += 4) {
    const uint16_t *w0 = wt + (size_t)j * K, *w1 = w0 + K, *w2 = w1 + K, *w3 = w2 + K;
    ws_v8 a0l = {}, a0h = {}, a1l = {}, a1h = {}, a2l = {}, a2h = {}, a3l = {}, a3h = {};
    for (int k = 0; k < K16; k += 16) {
      const ws_v8 xl = ws_ld8(x + k), xh = ws_ld8(x + k + 8);
      a0l += xl * ws_ldb8(w0 + k); a0h += xh * ws_ldb8(w0 + k + 8);
      a1l += xl * ws_ldb8(w1 + k); a1h += xh * ws_ldb8(w1 + k + 8);
      a2l += xl * ws_ldb8(w2 + k); a2h += xh * ws_ldb8(w2 + k + 8);
      a3l += xl * ws_ldb8(w3 + k); a3h += xh * ws_ldb8(w3 + k + 8);
    }
    const ws_v8 *acc[4][2] = {{&a0l, &a0h}, {&a1l, &a1h}, {&a2l, &a2h}, {&a3l, &a3h}};
    const uint16_t *wr[4] = {w0, w1, w2, w3};
    for (int r = 0; r < 4; ++r) {
      float a[16];
      std::memcpy(a, acc[r][0], 32);
      std::memcpy(a + 8, acc[r][1], 32);
      for (int k = K16; k < K; ++k) a[k & 15] = a[k & 15] + x[k] * ws_b2f(wr[r][k]);
      for (int h = 8; h >= 1; h >>= 1)
        for (int u = 0; u < h; ++u) a[u] = a[u] + a[u + h];
      out[j + r] = a[0] + b[j + r];
    }
  }
}

void ws_forward(const float *x, float *y) {
  float h1[WS_H], h2[WS_H];
  ws_gemv(x, g_net.w1t.data(), WS_IN, WS_H, g_net.b1.data(), h1);
  for (int j = 0; j < WS_H; ++j) h1[j] = bf16_round(ws_gelu(h1[j]));
  ws_gemv(h1, g_net.w2t.data(), WS_H, WS_H, g_net.b2.data(), h2);
  for (int j = 0; j < WS_H; ++j) h2[j] = bf16_round(ws_gelu(h2[j]));
  ws_gemv(h2, g_net.w3t.data(), WS_H, WS_OUT, g_net.b3.data(), y);
}

// ------------------------------------------------------ kernel #17 -------
// Optional per-pass record of one item's run (orc_bm_trace_f32).
struct Trace {
  orc_pass_record *rec;
  int cap, n;
};

template <typename T> struct Outs {
  T *speeds;
  int64_t rows;
  int32_t *failed_nodes;
  float *conv_time, *conv_loops, *conv_jacobi;
  int32_t *failed_blocks;
  float *conv_time_blocks, *conv_loops_blocks, *conv_jacobi_blocks;
  int64_t *item_steps, *item_passes, *item_sweeps;
};

// One time block integrated by kernel #17 (include/cochlea.cuh:13573-14245).
// C is the type of the sample clock (local time t, time_step_out, Ts and the
// interpolation weight): the method keeps it in float; the FP64 rebuild takes
// it as double or, to isolate the clock's rounding (Q9), as float.
template <typename T, typename C = T>
void bm_item(const orc_bm_scalars &sc, const float *sec, const float *gamma_set, const orc_bm_item &it,
             int item_index, const float *input, int64_t input_len, const Outs<T> &o, Trace *tr = nullptr) {
  // per-section parameters (:13700-13736); promoted from the packed FP32 values
  T rsm[NS], rmass[NS], gam[NS], R[NS], S[NS], Q[NS], Sohc[NS], Stm[NS], Rtm[NS], low[NS], mf[NS];
  for (int i = 0; i < NS; ++i) {
    T mass = (T)sec[0 * NS + i];
    rsm[i] = (T)sec[1 * NS + i];
    R[i] = (T)sec[2 * NS + i];
    S[i] = (T)sec[3 * NS + i];
    Q[i] = (T)sec[4 * NS + i];
    Sohc[i] = (T)sec[5 * NS + i];
    Stm[i] = (T)sec[6 * NS + i];
    Rtm[i] = (T)sec[7 * NS + i];
    mf[i] = (T)sec[8 * NS + i];
    gam[i] = (T)gamma_set[i];
    rmass[i] = (T)1 / mass;                 // reciprocal_mass_ind = 1.0f / mass[i]
    low[i] = i < NS - 2 ? (T)1 : (T)0;      // low_multiplier, SECTIONSM2
  }
  const T w_ohc = sc.w_ohc, dx = sc.delta_x, alpha_r = sc.alpha_r;
  const T al = sc.alpha_l, as = sc.alpha_s, Gme = sc.Gme, a0 = sc.a0, a1 = sc.a1, a2 = sc.a2;
  const T sigma_ow = sc.sigma_ow, eta1 = sc.eta1, eta2 = sc.eta2;
  const T dt_min = sc.dt_min, dt_max = sc.dt_max;
  const C tso = (C)sc.time_step_out, Ts_c = (C)sc.Ts;
  const T mtt = it.tol, mmspt = it.m1tol;
  const T half = (T)0.5f, one = (T)1;
  const T al_over_as = al / as;             // __fdividef(_ohc_alpha_l, _ohc_alpha_s)
  const T dxdx = dx * dx;
  const T dxa0 = dx * a0;

  std::vector<T> st(28 * NS, (T)0);
  T *pd = &st[0 * NS], *pv = &st[1 * NS], *pa = &st[2 * NS], *ppsi = &st[3 * NS], *ppsid = &st[4 * NS];
  T *pxis = &st[5 * NS], *pxi = &st[6 * NS];
  T *nd = &st[7 * NS], *nv = &st[8 * NS], *na = &st[9 * NS], *npsi = &st[10 * NS], *npsid = &st[11 * NS];
  T *nxis = &st[12 * NS], *nxi = &st[13 * NS], *Ptm = &st[14 * NS];
  T *p = &st[15 * NS], *pn = &st[16 * NS], *Y = &st[17 * NS], *G = &st[18 * NS], *od = &st[19 * NS];
  T *vref = &st[20 * NS], *aref = &st[21 * NS];
  T *pold = &st[26 * NS], *pext = &st[27 * NS];   // previous step's pressure; extrapolated guess
  T ow_pd = 0, ow_pv = 0, ow_pa = 0, ow_nd = 0, ow_nv = 0, ow_na = 0;

  T dt = (T)sc.time_step;
  T dt_prev = dt;   // dt of the previous step (warm_start 1)
  C t = 0;
  int in_s = it.start;
  auto in_at = [&](int64_t idx) -> T { return (idx >= 0 && idx < input_len) ? (T)input[idx] : (T)0; };
  T sx0 = in_at(in_s), sx1 = in_at(in_s + 1);
  float c_fail = 0.f, c_dt = 0.f, c_steps = 0.f, c_passes = 0.f;   // sx[11], sx[12], sx[13], sx[15]
  int64_t n_steps = 0, n_passes = 0, n_sweeps = 0;
  int32_t fb = 0;
  float ctb = 0.f, clb = 0.f, cjb = 0.f;
  const float span = (float)(it.mesi - it.mosi);

  while (in_s < it.mesi) {
    if (tr && tr->n >= tr->cap) break;       // trace full: stop the run
    int out_loop = 0, reset = 1, another = 1;
    T dt_commit = dt;
#if ORC_MUTANT == 6
    int reruns = 0;
#endif
    while (another) {
      for (int i = 0; i < NS; ++i) { vref[i] = nv[i]; aref[i] = na[i]; }
      if (out_loop == 0 || reset) {        // Euler predictor (:13770-13795)
        out_loop = 0;
        for (int i = 0; i < NS; ++i) {
          nxi[i] = pxis[i] * dt + pxi[i];
          nv[i] = pa[i] * dt + pv[i];
          nd[i] = ((pv[i] + nv[i]) * half) * dt + pd[i];
          npsi[i] = ppsid[i] * dt + ppsi[i];
        }
        ow_nd = ow_pv * dt + ow_pd;
        ow_nv = ow_pa * dt + ow_pv;
      } else {                             // trapezoidal corrector (:13797-13820)
        for (int i = 0; i < NS; ++i) {
          nv[i] = ((pa[i] + na[i]) * half) * dt + pv[i];
          nd[i] = ((pv[i] + nv[i]) * half) * dt + pd[i];
          nxi[i] = ((pxis[i] + nxis[i]) * half) * dt + pxi[i];
          npsi[i] = ((ppsid[i] + npsid[i]) * half) * dt + ppsi[i];
        }
        ow_nd = ((ow_pv + ow_nv) * half) * dt + ow_pd;
        ow_nv = ((ow_pa + ow_na) * half) * dt + ow_pv;
      }
      for (int i = 0; i < NS; ++i) {
        od[i] = nxi[i] - nd[i];
        T ohc_p = (gam[i] * Sohc[i]) * od[i];
        if (out_loop < 2) {                 // TM/OHC block, CUDA_AUX_TM_LOOPS (:13839-13898)
          T aux = M<T>::tanh_(al_over_as * npsi[i]);
          T dL = (-as) * aux;
          aux = aux * aux - one;
          aux = al * aux;
          T n = (eta2 * nv[i] + eta1 * od[i]) - w_ohc * npsi[i];
          n = Stm[i] * dL + (n * Rtm[i]) * aux;
          T ptmp = Stm[i] * nxi[i] + ohc_p;
          n = ptmp - gam[i] * n;
          T den = (gam[i] * aux) * eta2 - one;
          den = Rtm[i] * den;
          nxis[i] = n / den;
          Ptm[i] = Rtm[i] * nxis[i] + Stm[i] * nxi[i];
        }
        T g = nv[i] * nv[i];                // Calc_G (:13905-13909)
        g = R[i] * (alpha_r * g + one);
        g = g * nv[i] + S[i] * nd[i];
        G[i] = -g - Ptm[i];
      }
      // right-hand side and boundaries (:13925-13954)
      T delta = (T)(t / Ts_c);
      T s = sx1 * delta + sx0 * (one - delta);
      for (int i = 1; i < NS - 1; ++i) Y[i] = (dxdx * G[i]) * Q[i];
      Y[0] = dxa0 * ((((-a2) * ow_nv) - a1 * ow_nd) + Gme * s);
      Y[NS - 1] = 0;
      // Jacobi sweeps, warm start from the resident pressure (:13962-13982)
      auto sweep = [&](int count) {
        for (int it2 = 0; it2 < count; ++it2) {
          for (int i = 0; i < NS; ++i) {
            T pm1 = i > 0 ? p[i - 1] : (T)0;
            T pp1 = i < NS - 1 ? p[i + 1] : (T)0;
            T tmp = low[i] * pm1 + pp1;
            tmp = Y[i] - tmp;
            pn[i] = tmp * rsm[i];
          }
          std::memcpy(p, pn, sizeof(T) * NS);
        }
      };
      int J;
      if (sc.jacobi_mode == 2) {             // O2: exact FP64 pressure of this pass
        thomas_exact(Y, mf, low, p);
        J = 0;
      } else if (sc.jacobi_mode == 1) {      // sweep until |Y - A p| <= tol |Y|
        T y2[NS], r2[NS];
        for (int i = 0; i < NS; ++i) y2[i] = Y[i] * Y[i];
        const T yy = tree_sum(y2);
        const T tau2 = (T)sc.tol_rel * (T)sc.tol_rel;
        auto resid2 = [&](const T *q) {
          for (int i = 0; i < NS; ++i) {
            T pm1 = i > 0 ? q[i - 1] : (T)0;
            T pp1 = i < NS - 1 ? q[i + 1] : (T)0;
            T off = low[i] * pm1 + pp1;
            T r = Y[i] - (off + mf[i] * q[i]);
            r2[i] = r * r;
          }
          return tree_sum(r2);
        };
        if (sc.warm_start == 1 && out_loop == 0) {   // predictor-pass guess
          const T rho = dt / dt_prev;
          for (int i = 0; i < NS; ++i) pext[i] = p[i] + rho * (p[i] - pold[i]);
          const T rr0 = resid2(p), rre = resid2(pext);
          std::memcpy(pold, p, sizeof(T) * NS);
          if (rre < rr0) std::memcpy(p, pext, sizeof(T) * NS);
          dt_prev = dt;
        } else if (sc.warm_start == 2 && out_loop == 0) {
          // learned guess p_lin + rms(r_lin) c_out S(y), S = [1,2,1]/4 with the
          // edges replicated, y = the network of the features (d, v, r_lin,
          // each over its RMS, and s over the RMS of the previous pressure);
          // the smallest residual of the previous, extrapolated and learned
          // guesses is kept (ties: extrapolated over learned, previous over
          // extrapolated)
          const T rho = dt / dt_prev;
          for (int i = 0; i < NS; ++i) pext[i] = p[i] + rho * (p[i] - pold[i]);
          const T rr0 = resid2(p), rre = resid2(pext);
          std::memcpy(pold, p, sizeof(T) * NS);
          dt_prev = dt;
          auto rms = [&](const T *v) {
            T q[NS];
            for (int i = 0; i < NS; ++i) q[i] = v[i] * v[i];
            return (T)std::sqrt(tree_sum(q) * ((T)1 / (T)256)) + (T)1e-30f;
          };
          const T rr_scale = (T)std::sqrt(rre * ((T)1 / (T)256)) + (T)1e-30f;
          const T inv_d = (T)1 / rms(nd), inv_v = (T)1 / rms(nv), inv_r = (T)1 / rr_scale, inv_p = (T)1 / rms(pold);
          float x[WS_IN], y[WS_OUT];
          for (int i = 0; i < NS; ++i) {
            T pm1 = i > 0 ? pext[i - 1] : (T)0;
            T pp1 = i < NS - 1 ? pext[i + 1] : (T)0;
            T r = Y[i] - ((low[i] * pm1 + pp1) + mf[i] * pext[i]);
            x[i] = bf16_round((float)(nd[i] * inv_d));
            x[NS + i] = bf16_round((float)(nv[i] * inv_v));
            x[2 * NS + i] = bf16_round((float)(r * inv_r));
          }
          x[3 * NS] = bf16_round((float)(s * inv_p));
          ws_forward(x, y);
          const T sc_out = rr_scale * (T)g_net.c_out;
          T *ph = pn;
          for (int i = 0; i < NS; ++i) {
            const T yl = (T)y[i > 0 ? i - 1 : 0], yc = (T)y[i], yr = (T)y[i < NS - 1 ? i + 1 : NS - 1];
            ph[i] = pext[i] + sc_out * (((T)0.25f * yl + (T)0.5f * yc) + (T)0.25f * yr);
          }
          const T rrh = resid2(ph);
          if (rrh < rre && rrh < rr0)
            std::memcpy(p, ph, sizeof(T) * NS);
          else if (rre < rr0)
            std::memcpy(p, pext, sizeof(T) * NS);
        }
        J = 0;
        for (;;) {
          const T rr = resid2(p);
          if (rr <= tau2 * yy || J >= sc.max_sweeps) break;
          const int k = std::min(sc.check_every, sc.max_sweeps - J);
          sweep(k);
          J += k;
        }
      } else {
        J = out_loop < 2 ? sc.j_slow : sc.j_fast;
        sweep(J);
      }
      n_sweeps += J;
      for (int i = 0; i < NS; ++i) na[i] = (p[i] + G[i]) * rmass[i];   // (:13989-13991)
      {                                      // OW acceleration, thread 0 (:13995-14006)
        T NN1 = (-a2) * ow_nv;
        T NN2 = Gme * s + p[0];
        T NN3 = (-a1) * ow_nd;
        ow_na = ((NN1 + NN2) + NN3) / sigma_ow;
      }
      for (int i = 0; i < NS; ++i)          // psi derivative (:14011)
        npsid[i] = (eta1 * od[i] + eta2 * (nxis[i] - nv[i])) + (-w_ohc) * npsi[i];

      const T dt_in = dt;
      const int decided = reset == 0;
      Decision<T> d{0, 0, 0, 0, 1, 0, 0};
      if (decided) {                         // step control (:14021-14088)
        d = step_control<T>(aref, na, vref, nv, dt, mtt, mmspt, dt_min, dt_max, sc.decision_mode);
        another = d.another; reset = d.reset;
      } else {                               // pass 0 always iterates (:14095-14104)
        another = 1; reset = 0;
      }
      const int pass = out_loop;
      out_loop++;
      int capped = 0;
      if (out_loop >= sc.max_passes) { another = 0; reset = 1; c_fail = 1.f; capped = 1; }
#if ORC_MUTANT == 6
      if (d.rerun && (capped || ++reruns >= 64)) another = 0;
#endif
      if (tr && tr->n < tr->cap) {
        orc_pass_record &r = tr->rec[tr->n++];
        r.step = (int32_t)n_steps; r.pass = pass; r.in_s = in_s; r.decided = decided;
        r.t = (float)t; r.dt_in = (float)dt_in; r.dt_out = (float)dt;
        r.L = (float)d.L; r.F = (float)d.fsum; r.e_a = (float)d.emax; r.e_v = (float)d.esum;
        r.another = another; r.reset = reset; r.capped = capped; r.sweeps = J;
      }
#if ORC_MUTANT == 6
      if (another) continue;   // restart the step from the predictor with the halved dt
#endif
#if ORC_MUTANT == 7
      dt_commit = dt_in;
#else
      dt_commit = dt;
#endif
    }
    // committed step counters (:14130-14141)
    c_steps += 1.0f;
    c_dt += (float)dt;
    c_passes += (float)out_loop;
    n_steps++;
    n_passes += out_loop;
    if (t > tso) {                           // sample advance (:14145-14203)
      t -= tso;
      sx0 = sx1;
      sx1 = in_at((int64_t)in_s + 2);
      if (in_s + 1 >= it.mosi) {
        float div_times = c_dt / c_steps;
        float div_j = c_passes / c_steps;
        if (in_s >= it.write_start && in_s < o.rows) {
          if (o.speeds) for (int i = 0; i < NS; ++i) o.speeds[(int64_t)in_s * NS + i] = nv[i];
          if (o.failed_nodes) o.failed_nodes[in_s] = (int32_t)c_fail;
          if (o.conv_time) o.conv_time[in_s] = div_times;
          if (o.conv_jacobi) o.conv_jacobi[in_s] = div_j;
          if (o.conv_loops) o.conv_loops[in_s] = c_steps;
        }
        if (c_fail > 0) fb += 1;
        ctb += div_times / span;
        cjb += div_j / span;
        clb += c_steps / span;
      }
      c_fail = 0.f; c_dt = 0.f; c_steps = 0.f; c_passes = 0.f;
      in_s++;
    }
    for (int i = 0; i < NS; ++i) {           // commit (:14212-14226)
      pd[i] = nd[i]; pv[i] = nv[i]; pa[i] = na[i]; ppsi[i] = npsi[i];
      ppsid[i] = npsid[i]; pxis[i] = nxis[i]; pxi[i] = nxi[i];
    }
    ow_pd = ow_nd; ow_pv = ow_nv; ow_pa = ow_na;
    t += (C)dt_commit;   // the new dt, also after a halving pass (Q2)
  }
  if (o.failed_blocks) o.failed_blocks[item_index] = fb;
  if (o.conv_time_blocks) o.conv_time_blocks[item_index] = ctb;
  if (o.conv_loops_blocks) o.conv_loops_blocks[item_index] = clb;
  if (o.conv_jacobi_blocks) o.conv_jacobi_blocks[item_index] = cjb;
  if (o.item_steps) o.item_steps[item_index] = n_steps;
  if (o.item_passes) o.item_passes[item_index] = n_passes;
  if (o.item_sweeps) o.item_sweeps[item_index] = n_sweeps;
}

template <typename T, typename C = T>
int bm_run(const orc_bm_scalars *sc, const float *sec, const float *gamma, const orc_bm_item *items, int n_items,
           const float *input, int64_t input_len, const Outs<T> &o, int n_threads) {
  if (!sc || !sec || !gamma || !items || !input || n_items < 0) return -1;
  if (sc->warm_start == 2 && (!g_net.loaded || !std::is_same<T, float>::value || sc->jacobi_mode != 1)) return -2;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int b = 0; b < n_items; ++b)
    bm_item<T, C>(*sc, sec, gamma + (int64_t)items[b].set * NS, items[b], b, input, input_len, o);
  return 0;
}

}  // namespace

extern "C" {

float orc_tanhf(float x) { return tanh_defined(x); }
float orc_expf(float x) { return exp_reduced(x); }

// CModel::CModel, src/model.cpp:72-86, with the float-suffixed macros of
// include/const.h:216-255 evaluated in float before promotion (Q24).
int orc_model_vectors(double *out) {
  const float PIf = 3.141592653589793f;
  const double dx = (double)(3.5f / 256);
  const double M0 = 1.286e-6f, M1 = 1.5f, R0 = 0.25f, R1 = -0.06f, S0 = 1.282e4f, S1 = -1.5f;
  const double w_ohc = (double)((2 * PIf) * 1000.0f);
  const double ratio_K = 1e-3f, rho = 1.0f, beta = 0.003f, area = 0.5f;
  for (int i = 0; i < NS; ++i) {
    double Mb = M0 * std::exp(M1 * dx * i);
    double Rb = R0 * std::exp(R1 * dx * i);
    double Sb = S0 * std::exp(S1 * dx * i);
    double wcf2 = Sb / Mb;
    double Stm = Rb * wcf2 / w_ohc;
    out[0 * NS + i] = Mb;
    out[1 * NS + i] = Rb;
    out[2 * NS + i] = Sb;
    out[3 * NS + i] = Rb;
    out[4 * NS + i] = Stm;
    out[5 * NS + i] = ratio_K * Stm;
    out[6 * NS + i] = 2 * rho * beta / (area * Mb);
  }
  return 0;
}

int orc_tri_mid_diag(double *m) {
  double v[7 * NS];
  orc_model_vectors(v);
  const double dx = (double)(3.5f / 256);
  const double a0 = (double)((2.0f * 1.0f * 6e-3f) / 0.5f);
  for (int i = 0; i < NS; ++i) m[i] = -1.0 * (2.0 + dx * dx * v[6 * NS + i]);
  m[0] = -1.0 * (1.0 + dx * a0);
  m[NS - 1] = 1.0;
  return 0;
}

int orc_default_scalars(float Fs, orc_bm_scalars *o) {
  const float PIf = 3.141592653589793f;
  float W_OHC = 2 * PIf * 1000.0f;
  o->w_ohc = W_OHC;
  o->time_step = 1e-6f;
  int iters = (int)(1.0f / (Fs * 1e-6f));
  o->time_step_out = 1e-6f * (float)iters;
  o->delta_x = 3.5f / 256;
  o->alpha_r = 0.0f;
  o->Ts = (float)(1.0 / (double)Fs);
  o->alpha_l = 2e-6f;
  o->alpha_s = 1e-6f;
  o->Gme = 21.4f;
  o->a0 = (2.0f * 1.0f * 6e-3f) / 0.5f;
  float wow2 = 2 * PIf * 1500 * 2 * PIf * 1500;
  float cme = (2 * PIf * 1340) * (2 * PIf * 1340) * 0.059f / (0.49f * 1.4f);
  o->a1 = 0.5f * wow2 + cme * 21.4f;
  o->a2 = 0.5f * 20e3f;
  o->sigma_ow = 0.5f;
  o->eta1 = (1.0f) / 2e-6f * 2 * PIf * 1000.0f;
  o->eta2 = 0.0f;
  o->dt_min = 1e-18f;
  o->dt_max = 1e-6f;
  o->j_fast = 2;
  o->j_slow = 20;
  o->max_passes = 10;
  o->decision_mode = 0;
  o->jacobi_mode = 0;
  o->check_every = 4;
  o->max_sweeps = 1000;
  o->tol_rel = 1e-4f;
  o->warm_start = 0;
  return 0;
}

// src/cochlea.cu:1148-1176.  write_start: rows [write_start, mesi) are this
// block's in per-row outputs; the row mosi-1 that the next block would also
// write (Q10 overlap) stays with the earlier block.
int orc_block_decomposition(int TB, int bufP1, int overlap, int DF, int32_t *start, int32_t *mosi,
                            int32_t *mesi, int32_t *write_start) {
  if (TB <= 0) return -1;
  int I8 = (bufP1 - overlap) / (TB + 1);
  for (int b = 0; b < TB; ++b) {
    int transient_offset = (b > 0 && (DF == 0 || (DF != 1 && b % DF != 0)));
    int preDecoupled = DF > 0 && (DF == 1 || ((b + 1) % DF == 0)) && b != TB - 1;
    int postDecoupled = DF > 0 && (b % DF == 0);
    int in0 = b * I8;
    int so = in0 + std::max(1 - transient_offset - postDecoupled, 0) * overlap;
    int bl = I8 + (1 - preDecoupled) * overlap;
    int eo = in0 + bl;
    if (b > 0) so = std::max(so, mesi[b - 1]);
    start[b] = in0;
    mosi[b] = so;
    mesi[b] = eo;
    int first = std::max(in0, so - 1);
    if (b > 0 && first < mesi[b - 1]) first = mesi[b - 1];
    write_start[b] = first;
  }
  return 0;
}

int orc_bm_run_f32(const orc_bm_scalars *sc, const float *sec, const float *gamma, const orc_bm_item *items,
                   int n_items, const float *input, int64_t input_len, float *speeds, int64_t rows,
                   int32_t *failed_nodes, float *conv_time, float *conv_loops, float *conv_jacobi,
                   int32_t *failed_blocks, float *conv_time_blocks, float *conv_loops_blocks,
                   float *conv_jacobi_blocks, int64_t *item_steps, int64_t *item_passes, int64_t *item_sweeps,
                   int n_threads) {
  Outs<float> o{speeds, rows, failed_nodes, conv_time, conv_loops, conv_jacobi, failed_blocks,
                conv_time_blocks, conv_loops_blocks, conv_jacobi_blocks, item_steps, item_passes, item_sweeps};
  return bm_run<float>(sc, sec, gamma, items, n_items, input, input_len, o, n_threads);
}

int orc_bm_run_f64(const orc_bm_scalars *sc, const float *sec, const float *gamma, const orc_bm_item *items,
                   int n_items, const float *input, int64_t input_len, double *speeds, int64_t rows,
                   int64_t *item_steps, int64_t *item_passes, int64_t *item_sweeps, int n_threads) {
  Outs<double> o{speeds, rows, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                 item_steps, item_passes, item_sweeps};
  return bm_run<double>(sc, sec, gamma, items, n_items, input, input_len, o, n_threads);
}

int orc_step_control_f32(const float *aref, const float *na, const float *vref, const float *nv, float dt, float tol,
                         float m1tol, float dt_min, float dt_max, int decision_mode, float *out5,
                         int32_t *another_reset) {
  if (!aref || !na || !vref || !nv || !out5 || !another_reset) return -1;
  const Decision<float> d = step_control<float>(aref, na, vref, nv, dt, tol, m1tol, dt_min, dt_max, decision_mode);
  out5[0] = dt;
  out5[1] = d.L;
  out5[2] = d.emax;
  out5[3] = d.esum;
  out5[4] = d.fsum;
  another_reset[0] = d.another;
  another_reset[1] = d.reset;
  return 0;
}

int orc_bm_trace_f32(const orc_bm_scalars *sc, const float *sec, const float *gamma, const orc_bm_item *item,
                     const float *input, int64_t input_len, orc_pass_record *rec, int cap, int *n_rec) {
  if (!sc || !sec || !gamma || !item || !input || !rec || cap <= 0 || !n_rec) return -1;
  Outs<float> o{nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr, nullptr, nullptr};
  Trace tr{rec, cap, 0};
  bm_item<float>(*sc, sec, gamma + (int64_t)item->set * NS, *item, 0, input, input_len, o, &tr);
  *n_rec = tr.n;
  return 0;
}

int orc_bm_run_f64_clock32(const orc_bm_scalars *sc, const float *sec, const float *gamma, const orc_bm_item *items,
                           int n_items, const float *input, int64_t input_len, double *speeds, int64_t rows,
                           int64_t *item_steps, int64_t *item_passes, int64_t *item_sweeps, int n_threads) {
  Outs<double> o{speeds, rows, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                 item_steps, item_passes, item_sweeps};
  return bm_run<double, float>(sc, sec, gamma, items, n_items, input, input_len, o, n_threads);
}

int orc_ws_forward(const float *x, int n_rows, float *y) {
  if (!g_net.loaded || !x || !y || n_rows < 0) return -1;
  for (int r = 0; r < n_rows; ++r) {
    float xb[WS_IN];
    for (int k = 0; k < WS_IN; ++k) xb[k] = bf16_round(x[(size_t)r * WS_IN + k]);
    ws_forward(xb, y + (size_t)r * WS_OUT);
  }
  return 0;
}

int orc_ws_set_net(const float *w1, const float *b1, const float *w2, const float *b2, const float *w3,
                   const float *b3, float c_out) {
  if (!w1 || !b1 || !w2 || !b2 || !w3 || !b3) return -1;
  auto tr = [](const float *w, int K, int N, std::vector<uint16_t> &out) {   // [K][N] -> rounded [N][K]
    out.assign((size_t)K * N, 0);
    for (int k = 0; k < K; ++k)
      for (int j = 0; j < N; ++j) {
        const float r = bf16_round(w[(size_t)k * N + j]);
        uint32_t u;
        std::memcpy(&u, &r, 4);
        out[(size_t)j * K + k] = (uint16_t)(u >> 16);
      }
  };
  tr(w1, WS_IN, WS_H, g_net.w1t);
  tr(w2, WS_H, WS_H, g_net.w2t);
  tr(w3, WS_H, WS_OUT, g_net.w3t);
  g_net.b1.assign(b1, b1 + WS_H);
  g_net.b2.assign(b2, b2 + WS_H);
  g_net.b3.assign(b3, b3 + WS_OUT);
  g_net.c_out = c_out;
  g_net.loaded = true;
  return 0;
}

int orc_jacobi(const float *Y, const float *rsm, float *p, int sweeps) {
  float pn[NS];
  for (int s = 0; s < sweeps; ++s) {
    for (int i = 0; i < NS; ++i) {
      float low = i < NS - 2 ? 1.0f : 0.0f;
      float pm1 = i > 0 ? p[i - 1] : 0.0f;
      float pp1 = i < NS - 1 ? p[i + 1] : 0.0f;
      float tmp = low * pm1 + pp1;
      tmp = Y[i] - tmp;
      pn[i] = tmp * rsm[i];
    }
    std::memcpy(p, pn, sizeof(pn));
  }
  return 0;
}

// ------------------------------------------------------------ IHC ------
// FIR window of DeviceCudaFIRFilter (src/cochlea.cu:2712-2728): taps used at
// node k are min((k+1) mod period, taps), period = interval length before the
// final regular division and the whole analysed length after it (Q17).  The
// reference's Y = Y + c*x accumulation is taken as a fused multiply-add in
// tap order (what nvcc's default contraction makes of it).
static inline int fir_window(int64_t k, int64_t tla, int64_t L, int64_t frdp, int taps) {
  int64_t sg = frdp > 1 ? k / (frdp - 1) : 1;
  int64_t per = sg * tla + (1 - sg) * L;
  if (per <= 0) return 0;
  int64_t w = (k + 1) % per;
  return (int)std::min<int64_t>(w, taps);
}

// IIR AC filter of DeviceCudaIIRFilter (src/cochlea.cu:2742-2772) run by
// CudaIIRFilter (one interval over all rows, :2840-2857) or
// CudaIIRFilterDecoupled (intervals of L rows, last one extended, :2885-2908).
// Accumulation is in float with double products, as the <float,float> instance.
static void iir_section(const float *bm, int64_t rows, int s, int64_t L, int64_t frdp, const double *b,
                        const double *a, int taps, float *Y) {
  const int64_t nint = L > 0 ? std::max<int64_t>(1, rows / L) : 1;
  for (int64_t iv = 0; iv < nint; ++iv) {
    const int64_t k0 = L > 0 ? iv * L : 0;
    int64_t len = L > 0 ? L : rows;
    if (iv + 1 == nint) len += rows - nint * len;
    for (int64_t k = 0; k < len; ++k) {
      int64_t sg = frdp > 1 ? k / (frdp - 1) : 1;
      int64_t per = sg * len + (1 - sg) * len;
      int ob = per > 0 ? (int)std::min<int64_t>((k + 1) % per, taps) : 0;
      float y = 0.0f;
      for (int i = 0; i < ob; ++i) y = (float)((double)y + b[i] * (double)bm[(k0 + k - i) * NS + s]);
      for (int i = 0; i < ob - 1; ++i) y = (float)((double)y - a[i + 1] * (double)Y[k0 + k - i - 1]);
      Y[k0 + k] = y;
    }
  }
}

int orc_ihc_lambda(const float *bm, int64_t rows, int interval_len, int frdp, const double *ac_b,
                   const double *ac_a, int n_ac, const double *dc_b, int n_dc, const double *ihc_damage,
                   const float *A, const float *spont, float lambda_sat, float eta_ac, float eta_dc, float scale_bm,
                   float *lambda, float *stages, int n_threads) {
  if (!bm || !lambda || rows <= 0) return -1;
  const int64_t tla = rows;
  const int64_t L = interval_len > 0 ? interval_len : rows;
  const float coefA = scale_bm;           // vectors_sum_linear_coefficents[0]
  const double coefB = -1.0;
  const float coefAC = eta_ac;            // vectors_sum_linear_coefficents[1]
  const double coefDC = (double)eta_dc;
  const double SAT = (double)lambda_sat;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(static)
#endif
  for (int s = 0; s < NS; ++s) {
    std::vector<float> AC(rows), dS(rows), DC(rows);
    if (ac_a) {                            // AC IIR, Eq. 5.1
      iir_section(bm, rows, s, interval_len > 0 ? interval_len : 0, frdp, ac_b, ac_a, n_ac, AC.data());
    } else {
      for (int64_t k = 0; k < rows; ++k) { // AC FIR, Eq. 5.1 (double accumulate)
        int w = fir_window(k, tla, L, frdp, n_ac);
        double y = 0.0;
        for (int i = 0; i < w; ++i) y = std::fma(ac_b[i], (double)bm[(k - i) * NS + s], y);
        AC[k] = (float)y;
      }
    }
    for (int64_t k = 0; k < rows; ++k) {   // dS = (A*BM - AC)^2, Eq. 5.2
      float mid = fmaf(bm[k * NS + s], coefA, (float)((double)AC[k] * coefB));
      dS[k] = mid * mid;
    }
    for (int64_t k = 0; k < rows; ++k) {   // DC box-car FIR, Eq. 5.3
      int w = fir_window(k, tla, L, frdp, n_dc);
      double y = 0.0;
      for (int i = 0; i < w; ++i) y = std::fma(dc_b[i], (double)dS[k - i], y);
      DC[k] = (float)y;
    }
    const double dmg = ihc_damage[s];
    const double zero_factor = s != 0 ? 1.0 : 0.0;
    for (int64_t k = 0; k < rows; ++k) {   // Eq. 5.4-5.8
      float pre = fmaf(AC[k], coefAC, (float)((double)DC[k] * coefDC));
      float ihc = (float)std::log10(std::fmax((double)(dmg * (double)pre), 0.0) + 1e-20);
      if (stages) {
        stages[(0 * rows + k) * NS + s] = AC[k];
        stages[(1 * rows + k) * NS + s] = dS[k];
        stages[(2 * rows + k) * NS + s] = DC[k];
        stages[(3 * rows + k) * NS + s] = pre;
        stages[(4 * rows + k) * NS + s] = ihc;
      }
      for (int li = 0; li < ORC_LAMBDA_COUNT; ++li) {
        double Aihc = (double)A[li * NS + s];
        double sp = (double)spont[li];
        double zero_offset = s != 0 ? 0.0 : sp;
        double base = std::fmin(SAT, std::fmax(Aihc * (double)ihc, sp));
        lambda[((int64_t)li * rows + k) * NS + s] = (float)std::fma(zero_factor, base, zero_offset);
      }
    }
  }
  return 0;
}

// ------------------------------------------------------------ JND ------
// Shared-memory halving tree of GlobalCalculateJND (src/cochlea.cu:3924-3933).
static float halving_tree(float *a) {
  for (int t = NS >> 1; t >= 1; t >>= 1)
    for (int s = 0; s < t; ++s) a[s] = a[s] + a[s + t];
  return a[0];
}

int orc_jnd(const float *lambda, int64_t rows, int n_intervals, int interval_nodes, int head_nodes,
            int length_nodes, const float *dA, const int32_t *ref_index, const double *nIHC, const float *W3,
            double Fs, double delta_alpha_factor, int calc_ai, int calc_rms, float *mean_rate, float *F_RA,
            float *FisherAISum, int backup_stage, float *buffer) {
  const int stage = buffer ? backup_stage : 0;
  if (!lambda || n_intervals <= 0) return -1;
  const int len = length_nodes;
  std::vector<float> MR((size_t)ORC_LAMBDA_COUNT * n_intervals * NS);
  for (int li = 0; li < ORC_LAMBDA_COUNT; ++li)     // GlobalCalculateMeanRate (:3741-3771)
    for (int iv = 0; iv < n_intervals; ++iv)
      for (int s = 0; s < NS; ++s) {
        int64_t off = ((int64_t)li * rows + (int64_t)iv * interval_nodes + head_nodes) * NS + s;
        double acc = lambda[off];
        for (int t = 1; t < len; ++t) acc = acc + lambda[off + (int64_t)t * NS];
        MR[((size_t)li * n_intervals + iv) * NS + s] = (float)acc / (float)len;
      }
  if (mean_rate) std::memcpy(mean_rate, MR.data(), MR.size() * sizeof(float));
  const float T = (float)(delta_alpha_factor / Fs);
  const float Tlength = (float)((double)(float)len / Fs);
  float acc_ai[NS], acc_rms[NS];
  for (int li = 0; li < ORC_LAMBDA_COUNT; ++li)      // GlobalCalculateJND (:3773-3944)
    for (int iv = 0; iv < n_intervals; ++iv) {
      int ref = ref_index[iv];
      float dAv = dA[iv];
      const size_t mi = ((size_t)li * n_intervals + iv) * NS;
      if (stage == 1 && calc_rms) {          // AvgMeanRate (:3815-3831)
        float tmp[NS];
        std::memcpy(tmp, &MR[mi], sizeof(tmp));
        buffer[(size_t)li * n_intervals + iv] = halving_tree(tmp) / (float)NS;
      }
      for (int s = 0; s < NS; ++s) {
        float mr = MR[mi + s];
        float mref = MR[((size_t)li * n_intervals + ref) * NS + s];
        float dMR = (mr - mref) / dAv;
        if (stage == 2) buffer[mi + s] = dMR;   // dMeanRate (:3858-3861)
        if (calc_ai) {
          int64_t c = ((int64_t)li * rows + (int64_t)iv * interval_nodes + head_nodes) * NS + s;
          int64_t r = ((int64_t)li * rows + (int64_t)ref * interval_nodes + head_nodes) * NS + s;
          float accv = 0.0f;
          for (int t = 0; t < len; ++t) {
            float refl = lambda[r + (int64_t)t * NS];
            float cal = lambda[c + (int64_t)t * NS];
            float dl = dAv > 0 ? (cal - refl) / dAv : 0.0f;
            float pre = (float)((double)(dl * dl) / (Fs * (double)refl));
            if (stage == 3) buffer[c + (int64_t)t * NS] = dl;    // dLambda (:3870-3872)
            if (stage == 4) buffer[c + (int64_t)t * NS] = pre;   // preFisherAI (:3883-3885)
            accv += pre;
          }
          float rs = 1.0f / sqrtf(accv);
          acc_ai[s] = (float)((double)(T / Tlength) * nIHC[s] / (double)(rs * rs));
          if (stage == 5) buffer[mi + s] = acc_ai[s];           // preFisherAITimeReduced (:3892-3894)
        }
        if (calc_rms) {
          float crlb = 1.0f / sqrtf(((T / mref) * dMR) * dMR);
          if (stage == 6) buffer[mi + s] = crlb;                // CRLB_RA (:3909-3911)
          acc_rms[s] = (float)(nIHC[s] / (double)(crlb * crlb));
        }
      }
      size_t fi = (size_t)li * n_intervals + iv;
      if (calc_ai && FisherAISum) FisherAISum[fi] = halving_tree(acc_ai) * W3[li];
      if (calc_rms && F_RA) F_RA[fi] = halving_tree(acc_rms) * W3[li];
    }
  return 0;
}

// CudaGenerateInputFromProfile, src/cochlea.cu:4070-4103 (tone mode).
int orc_generate_input(int n_profiles, const double *dA, const double *Wn, const double *freq, const float *WN,
                       double wn_dc, double wn_norm, double sig_norm, int L, int head, float Fs, float *out) {
  const float PIf = 3.141592653589793f;
  for (int p = 0; p < n_profiles; ++p)
    for (int n = 0; n < L; ++n) {
      double time = (double)(n - head) / (double)Fs;
      double cut = (double)(2 * PIf) * freq[p] * time;
      double sum = dA[p] * std::cos(cut) / sig_norm;
      sum = sum + Wn[p] * ((double)WN[n] - wn_dc) / wn_norm;
      out[(int64_t)p * L + n] = (float)sum;
    }
  return 0;
}

}  // extern "C"
