// BM/OHC velocity solver with the semantics of the reference's kernel #17
// (include/cochlea.cuh:13573-14245, launched from src/cochlea.cu:1436-1493).
//
// gfx950 layout: one wave64 integrates one time block; lane l holds the four
// consecutive sections 4l..4l+3 in registers.  The Jacobi pressure sweep
// exchanges the two lane-boundary pressures with DPP wave shifts, the three
// step-control reductions run as DPP butterflies plus four readlanes, and
// accept/reject is a lane-uniform decision.  No LDS, no barriers.
//
// Arithmetic contract (shared with the CPU oracle): FP32, every operation
// IEEE-rounded, no contraction (built with -ffp-contract=off), IEEE division
// where the reference uses __fdividef, tanh by cm_tanhf below, and the fixed
// reduction tree (s0+s1)+(s2+s3) -> xor 1,2,4,8 -> (r0+r16)+(r32+r48).
#include <hip/hip_runtime.h>

#include "cm_internal.h"

namespace cm {

// ----------------------------------------------------------- math ------
__device__ __forceinline__ float exp_reduced(float y) {
  const float log2e = 1.44269502e+00f;
  const float ln2_hi = 6.93145752e-01f;
  const float ln2_lo = 1.42860677e-06f;
  float k = __builtin_rintf(y * log2e);
  float r = __builtin_fmaf(-k, ln2_hi, y);
  r = __builtin_fmaf(-k, ln2_lo, r);
  float p = 2.48015873e-05f;
  p = __builtin_fmaf(p, r, 1.98412701e-04f);
  p = __builtin_fmaf(p, r, 1.38888892e-03f);
  p = __builtin_fmaf(p, r, 8.33333377e-03f);
  p = __builtin_fmaf(p, r, 4.16666679e-02f);
  p = __builtin_fmaf(p, r, 1.66666672e-01f);
  p = __builtin_fmaf(p, r, 5.00000000e-01f);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  return __builtin_ldexpf(p, (int)k);
}

// tanh: odd series below 0.55, 1 - 2/(e^{2a}+1) below 10, 1 beyond.
__device__ __forceinline__ float cm_tanhf(float x) {
  float a = __builtin_fabsf(x);
  float r;
  if (a < 0.55f) {
    float z = a * a;
    float q = -2.39129114e-04f;
    q = __builtin_fmaf(q, z, 5.90027441e-04f);
    q = __builtin_fmaf(q, z, -1.45583438e-03f);
    q = __builtin_fmaf(q, z, 3.59212812e-03f);
    q = __builtin_fmaf(q, z, -8.86323989e-03f);
    q = __builtin_fmaf(q, z, 2.18694881e-02f);
    q = __builtin_fmaf(q, z, -5.39682540e-02f);
    q = __builtin_fmaf(q, z, 1.33333340e-01f);
    q = __builtin_fmaf(q, z, -3.33333343e-01f);
    r = __builtin_fmaf(a * z, q, a);
  } else if (a < 10.0f) {
    float e = exp_reduced(a + a);
    r = 1.0f - 2.0f / (e + 1.0f);
  } else {
    r = 1.0f;
  }
  return __builtin_copysignf(r, x);
}

// ------------------------------------------------------------ DPP ------
#define CM_DPP(x, ctrl) \
  __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), (ctrl), 0xF, 0xF, true))
// lane i <- lane i-1 (lane 0 <- 0)
__device__ __forceinline__ float wave_shr1(float x) { return CM_DPP(x, 0x138); }
// lane i <- lane i+1 (lane 63 <- 0)
__device__ __forceinline__ float wave_shl1(float x) { return CM_DPP(x, 0x130); }
__device__ __forceinline__ float quad_xor1(float x) { return CM_DPP(x, 0xB1); }
__device__ __forceinline__ float quad_xor2(float x) { return CM_DPP(x, 0x4E); }
__device__ __forceinline__ float row_half_mirror(float x) { return CM_DPP(x, 0x141); }
__device__ __forceinline__ float row_mirror(float x) { return CM_DPP(x, 0x140); }
__device__ __forceinline__ float readlane(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// Row broadcasts: every row r >= 1 <- lane 15 of row r-1; rows 2 and 3 <-
// lane 31.  Only row 3 is read afterwards, so what the other rows receive
// does not matter and the full row mask lets the move fold into the add.
__device__ __forceinline__ float row_bcast15(float x) { return CM_DPP(x, 0x142); }
__device__ __forceinline__ float row_bcast31(float x) { return CM_DPP(x, 0x143); }

// Sum of 256 values, 4 per lane, with the fixed tree
// ((r0 + r1) + (r2 + r3)) of the row sums r_k.  After the xor steps every lane
// of row k holds r_k; the two row broadcasts leave (r3 + r2) + (r1 + r0) in
// row 3, which equals the tree because each addition is commutative.
__device__ __forceinline__ float tree_sum(float x0, float x1, float x2, float x3) {
  float s = (x0 + x1) + (x2 + x3);
  s = s + quad_xor1(s);
  s = s + quad_xor2(s);
  s = s + row_half_mirror(s);   // equals xor 4 once quads are uniform
  s = s + row_mirror(s);        // equals xor 8 once octets are uniform
  s = s + row_bcast15(s);
  s = s + row_bcast31(s);
  return readlane(s, 63);
}
__device__ __forceinline__ float tree_max(float x0, float x1, float x2, float x3) {
  float s = fmaxf(fmaxf(x0, x1), fmaxf(x2, x3));
  s = fmaxf(s, quad_xor1(s));
  s = fmaxf(s, quad_xor2(s));
  s = fmaxf(s, row_half_mirror(s));
  s = fmaxf(s, row_mirror(s));
  s = fmaxf(s, row_bcast15(s));
  s = fmaxf(s, row_bcast31(s));
  return readlane(s, 63);
}



#define UNROLL _Pragma("unroll")

__device__ __forceinline__ float in_at(const BmLaunch &a, int64_t idx) {
  return (idx >= 0 && idx < a.input_len) ? a.input[idx] : 0.0f;
}

// J Jacobi sweeps of the kernel convention (include/cochlea.cuh:13962-13982):
// p_i <- (Y_i - (l_i p_{i-1} + p_{i+1})) / m_i with the lane-boundary
// neighbours exchanged by DPP wave shifts (zero past either end).  l_i is 0 or
// 1, so l_i p_{i-1} is exact and fma(l_i, p_{i-1}, p_{i+1}) rounds exactly as
// the separate multiply and add do: one instruction fewer per sweep.
__device__ __forceinline__ void jacobi_sweep(const float (&Y)[4], const float (&rsm)[4], float low23,
                                             const float (&p)[4], float (&q)[4]) {
  const float left = wave_shr1(p[3]);
  const float right = wave_shl1(p[0]);
  q[0] = (Y[0] - (left + p[1])) * rsm[0];
  q[1] = (Y[1] - (p[0] + p[2])) * rsm[1];
  q[2] = (Y[2] - __builtin_fmaf(low23, p[1], p[3])) * rsm[2];
  q[3] = (Y[3] - (low23 * p[2] + right)) * rsm[3];
}
// J is wave-uniform: the count goes to a scalar register.  U (1, 2 or 4) is
// the unroll of the sweep loop, a launch knob of the autotuner: with U >= 2
// sweeps come in pairs so the two iterates alternate between register sets
// instead of being copied back each sweep; U = 4 runs two pairs per loop trip.
// Every U performs the same sweeps in the same order, so results are identical.
template <int U = 2>
__device__ __forceinline__ void jacobi_sweeps(const float (&Y)[4], const float (&rsm)[4], float low23, float (&p)[4],
                                              int J) {
  static_assert(U == 1 || U == 2 || U == 4, "unroll 1, 2 or 4");
  J = __builtin_amdgcn_readfirstlane(J);
  float q[4];
  if (U == 1) {
    for (int k = 0; k < J; ++k) {
      jacobi_sweep(Y, rsm, low23, p, q);
      UNROLL for (int j = 0; j < 4; ++j) p[j] = q[j];
    }
    return;
  }
  int k = 0;
  if (U == 4) {
    for (; k + 4 <= J; k += 4) {
      jacobi_sweep(Y, rsm, low23, p, q);
      jacobi_sweep(Y, rsm, low23, q, p);
      jacobi_sweep(Y, rsm, low23, p, q);
      jacobi_sweep(Y, rsm, low23, q, p);
    }
  }
  for (; k + 2 <= J; k += 2) {
    jacobi_sweep(Y, rsm, low23, p, q);
    jacobi_sweep(Y, rsm, low23, q, p);
  }
  if (J & 1) {
    jacobi_sweep(Y, rsm, low23, p, q);
    UNROLL for (int j = 0; j < 4; ++j) p[j] = q[j];
  }
}


// Residual Y - A p of the kernel convention, per lane, and its squared 2-norm
// with the fixed tree.
__device__ __forceinline__ void resid_vec(const float (&Y)[4], const float (&mf)[4], float low23, const float (&p)[4],
                                          float (&r)[4]) {
  const float left = wave_shr1(p[3]);
  const float right = wave_shl1(p[0]);
  r[0] = Y[0] - ((left + p[1]) + mf[0] * p[0]);
  r[1] = Y[1] - ((p[0] + p[2]) + mf[1] * p[1]);
  r[2] = Y[2] - (__builtin_fmaf(low23, p[1], p[3]) + mf[2] * p[2]);
  r[3] = Y[3] - ((low23 * p[2] + right) + mf[3] * p[3]);
}
__device__ __forceinline__ float resid2(const float (&Y)[4], const float (&mf)[4], float low23, const float (&p)[4]) {
  float r[4];
  resid_vec(Y, mf, low23, p, r);
  return tree_sum(r[0] * r[0], r[1] * r[1], r[2] * r[2], r[3] * r[3]);
}

// ------------------------------------------------- one time block -------
// Per-lane constants of an item: section parameters of sections 4l..4l+3.
struct ItemConst {
  float rsm[4], rmass[4], gam[4], R[4], S[4], Q[4], Sohc[4], Stm[4], Rtm[4], mf[4];
  float low23;   // lower-coupling multiplier of slots 2, 3 (0 for sections 254, 255)
};

// Scalars shared by all lanes (uniform).
struct ItemScal {
  float w_ohc, alpha_r, Ts, al, Gme, a1, sigma_ow, eta1, eta2, dt_min, dt_max, tso, mtt, mmspt;
  float al_over_as, dxdx, dxa0, neg_a1, neg_a2, neg_w, neg_as, span;
};

// Integration state of an item (include/cochlea.cuh:13620-13700 names in comments).
struct ItemState {
  float pd[4], pv[4], pa[4], ppsi[4], ppsid[4], pxis[4], pxi[4];    // committed
  float nd[4], nv[4], na[4], npsi[4], npsid[4], nxis[4], nxi[4], Ptm[4];
  float p[4];                                                          // resident pressure
  float ow_pd, ow_pv, ow_pa, ow_nd, ow_nv, ow_na;                      // oval window
  float dt, t, sx0, sx1;
  int in_s;
  float c_fail, c_dt, c_steps, c_passes;                               // sx[11], sx[12], sx[13], sx[15]
  int64_t n_steps, n_passes, n_sweeps;
  int32_t fb;
  float ctb, clb, cjb;
};

template <bool TOL>
__device__ __forceinline__ void load_consts(const BmLaunch &a, const cm_bm_item &it, int lane, ItemConst &k) {
  const int base = 4 * lane;
  {
    const float4 m4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_MASS * CM_SECTIONS + base);
    const float4 r4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_RSM * CM_SECTIONS + base);
    const float4 R4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_R * CM_SECTIONS + base);
    const float4 S4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_S * CM_SECTIONS + base);
    const float4 Q4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_Q * CM_SECTIONS + base);
    const float4 O4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_SOHC * CM_SECTIONS + base);
    const float4 T4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_STM * CM_SECTIONS + base);
    const float4 U4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_RTM * CM_SECTIONS + base);
    const float4 G4 = *reinterpret_cast<const float4 *>(a.gamma + (int64_t)it.set * CM_SECTIONS + base);
    const float mm[4] = {m4.x, m4.y, m4.z, m4.w};
    k.rsm[0] = r4.x; k.rsm[1] = r4.y; k.rsm[2] = r4.z; k.rsm[3] = r4.w;
    k.R[0] = R4.x; k.R[1] = R4.y; k.R[2] = R4.z; k.R[3] = R4.w;
    k.S[0] = S4.x; k.S[1] = S4.y; k.S[2] = S4.z; k.S[3] = S4.w;
    k.Q[0] = Q4.x; k.Q[1] = Q4.y; k.Q[2] = Q4.z; k.Q[3] = Q4.w;
    k.Sohc[0] = O4.x; k.Sohc[1] = O4.y; k.Sohc[2] = O4.z; k.Sohc[3] = O4.w;
    k.Stm[0] = T4.x; k.Stm[1] = T4.y; k.Stm[2] = T4.z; k.Stm[3] = T4.w;
    k.Rtm[0] = U4.x; k.Rtm[1] = U4.y; k.Rtm[2] = U4.z; k.Rtm[3] = U4.w;
    k.gam[0] = G4.x; k.gam[1] = G4.y; k.gam[2] = G4.z; k.gam[3] = G4.w;
    UNROLL for (int j = 0; j < 4; ++j) k.rmass[j] = 1.0f / mm[j];
    if (TOL) {
      const float4 M4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_M * CM_SECTIONS + base);
      k.mf[0] = M4.x; k.mf[1] = M4.y; k.mf[2] = M4.z; k.mf[3] = M4.w;
    } else {
      UNROLL for (int j = 0; j < 4; ++j) k.mf[j] = 0.0f;
    }
  }
  k.low23 = lane == 63 ? 0.0f : 1.0f;
}

template <bool TOL>
__device__ __forceinline__ void item_setup(const BmLaunch &a, const cm_bm_item &it, int lane, ItemConst &k,
                                           ItemScal &c, ItemState &s) {
  const cm_bm_scalars &sc = a.sc;
  load_consts<TOL>(a, it, lane, k);

  c.w_ohc = sc.w_ohc; c.alpha_r = sc.alpha_r; c.Ts = sc.Ts;
  c.al = sc.alpha_l; c.Gme = sc.Gme; c.a1 = sc.a1;
  c.sigma_ow = sc.sigma_ow; c.eta1 = sc.eta1; c.eta2 = sc.eta2;
  c.dt_min = sc.dt_min; c.dt_max = sc.dt_max; c.tso = sc.time_step_out;
  c.mtt = it.tol; c.mmspt = it.m1tol;
  c.al_over_as = sc.alpha_l / sc.alpha_s;
  c.dxdx = sc.delta_x * sc.delta_x;
  c.dxa0 = sc.delta_x * sc.a0;
  c.neg_a1 = -sc.a1; c.neg_a2 = -sc.a2; c.neg_w = -sc.w_ohc; c.neg_as = -sc.alpha_s;
  c.span = (float)(it.mesi - it.mosi);

  UNROLL for (int j = 0; j < 4; ++j) {
    s.pd[j] = 0; s.pv[j] = 0; s.pa[j] = 0; s.ppsi[j] = 0; s.ppsid[j] = 0; s.pxis[j] = 0; s.pxi[j] = 0;
    s.nd[j] = 0; s.nv[j] = 0; s.na[j] = 0; s.npsi[j] = 0; s.npsid[j] = 0; s.nxis[j] = 0; s.nxi[j] = 0;
    s.Ptm[j] = 0; s.p[j] = 0;
  }
  s.ow_pd = 0; s.ow_pv = 0; s.ow_pa = 0; s.ow_nd = 0; s.ow_nv = 0; s.ow_na = 0;
  s.dt = sc.time_step;
  s.t = 0.0f;
  s.in_s = it.start;
  s.sx0 = in_at(a, s.in_s);
  s.sx1 = in_at(a, (int64_t)s.in_s + 1);
  s.c_fail = 0.f; s.c_dt = 0.f; s.c_steps = 0.f; s.c_passes = 0.f;
  s.n_steps = 0; s.n_passes = 0; s.n_sweeps = 0;
  s.fb = 0;
  s.ctb = 0.f; s.clb = 0.f; s.cjb = 0.f;
}

// Predictor (Euler) or corrector (trapezoid) update, :13770-13820.
__device__ __forceinline__ void pass_begin(ItemState &s, bool predictor) {
  const float dt = s.dt;
  if (predictor) {
    UNROLL for (int j = 0; j < 4; ++j) {
      s.nxi[j] = s.pxis[j] * dt + s.pxi[j];
      s.nv[j] = s.pa[j] * dt + s.pv[j];
      s.nd[j] = ((s.pv[j] + s.nv[j]) * 0.5f) * dt + s.pd[j];
      s.npsi[j] = s.ppsid[j] * dt + s.ppsi[j];
    }
    s.ow_nd = s.ow_pv * dt + s.ow_pd;
    s.ow_nv = s.ow_pa * dt + s.ow_pv;
  } else {
    UNROLL for (int j = 0; j < 4; ++j) {
      s.nv[j] = ((s.pa[j] + s.na[j]) * 0.5f) * dt + s.pv[j];
      s.nd[j] = ((s.pv[j] + s.nv[j]) * 0.5f) * dt + s.pd[j];
      s.nxi[j] = ((s.pxis[j] + s.nxis[j]) * 0.5f) * dt + s.pxi[j];
      s.npsi[j] = ((s.ppsid[j] + s.npsid[j]) * 0.5f) * dt + s.ppsi[j];
    }
    s.ow_nd = ((s.ow_pv + s.ow_nv) * 0.5f) * dt + s.ow_pd;
    s.ow_nv = ((s.ow_pa + s.ow_na) * 0.5f) * dt + s.ow_pv;
  }
}

// TM/OHC block (passes < 2, :13839-13898), G (:13905-13909) and the right-hand
// side with its boundaries (:13925-13954).  Returns the forcing s(t).
__device__ __forceinline__ float pass_rhs(const ItemConst &k, const ItemScal &c, ItemState &s, bool aux_pass, int lane,
                                          float (&od)[4], float (&G)[4], float (&Y)[4]) {
  UNROLL for (int j = 0; j < 4; ++j) {
    od[j] = s.nxi[j] - s.nd[j];
    const float ohc_p = (k.gam[j] * k.Sohc[j]) * od[j];
    if (aux_pass) {
      float aux = cm_tanhf(c.al_over_as * s.npsi[j]);
      const float dL = c.neg_as * aux;
      aux = aux * aux - 1.0f;
      aux = c.al * aux;
      float n = (c.eta2 * s.nv[j] + c.eta1 * od[j]) - c.w_ohc * s.npsi[j];
      n = k.Stm[j] * dL + (n * k.Rtm[j]) * aux;
      const float ptmp = k.Stm[j] * s.nxi[j] + ohc_p;
      n = ptmp - k.gam[j] * n;
      float den = (k.gam[j] * aux) * c.eta2 - 1.0f;
      den = k.Rtm[j] * den;
      s.nxis[j] = n / den;
      s.Ptm[j] = k.Rtm[j] * s.nxis[j] + k.Stm[j] * s.nxi[j];
    }
    float g = s.nv[j] * s.nv[j];
    g = k.R[j] * (c.alpha_r * g + 1.0f);
    g = g * s.nv[j] + k.S[j] * s.nd[j];
    G[j] = -g - s.Ptm[j];
    Y[j] = (c.dxdx * G[j]) * k.Q[j];
  }
  const float delta = s.t / c.Ts;
  const float sf = s.sx1 * delta + s.sx0 * (1.0f - delta);
  const float Y0 = c.dxa0 * (((c.neg_a2 * s.ow_nv) - c.a1 * s.ow_nd) + c.Gme * sf);
  if (lane == 0) Y[0] = Y0;
  if (lane == 63) Y[3] = 0.0f;
  return sf;
}

// Jacobi until |Y - A p|_2 <= tol |Y|_2 (checked every check_every sweeps,
// capped at max_sweeps); returns the sweeps done.  rr is |Y - A p|_2^2 of the
// incoming p (the guess selection has already computed it).
template <int U = 2>
__device__ __forceinline__ int jacobi_tol(const cm_bm_scalars &sc, const ItemConst &k, const float (&Y)[4],
                                          float yy, float (&p)[4], float rr) {
  const float tau2 = sc.tol_rel * sc.tol_rel;
  int J = 0;
  for (;;) {
    if (rr <= tau2 * yy || J >= sc.max_sweeps) break;
    const int kk = sc.check_every < sc.max_sweeps - J ? sc.check_every : sc.max_sweeps - J;
    jacobi_sweeps<U>(Y, k.rsm, k.low23, p, kk);
    J += kk;
    rr = resid2(Y, k.mf, k.low23, p);
  }
  return J;
}

// Acceleration, oval-window acceleration and psi derivative, :13989-14013.
__device__ __forceinline__ void pass_finish(const ItemConst &k, const ItemScal &c, ItemState &s, const float (&od)[4],
                                            const float (&G)[4], float sf) {
  UNROLL for (int j = 0; j < 4; ++j) s.na[j] = (s.p[j] + G[j]) * k.rmass[j];
  {
    const float p0 = readlane(s.p[0], 0);
    const float NN1 = c.neg_a2 * s.ow_nv;
    const float NN2 = c.Gme * sf + p0;
    const float NN3 = c.neg_a1 * s.ow_nd;
    s.ow_na = ((NN1 + NN2) + NN3) / c.sigma_ow;
  }
  UNROLL for (int j = 0; j < 4; ++j)
    s.npsid[j] = (c.eta1 * od[j] + c.eta2 * (s.nxis[j] - s.nv[j])) + c.neg_w * s.npsi[j];
}

// Step control (:14021-14088) and the pass counter; Q1 lane-0 semantics.
__device__ __forceinline__ void pass_control(const cm_bm_scalars &sc, const ItemScal &c, ItemState &s,
                                             const float (&vref)[4], const float (&aref)[4], int &out_loop,
                                             int &reset, int &another) {
  if (reset == 0) {
    float ex[4], ey[4], ez[4];
    UNROLL for (int j = 0; j < 4; ++j) {
      ex[j] = __builtin_fabsf(aref[j] - s.na[j]);
      ey[j] = __builtin_fabsf(s.nv[j] - vref[j]);
      ez[j] = ey[j] - fmaxf(c.mtt, s.nv[j] * s.dt);
    }
    float emax = tree_max(ex[0], ex[1], ex[2], ex[3]);
    float esum = tree_sum(ey[0], ey[1], ey[2], ey[3]);
    float fsum = tree_sum(ez[0], ez[1], ez[2], ez[3]);
    if (sc.decision_mode == 1) { emax = 0.f; esum = 0.f; fsum = 0.f; }
    float L;
    if (esum > c.mmspt) { L = emax * s.dt; L = L / esum; } else L = 0.0f;
    if (L > 6.0f) {
      s.dt = s.dt * 0.5f;
      if (s.dt < c.dt_min) s.dt = c.dt_min;
      another = 0; reset = 1;
    } else if (fsum > 0.0f) {
      another = 1; reset = 0;
    } else if (L < 0.5f) {
      s.dt = s.dt * 2.0f;
      if (s.dt > c.dt_max) s.dt = c.dt_max;
      another = 0; reset = 1;
    } else {
      another = 1; reset = 0;
    }
  } else {
    another = 1; reset = 0;
  }
  out_loop++;
  if (out_loop >= sc.max_passes) { another = 0; reset = 1; s.c_fail = 1.f; }
}

// Commit of a step: counters, sample advance with output row (:14130-14203),
// state shift and local time (:14212-14230).
__device__ __forceinline__ void step_commit(const BmLaunch &a, const cm_bm_item &it, const ItemScal &c, ItemState &s,
                                            int out_loop, int lane) {
  s.c_steps += 1.0f;
  s.c_dt += s.dt;
  s.c_passes += (float)out_loop;
  s.n_steps++;
  s.n_passes += out_loop;
  if (s.t > c.tso) {
    s.t -= c.tso;
    s.sx0 = s.sx1;
    s.sx1 = in_at(a, (int64_t)s.in_s + 2);
    if (s.in_s + 1 >= it.mosi) {
      const float div_times = s.c_dt / s.c_steps;
      const float div_j = s.c_passes / s.c_steps;
      if (s.in_s >= it.write_start && s.in_s >= 0 && s.in_s < a.rows) {   // rows outside the buffer are dropped
        if (a.speeds)
          *reinterpret_cast<float4 *>(a.speeds + (int64_t)s.in_s * CM_SECTIONS + 4 * lane) =
              make_float4(s.nv[0], s.nv[1], s.nv[2], s.nv[3]);
        if (lane == 0) {
          if (a.failed_nodes) a.failed_nodes[s.in_s] = (int32_t)s.c_fail;
          if (a.conv_time) a.conv_time[s.in_s] = div_times;
          if (a.conv_jacobi) a.conv_jacobi[s.in_s] = div_j;
          if (a.conv_loops) a.conv_loops[s.in_s] = s.c_steps;
        }
      }
      if (s.c_fail > 0.f) s.fb += 1;
      s.ctb += div_times / c.span;
      s.cjb += div_j / c.span;
      s.clb += s.c_steps / c.span;
    }
    s.c_fail = 0.f; s.c_dt = 0.f; s.c_steps = 0.f; s.c_passes = 0.f;
    s.in_s++;
  }
  UNROLL for (int j = 0; j < 4; ++j) {
    s.pd[j] = s.nd[j]; s.pv[j] = s.nv[j]; s.pa[j] = s.na[j]; s.ppsi[j] = s.npsi[j];
    s.ppsid[j] = s.npsid[j]; s.pxis[j] = s.nxis[j]; s.pxi[j] = s.nxi[j];
  }
  s.ow_pd = s.ow_nd; s.ow_pv = s.ow_nv; s.ow_pa = s.ow_na;
  s.t += s.dt;
}

__device__ __forceinline__ void item_final(const BmLaunch &a, const ItemState &s, int item_index, int lane) {
  if (lane == 0) {
    if (a.failed_blocks) a.failed_blocks[item_index] = s.fb;
    if (a.conv_time_blocks) a.conv_time_blocks[item_index] = s.ctb;
    if (a.conv_loops_blocks) a.conv_loops_blocks[item_index] = s.clb;
    if (a.conv_jacobi_blocks) a.conv_jacobi_blocks[item_index] = s.cjb;
    if (a.stats) {
      a.stats[item_index].steps = s.n_steps;
      a.stats[item_index].passes = s.n_passes;
      a.stats[item_index].sweeps = s.n_sweeps;
    }
  }
}

// Extrapolated predictor-pass guess: p <- p + (dt/dt_prev)(p - p_prev) when
// that lowers the residual (ties keep the previous pressure).  The previous
// step's pressure and dt_prev live in LDS (pold[0..63], pold[64].x).  Returns
// the squared residual of the plain guess; *rre gets the extrapolation's, and
// p holds the extrapolation on return when keep_lin, else the choice.
__device__ __forceinline__ float extrapolate_guess(const ItemConst &k, const float (&Y)[4], float dt, float4 *pold,
                                                   int lane, float (&p)[4], float &rre, bool keep_lin) {
  float *dt_prev = reinterpret_cast<float *>(pold + 64);
  const float rho = dt / *dt_prev;
  const float rr0 = resid2(Y, k.mf, k.low23, p);
  const float4 po = pold[lane];
  pold[lane] = make_float4(p[0], p[1], p[2], p[3]);
  p[0] = p[0] + rho * (p[0] - po.x);
  p[1] = p[1] + rho * (p[1] - po.y);
  p[2] = p[2] + rho * (p[2] - po.z);
  p[3] = p[3] + rho * (p[3] - po.w);
  rre = resid2(Y, k.mf, k.low23, p);
  if (!keep_lin && !(rre < rr0)) {
    const float4 pk = pold[lane];
    p[0] = pk.x; p[1] = pk.y; p[2] = pk.z; p[3] = pk.w;
  }
  if (lane == 0) *dt_prev = dt;
  return rr0;
}

// One time block, include/cochlea.cuh:13620-14232.  TOL selects the
// sweep-to-tolerance Jacobi; otherwise the reference's fixed 20/2 sweeps.
// WS = 1 adds the extrapolated predictor-pass guess (tolerance mode only).
// U is the Jacobi sweep unroll.
template <bool TOL, int WS, int U>
__device__ void run_item(const BmLaunch &a, int item_index, int lane, float4 *pold) {
  const cm_bm_item it = a.items[item_index];
  ItemConst k;
  ItemScal c;
  ItemState s;
  item_setup<TOL>(a, it, lane, k, c, s);
  if (WS == 1) {
    pold[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (lane == 0) reinterpret_cast<float *>(pold + 64)[0] = s.dt;
  }

  while (s.in_s < it.mesi) {
    int out_loop = 0, reset = 1, another = 1;
    while (another) {
      float vref[4], aref[4];
      UNROLL for (int j = 0; j < 4; ++j) { vref[j] = s.nv[j]; aref[j] = s.na[j]; }
      const bool predictor = out_loop == 0 || reset;
      if (predictor) out_loop = 0;
      pass_begin(s, predictor);
      float od[4], G[4], Y[4];
      const float sf = pass_rhs(k, c, s, out_loop < 2, lane, od, G, Y);
      int J;
      if (TOL) {
        const float yy = tree_sum(Y[0] * Y[0], Y[1] * Y[1], Y[2] * Y[2], Y[3] * Y[3]);
        float rr;
        if (WS == 1 && out_loop == 0) {
          float rre;
          const float rr0 = extrapolate_guess(k, Y, s.dt, pold, lane, s.p, rre, false);
          rr = rre < rr0 ? rre : rr0;   // residual of the guess kept
        } else {
          rr = resid2(Y, k.mf, k.low23, s.p);
        }
        J = jacobi_tol<U>(a.sc, k, Y, yy, s.p, rr);
      } else {
        J = out_loop < 2 ? a.sc.j_slow : a.sc.j_fast;
        jacobi_sweeps<U>(Y, k.rsm, k.low23, s.p, J);
      }
      s.n_sweeps += J;
      pass_finish(k, c, s, od, G, sf);
      pass_control(a.sc, c, s, vref, aref, out_loop, reset, another);
    }
    step_commit(a, it, c, s, out_loop, lane);
  }
  item_final(a, s, item_index, lane);
}

template <int WAVES, bool TOL, int WS, int U>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(3))) void bm_kernel(BmLaunch a) {
  const int lane = threadIdx.x & 63;
  // wave-uniform by construction: readfirstlane lets the compiler keep the
  // item's fields and the whole step control in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int idx = blockIdx.x * WAVES + wave;
  __shared__ float4 pold_lds[WS ? WAVES * 65 : 1];
  if (idx < a.n_items)
    run_item<TOL, WS, U>(a, __builtin_amdgcn_readfirstlane(a.order ? a.order[idx] : idx), lane,
                      pold_lds + (WS ? wave * 65 : 0));
}

// ------------------------------------------- learned warm start ---------
// A workgroup of G waves integrates G time blocks in lockstep, one step per
// round.  At each step's predictor pass every wave writes its features to
// LDS and the G waves evaluate the 769 -> 512 -> 512 -> 256 GELU MLP together
// on MFMA (v_mfma_f32_16x16x16_bf16, rows = blocks, padded to 16), with the
// BF16 weights streamed from L2 in MFMA fragment order.  Each wave then takes
// the plain, extrapolated or learned guess with the smallest residual and
// iterates Jacobi to tolerance as in the fused kernel.
constexpr int WS_XS = 808;    // LDS row stride (bf16) of the feature tile (K = 800 read, 769 used)
constexpr int WS_HS = 520;    // row stride (bf16) of hidden tiles
constexpr int WS_OS = 260;    // row stride (float) of the output tile

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 ws_bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float ws_f32x4;

// One dense layer: out[16][N] = act(A[16][K] W + b) with NT = N/16 column tiles
// shared out among the G waves (up to three tiles per wave, sharing the A
// fragment of each k step).
#define CM_LDS __attribute__((address_space(3)))
#define CM_GLB __attribute__((address_space(1)))
typedef uint32_t ws_u32x4 __attribute__((ext_vector_type(4)));

// In-solver inference only: GELU (tanh form) as x * sigmoid(2u), u =
// sqrt(2/pi) (x + 0.044715 x^3), with the hardware exp2 and reciprocal, and the
// hardware round-to-nearest-even BF16 conversion.  Both agree with the exact
// forms to a few FP32 ulps, far inside the BF16 rounding of the activations
// (training keeps cm_gelu).
__device__ __forceinline__ float ws_gelu(float x) {
  const float t = x * x;
  const float z = (x * __builtin_fmaf(0.044715f, t, 1.0f)) * -2.302208198f;   // -2 sqrt(2/pi) log2(e)
  return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z));
}
__device__ __forceinline__ uint16_t ws_bf16(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
__device__ __forceinline__ uint32_t ws_bf16x2(float lo, float hi) {
  return (uint32_t)ws_bf16(lo) | ((uint32_t)ws_bf16(hi) << 16);
}

// NTL (1..3) column tiles t0, t0 + G, t0 + 2G of one layer, sharing the A
// fragments of each k step.  Weights come as 16-byte fragments of k-tile pairs
// (uint4 = this lane's 8 rows 32 j + 8 (lane / 16) + 0..7 of one column, the B
// operand of a K = 32 MFMA); a ring of D pairs per tile stays in
// flight: pair kp + D is requested as soon as pair kp has been consumed.  The
// ring is one register array of WS_RING entries, slot (j, i) at j * NTL + i.
constexpr int WS_RING = 9;
template <int NTL>
struct WsDepth {
  static constexpr int D = NTL == 3 ? 3 : (NTL == 2 ? 4 : 6);
};

// First D pairs of every tile: global loads only, so a wave can issue them
// before the barrier that publishes the layer's input tile.
template <int G, int KT, int NT, int NTL>
__device__ __forceinline__ void ws_fill(ws_u32x4 (&ring)[WS_RING], const CM_GLB ws_u32x4 *W, int lane, int t0) {
  constexpr int KP = (KT + 1) / 2, D = WsDepth<NTL>::D;
  const CM_GLB ws_u32x4 *wl = W + t0 * 64 + lane;
  UNROLL for (int j = 0; j < D; ++j)
    if (j < KP) {
      UNROLL for (int i = 0; i < NTL; ++i) ring[j * NTL + i] = wl[(j * NT + i * G) * 64];
    }
}

// The k loop over a filled ring and the bias / activation epilogue.
template <int G, int KT, int NT, bool GELU, int NTL>
__device__ __forceinline__ void ws_run(ws_u32x4 (&ring)[WS_RING], const CM_LDS uint16_t *A, int sa,
                                       const CM_GLB ws_u32x4 *W, const CM_GLB float *bias, CM_LDS void *out, int so,
                                       int lane, int t0) {
  constexpr int KP = (KT + 1) / 2, D = WsDepth<NTL>::D;
  ws_f32x4 acc[NTL];
  UNROLL for (int i = 0; i < NTL; ++i) acc[i] = ws_f32x4{0.f, 0.f, 0.f, 0.f};
  const CM_GLB ws_u32x4 *wl = W + t0 * 64 + lane;
  const CM_LDS uint16_t *arow = A + (lane & 15) * sa + 8 * (lane >> 4);
  __builtin_amdgcn_sched_barrier(0);
  UNROLL for (int kp = 0; kp < KP; ++kp) {
    const int slot = kp % D;
    // one K = 32 MFMA per pair of k tiles: this lane's 8 consecutive inputs
    // 32 kp + 8 (lane / 16) + 0..7 (one 16-byte LDS read) against the same 8
    // rows of the weight fragment
    const ws_bf16x8 a8 = __builtin_bit_cast(ws_bf16x8, *reinterpret_cast<const CM_LDS ws_u32x4 *>(arow + 32 * kp));
    UNROLL for (int i = 0; i < NTL; ++i)
      acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a8, __builtin_bit_cast(ws_bf16x8, ring[slot * NTL + i]),
                                                       acc[i], 0, 0, 0);
    if (kp + D < KP) {
      UNROLL for (int i = 0; i < NTL; ++i) ring[slot * NTL + i] = wl[((kp + D) * NT + i * G) * 64];
    }
    __builtin_amdgcn_sched_barrier(0);   // keep the requests D pairs ahead of their use
  }
  // lane holds rows 4*(lane/16)+q of column 16*t + lane%16
  const int r0 = 4 * (lane >> 4), cl = lane & 15;
  UNROLL for (int i = 0; i < NTL; ++i) {
    const int col = 16 * (t0 + i * G) + cl;
    const float bv = bias[col];
    UNROLL for (int q = 0; q < 4; ++q) {
      const float v = acc[i][q] + bv;
      if (GELU)
        reinterpret_cast<CM_LDS uint16_t *>(out)[(r0 + q) * so + col] = ws_bf16(ws_gelu(v));
      else
        reinterpret_cast<CM_LDS float *>(out)[(r0 + q) * so + col] = v;
    }
  }
}

// One dense layer out[16][N] = act(A[16][K] W + b), N = 16 NT: wave w computes
// column tiles w, w + G, w + 2G, then w + 3G, ... three at a time.  Every wave
// has a first group of tiles (G < NT); its weight ring is filled before the
// layer's barrier, which every wave then passes exactly once, so the first
// requests are in flight while the slowest wave finishes the previous stage.
template <int G, int KT, int NT, bool GELU>
__device__ __noinline__ void ws_layer(const CM_LDS uint16_t *A, int sa, const CM_GLB ws_u32x4 *W,
                                      const CM_GLB float *bias, CM_LDS void *out, int so, int wave, int lane) {
  static_assert(G < NT, "every wave needs a first tile group");
  ws_u32x4 ring[WS_RING];
  const int n0 = wave + 2 * G < NT ? 3 : (wave + G < NT ? 2 : 1);
  if (n0 == 3)
    ws_fill<G, KT, NT, 3>(ring, W, lane, wave);
  else if (n0 == 2)
    ws_fill<G, KT, NT, 2>(ring, W, lane, wave);
  else
    ws_fill<G, KT, NT, 1>(ring, W, lane, wave);
  __syncthreads();
  if (n0 == 3)
    ws_run<G, KT, NT, GELU, 3>(ring, A, sa, W, bias, out, so, lane, wave);
  else if (n0 == 2)
    ws_run<G, KT, NT, GELU, 2>(ring, A, sa, W, bias, out, so, lane, wave);
  else
    ws_run<G, KT, NT, GELU, 1>(ring, A, sa, W, bias, out, so, lane, wave);
  for (int t0 = wave + 3 * G; t0 < NT; t0 += 3 * G) {
    const int n = t0 + 2 * G < NT ? 3 : (t0 + G < NT ? 2 : 1);
    if (n == 3) {
      ws_fill<G, KT, NT, 3>(ring, W, lane, t0);
      ws_run<G, KT, NT, GELU, 3>(ring, A, sa, W, bias, out, so, lane, t0);
    } else if (n == 2) {
      ws_fill<G, KT, NT, 2>(ring, W, lane, t0);
      ws_run<G, KT, NT, GELU, 2>(ring, A, sa, W, bias, out, so, lane, t0);
    } else {
      ws_fill<G, KT, NT, 1>(ring, W, lane, t0);
      ws_run<G, KT, NT, GELU, 1>(ring, A, sa, W, bias, out, so, lane, t0);
    }
  }
}

// RMS over the 256 sections, for the network's input scales (the hardware
// square root, 1 ulp, is enough there)
__device__ __forceinline__ float rms256(float x0, float x1, float x2, float x3) {
  return __builtin_amdgcn_sqrtf(tree_sum(x0 * x0, x1 * x1, x2 * x2, x3 * x3) * (1.0f / 256.0f)) + 1e-30f;
}

template <int G>
__global__ __launch_bounds__(64 * G) void bm_ws_kernel(BmLaunch a, WsArgs w) {
  __shared__ __align__(16) uint16_t Xs[16 * WS_XS];
  __shared__ __align__(16) uint16_t H1s[16 * WS_HS];
  __shared__ __align__(16) uint16_t H2s[16 * WS_HS];      // also the output tile (16 x WS_OS floats)
  __shared__ float4 pold_lds[G * 65];
  __shared__ float4 keep_lds[G * 5 * 64];                  // predictor right-hand side across the network
  __shared__ int active_lds[2][G];                         // double-buffered by round parity

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int slot = blockIdx.x * G + wave;
  const bool valid = slot < a.n_items;
  const int idx = valid ? (a.order ? a.order[slot] : slot) : slot;   // item integrated by this wave
  float4 *pold = pold_lds + wave * 65;
  float4 *keep = keep_lds + wave * 5 * 64;   // Y, od, G, nxis, Ptm of the predictor pass

  // zero the feature tile once (rows >= G stay zero)
  for (int e = threadIdx.x; e < 16 * WS_XS; e += 64 * G) Xs[e] = 0;

  cm_bm_item it;
  ItemConst k;
  ItemScal c;
  ItemState s;
  if (valid) {
    it = a.items[idx];
  } else {
    it.start = 0; it.mosi = 0; it.mesi = 0; it.write_start = 0; it.tol = 0.f; it.m1tol = 0.f; it.set = 0;
    it.pad = 0;
  }
  item_setup<true>(a, it, lane, k, c, s);
  pold[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (lane == 0) reinterpret_cast<float *>(pold + 64)[0] = s.dt;
  __syncthreads();

  const float inv_256 = 1.0f / 256.0f;
  for (int round = 0;; ++round) {
    const bool active = valid && s.in_s < it.mesi;
    // the network runs on every net_every-th round of the group (group-uniform)
    const bool net_round = w.use_net && (round % w.net_every) == 0;
    float rr0 = 0.f, rre = 0.f, rr_scale = 0.f;
    if (active) {
      float od[4], G4[4], Y[4];
      pass_begin(s, true);
      const float sf = pass_rhs(k, c, s, true, lane, od, G4, Y);
      keep[lane] = make_float4(Y[0], Y[1], Y[2], Y[3]);
      keep[64 + lane] = make_float4(od[0], od[1], od[2], od[3]);
      keep[128 + lane] = make_float4(G4[0], G4[1], G4[2], G4[3]);
      keep[192 + lane] = make_float4(s.nxis[0], s.nxis[1], s.nxis[2], s.nxis[3]);
      keep[256 + lane] = make_float4(s.Ptm[0], s.Ptm[1], s.Ptm[2], s.Ptm[3]);
      if (lane == 0) reinterpret_cast<float *>(pold + 64)[1] = sf;
      rr0 = extrapolate_guess(k, Y, s.dt, pold, lane, s.p, rre, true);   // s.p = p_lin
      rr_scale = __builtin_amdgcn_sqrtf(rre * inv_256) + 1e-30f;
      if (net_round || w.capture) {
      float r[4];
      resid_vec(Y, k.mf, k.low23, s.p, r);
      const float4 pk = pold[lane];   // plain guess (previous pressure)
      // the scales only normalise BF16 network inputs: the hardware reciprocal
      // (1 ulp) stands in for an IEEE division
      const float inv_d = __builtin_amdgcn_rcpf(rms256(s.nd[0], s.nd[1], s.nd[2], s.nd[3]));
      const float inv_v = __builtin_amdgcn_rcpf(rms256(s.nv[0], s.nv[1], s.nv[2], s.nv[3]));
      const float inv_r = __builtin_amdgcn_rcpf(rr_scale);
      const float inv_p = __builtin_amdgcn_rcpf(rms256(pk.x, pk.y, pk.z, pk.w));
      float f[13];
      UNROLL for (int j = 0; j < 4; ++j) {
        f[j] = s.nd[j] * inv_d;
        f[4 + j] = s.nv[j] * inv_v;
        f[8 + j] = r[j] * inv_r;
      }
      f[12] = sf * inv_p;
      uint16_t *xr = Xs + wave * WS_XS;
      UNROLL for (int b = 0; b < 3; ++b)
        *reinterpret_cast<uint2 *>(xr + 256 * b + 4 * lane) =
            make_uint2(ws_bf16x2(f[4 * b], f[4 * b + 1]), ws_bf16x2(f[4 * b + 2], f[4 * b + 3]));
      if (lane == 0) xr[768] = ws_bf16(f[12]);
      if (w.capture && (s.n_steps % w.stride) == 0 && s.n_steps / w.stride < w.per_item) {   // slot of this step
        float *rec = w.records + ((int64_t)idx * w.per_item + s.n_steps / w.stride) * CM_WS_RECORD;
        *reinterpret_cast<float4 *>(rec + 4 * lane) = make_float4(f[0], f[1], f[2], f[3]);
        *reinterpret_cast<float4 *>(rec + 256 + 4 * lane) = make_float4(f[4], f[5], f[6], f[7]);
        *reinterpret_cast<float4 *>(rec + 512 + 4 * lane) = make_float4(f[8], f[9], f[10], f[11]);
        *reinterpret_cast<float4 *>(rec + CM_WS_REC_Y + 4 * lane) = make_float4(Y[0], Y[1], Y[2], Y[3]);
        *reinterpret_cast<float4 *>(rec + CM_WS_REC_PLIN + 4 * lane) = make_float4(s.p[0], s.p[1], s.p[2], s.p[3]);
        if (lane == 0) {
          rec[768] = f[12];
          rec[769] = rr_scale;
          rec[770] = (float)s.n_steps;
          rec[771] = 0.f;
        }
      }
      }
    }
    // a fast wave may write the next round's flag before a slow one has read
    // this round's, so the flags alternate between two slots
    if (lane == 0) active_lds[round & 1][wave] = active ? 1 : 0;
    // X -> H2 -> H1 -> output tile in H2's space (H2 is dead once layer 2 ends).
    // Each layer's barrier sits inside ws_layer, after the first weight
    // requests; the first one also publishes the activity flags.  When no
    // block of the group is left the first layer runs once more on a stale
    // tile, and its result is never read.
    typedef const CM_LDS uint16_t *lds_in;
    typedef CM_LDS void *lds_out;
    typedef const CM_GLB ws_u32x4 *glb_w;
    typedef const CM_GLB float *glb_b;
    if (net_round)
      ws_layer<G, 49, 32, true>((lds_in)Xs, WS_XS, (glb_w)w.w1, (glb_b)w.b1, (lds_out)H2s, WS_HS, wave, lane);
    else
      __syncthreads();
    int any = 0;
    for (int g = 0; g < G; ++g) any |= active_lds[round & 1][g];
    if (!any) break;
    if (net_round) {
      ws_layer<G, 32, 32, true>((lds_in)H2s, WS_HS, (glb_w)w.w2, (glb_b)w.b2, (lds_out)H1s, WS_HS, wave, lane);
      ws_layer<G, 32, 16, false>((lds_in)H1s, WS_HS, (glb_w)w.w3, (glb_b)w.b3, (lds_out)H2s, WS_OS, wave, lane);
      __syncthreads();
    }
    // The section constants are re-read from L1/L2 here, so that they are not
    // held in registers across the network evaluation (the memory clobber keeps
    // the compiler from reusing the earlier loads).
    asm volatile("" ::: "memory");
    load_consts<true>(a, it, lane, k);
    if (active) {
      // the predictor state again (same inputs, same values) and its
      // right-hand side from LDS: cheaper than holding either in registers
      // across the network
      float od[4], G4[4], Y[4];
      pass_begin(s, true);
      {
        const float4 y = keep[lane], o = keep[64 + lane], g = keep[128 + lane], x = keep[192 + lane],
                     t = keep[256 + lane];
        Y[0] = y.x; Y[1] = y.y; Y[2] = y.z; Y[3] = y.w;
        od[0] = o.x; od[1] = o.y; od[2] = o.z; od[3] = o.w;
        G4[0] = g.x; G4[1] = g.y; G4[2] = g.z; G4[3] = g.w;
        s.nxis[0] = x.x; s.nxis[1] = x.y; s.nxis[2] = x.z; s.nxis[3] = x.w;
        s.Ptm[0] = t.x; s.Ptm[1] = t.y; s.Ptm[2] = t.z; s.Ptm[3] = t.w;
      }
      const float sf = reinterpret_cast<const float *>(pold + 64)[1];
      const float yy = tree_sum(Y[0] * Y[0], Y[1] * Y[1], Y[2] * Y[2], Y[3] * Y[3]);
      float rr_sel = rre;   // residual of the guess kept (p_lin unless replaced below)
      if (net_round) {
        // learned guess: p_lin + rms(r_lin) c_out S(y), S = [1,2,1]/4 (edges replicated)
        const float4 y4 = *reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(H2s) + wave * WS_OS +
                                                             4 * lane);
        if (w.capture && (s.n_steps % w.stride) == 0 && s.n_steps / w.stride < w.per_item)
          *reinterpret_cast<float4 *>(w.records + ((int64_t)idx * w.per_item + s.n_steps / w.stride) * CM_WS_RECORD +
                                      CM_WS_REC_OUT + 4 * lane) = y4;
        float left = wave_shr1(y4.w), right = wave_shl1(y4.x);
        if (lane == 0) left = y4.x;
        if (lane == 63) right = y4.w;
        const float sc_out = rr_scale * w.c_out;
        float ph[4];
        ph[0] = s.p[0] + sc_out * ((0.25f * left + 0.5f * y4.x) + 0.25f * y4.y);
        ph[1] = s.p[1] + sc_out * ((0.25f * y4.x + 0.5f * y4.y) + 0.25f * y4.z);
        ph[2] = s.p[2] + sc_out * ((0.25f * y4.y + 0.5f * y4.z) + 0.25f * y4.w);
        ph[3] = s.p[3] + sc_out * ((0.25f * y4.z + 0.5f * y4.w) + 0.25f * right);
        const float rrh = resid2(Y, k.mf, k.low23, ph);
        if (rrh < rre && rrh < rr0) {
          UNROLL for (int j = 0; j < 4; ++j) s.p[j] = ph[j];
          rr_sel = rrh;
        } else if (!(rre < rr0)) {
          const float4 pk = pold[lane];
          s.p[0] = pk.x; s.p[1] = pk.y; s.p[2] = pk.z; s.p[3] = pk.w;
          rr_sel = rr0;
        }
      } else if (!(rre < rr0)) {
        const float4 pk = pold[lane];
        s.p[0] = pk.x; s.p[1] = pk.y; s.p[2] = pk.z; s.p[3] = pk.w;
        rr_sel = rr0;
      }
      // rest of the predictor pass, then the corrector passes of this step
      int out_loop = 0, reset = 1, another = 1;
      {
        const int J = jacobi_tol(a.sc, k, Y, yy, s.p, rr_sel);
        s.n_sweeps += J;
        pass_finish(k, c, s, od, G4, sf);
        float vref[4], aref[4];   // unused: step control is skipped on the predictor pass
        UNROLL for (int j = 0; j < 4; ++j) { vref[j] = 0.f; aref[j] = 0.f; }
        pass_control(a.sc, c, s, vref, aref, out_loop, reset, another);
      }
      // vref/aref of the first corrector pass are the predictor pass's values
      while (another) {
        float vref[4], aref[4];
        UNROLL for (int j = 0; j < 4; ++j) { vref[j] = s.nv[j]; aref[j] = s.na[j]; }
        const bool predictor = out_loop == 0 || reset;
        if (predictor) out_loop = 0;
        pass_begin(s, predictor);
        float od2[4], G2[4], Y2[4];
        const float sf2 = pass_rhs(k, c, s, out_loop < 2, lane, od2, G2, Y2);
        const float yy2 = tree_sum(Y2[0] * Y2[0], Y2[1] * Y2[1], Y2[2] * Y2[2], Y2[3] * Y2[3]);
        const int J = jacobi_tol(a.sc, k, Y2, yy2, s.p, resid2(Y2, k.mf, k.low23, s.p));
        s.n_sweeps += J;
        pass_finish(k, c, s, od2, G2, sf2);
        pass_control(a.sc, c, s, vref, aref, out_loop, reset, another);
      }
      step_commit(a, it, c, s, out_loop, lane);
    }
  }
  if (valid) item_final(a, s, idx, lane);
}

}  // namespace cm

// Launcher used by the C-ABI layer: waves per workgroup 1, 2 or 4, Jacobi
// unroll 1, 2 or 4.
extern "C" int cm_launch_bm(const cm::BmLaunch *a, int waves_per_block, int unroll, int grid_blocks,
                            hipStream_t stream) {
  if (a->use_queue) return -1;   // the work-queue launch is switched off
  dim3 grid(grid_blocks);
  const bool tol = a->sc.jacobi_mode == 1;
  const int ws = tol ? a->sc.warm_start : 0;
  if (ws < 0 || ws > 1) return -1;
  const int u = unroll == 1 ? 0 : (unroll == 2 ? 1 : (unroll == 4 ? 2 : -1));
  const int w = waves_per_block == 1 ? 0 : (waves_per_block == 2 ? 1 : (waves_per_block == 4 ? 2 : -1));
  if (u < 0 || w < 0) return -1;
  const int mode = tol ? 1 + ws : 0;   // 0 fixed sweeps, 1 tolerance, 2 tolerance + extrapolation
#define CM_BM_CASE(W, U) \
  case ((W) == 1 ? 0 : ((W) == 2 ? 1 : 2)) * 3 + ((U) == 1 ? 0 : ((U) == 2 ? 1 : 2)): \
    if (mode == 0) \
      hipLaunchKernelGGL((cm::bm_kernel<W, false, 0, U>), grid, dim3(64 * (W)), 0, stream, *a); \
    else if (mode == 1) \
      hipLaunchKernelGGL((cm::bm_kernel<W, true, 0, U>), grid, dim3(64 * (W)), 0, stream, *a); \
    else \
      hipLaunchKernelGGL((cm::bm_kernel<W, true, 1, U>), grid, dim3(64 * (W)), 0, stream, *a); \
    break;
  switch (w * 3 + u) {
    CM_BM_CASE(1, 1)
    CM_BM_CASE(1, 2)
    CM_BM_CASE(1, 4)
    CM_BM_CASE(2, 1)
    CM_BM_CASE(2, 2)
    CM_BM_CASE(2, 4)
    CM_BM_CASE(4, 1)
    CM_BM_CASE(4, 2)
    CM_BM_CASE(4, 4)
    default: return -1;
  }
#undef CM_BM_CASE
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Lockstep groups of G blocks, G = 8 when the groups fit one per CU, else 12.
// A round costs the solver work of the busiest SIMD (the waves of a group are
// spread over the CU's 4 SIMDs) plus the network's weight stream, whose time
// grows with the number of CUs streaming from L2 at once.  G = 8 puts 2 waves
// on every SIMD; G = 12 puts 3 on every SIMD, so G = 10 (3, 3, 2, 2) costs the
// same solver time as 12 with more groups streaming weights, and is not used.
// Returns G (> 0) or a negative error.
extern "C" int cm_launch_bm_ws(const cm::BmLaunch *a, const cm::WsArgs *w, int n_cus, hipStream_t stream) {
  if (a->sc.jacobi_mode != 1 || w->net_every < 1) return -1;
  const int n = a->n_items;
  const int G = (n + 7) / 8 <= n_cus ? 8 : 12;
  const dim3 grid((n + G - 1) / G);
  if (G == 8)
    hipLaunchKernelGGL((cm::bm_ws_kernel<8>), grid, dim3(64 * 8), 0, stream, *a, *w);
  else
    hipLaunchKernelGGL((cm::bm_ws_kernel<12>), grid, dim3(64 * 12), 0, stream, *a, *w);
  return hipGetLastError() == hipSuccess ? G : -2;
}
