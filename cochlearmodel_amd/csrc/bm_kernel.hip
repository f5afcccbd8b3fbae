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

// Sum of 256 values, 4 per lane, with the fixed tree.
__device__ __forceinline__ float tree_sum(float x0, float x1, float x2, float x3) {
  float s = (x0 + x1) + (x2 + x3);
  s = s + quad_xor1(s);
  s = s + quad_xor2(s);
  s = s + row_half_mirror(s);   // equals xor 4 once quads are uniform
  s = s + row_mirror(s);        // equals xor 8 once octets are uniform
  return (readlane(s, 0) + readlane(s, 16)) + (readlane(s, 32) + readlane(s, 48));
}
__device__ __forceinline__ float tree_max(float x0, float x1, float x2, float x3) {
  float s = fmaxf(fmaxf(x0, x1), fmaxf(x2, x3));
  s = fmaxf(s, quad_xor1(s));
  s = fmaxf(s, quad_xor2(s));
  s = fmaxf(s, row_half_mirror(s));
  s = fmaxf(s, row_mirror(s));
  return fmaxf(fmaxf(readlane(s, 0), readlane(s, 16)), fmaxf(readlane(s, 32), readlane(s, 48)));
}



#define UNROLL _Pragma("unroll")

__device__ __forceinline__ float in_at(const BmLaunch &a, int64_t idx) {
  return (idx >= 0 && idx < a.input_len) ? a.input[idx] : 0.0f;
}

// J Jacobi sweeps of the kernel convention (include/cochlea.cuh:13962-13982):
// p_i <- (Y_i - (l_i p_{i-1} + p_{i+1})) / m_i with the lane-boundary
// neighbours exchanged by DPP wave shifts (zero past either end).
__device__ __forceinline__ void jacobi_sweeps(const float (&Y)[4], const float (&rsm)[4], float low23, float (&p)[4],
                                              int J) {
  for (int k = 0; k < J; ++k) {
    const float left = wave_shr1(p[3]);
    const float right = wave_shl1(p[0]);
    const float q0 = (Y[0] - (left + p[1])) * rsm[0];
    const float q1 = (Y[1] - (p[0] + p[2])) * rsm[1];
    const float q2 = (Y[2] - (low23 * p[1] + p[3])) * rsm[2];
    const float q3 = (Y[3] - (low23 * p[2] + right)) * rsm[3];
    p[0] = q0; p[1] = q1; p[2] = q2; p[3] = q3;
  }
}

// Squared 2-norm of the residual Y - A p (kernel convention, fixed tree).
__device__ __forceinline__ float resid2(const float (&Y)[4], const float (&mf)[4], float low23, const float (&p)[4]) {
  const float left = wave_shr1(p[3]);
  const float right = wave_shl1(p[0]);
  const float r0 = Y[0] - ((left + p[1]) + mf[0] * p[0]);
  const float r1 = Y[1] - ((p[0] + p[2]) + mf[1] * p[1]);
  const float r2 = Y[2] - ((low23 * p[1] + p[3]) + mf[2] * p[2]);
  const float r3 = Y[3] - ((low23 * p[2] + right) + mf[3] * p[3]);
  return tree_sum(r0 * r0, r1 * r1, r2 * r2, r3 * r3);
}

// One time block, include/cochlea.cuh:13620-14232.  TOL selects the
// sweep-to-tolerance Jacobi; otherwise the reference's fixed 20/2 sweeps.
// WS = 1 adds the extrapolated predictor-pass guess (tolerance mode only).
template <bool TOL, int WS>
__device__ void run_item(const BmLaunch &a, int item_index, int lane, float4 *pold_lds) {
  const cm_bm_item it = a.items[item_index];
  const cm_bm_scalars &sc = a.sc;
  const int base = 4 * lane;

  float rsm[4], rmass[4], gam[4], R[4], S[4], Q[4], Sohc[4], Stm[4], Rtm[4], mf[4];
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
    rsm[0] = r4.x; rsm[1] = r4.y; rsm[2] = r4.z; rsm[3] = r4.w;
    R[0] = R4.x; R[1] = R4.y; R[2] = R4.z; R[3] = R4.w;
    S[0] = S4.x; S[1] = S4.y; S[2] = S4.z; S[3] = S4.w;
    Q[0] = Q4.x; Q[1] = Q4.y; Q[2] = Q4.z; Q[3] = Q4.w;
    Sohc[0] = O4.x; Sohc[1] = O4.y; Sohc[2] = O4.z; Sohc[3] = O4.w;
    Stm[0] = T4.x; Stm[1] = T4.y; Stm[2] = T4.z; Stm[3] = T4.w;
    Rtm[0] = U4.x; Rtm[1] = U4.y; Rtm[2] = U4.z; Rtm[3] = U4.w;
    gam[0] = G4.x; gam[1] = G4.y; gam[2] = G4.z; gam[3] = G4.w;
    UNROLL for (int j = 0; j < 4; ++j) rmass[j] = 1.0f / mm[j];
    if (TOL) {
      const float4 M4 = *reinterpret_cast<const float4 *>(a.sec + CM_SEC_M * CM_SECTIONS + base);
      mf[0] = M4.x; mf[1] = M4.y; mf[2] = M4.z; mf[3] = M4.w;
    } else {
      UNROLL for (int j = 0; j < 4; ++j) mf[j] = 0.0f;
    }
  }
  // low multiplier: 0 for sections 254 and 255 (lane 63, slots 2 and 3)
  const float low23 = lane == 63 ? 0.0f : 1.0f;

  const float w_ohc = sc.w_ohc, dx = sc.delta_x, alpha_r = sc.alpha_r, Ts = sc.Ts;
  const float al = sc.alpha_l, as = sc.alpha_s, Gme = sc.Gme, a1 = sc.a1, a2 = sc.a2;
  const float sigma_ow = sc.sigma_ow, eta1 = sc.eta1, eta2 = sc.eta2;
  const float dt_min = sc.dt_min, dt_max = sc.dt_max, tso = sc.time_step_out;
  const float mtt = it.tol, mmspt = it.m1tol;
  const float al_over_as = al / as;
  const float dxdx = dx * dx;
  const float dxa0 = dx * sc.a0;
  const float neg_a1 = -a1, neg_a2 = -a2, neg_w = -w_ohc, neg_as = -as;

  float pd[4] = {0, 0, 0, 0}, pv[4] = {0, 0, 0, 0}, pa[4] = {0, 0, 0, 0}, ppsi[4] = {0, 0, 0, 0};
  float ppsid[4] = {0, 0, 0, 0}, pxis[4] = {0, 0, 0, 0}, pxi[4] = {0, 0, 0, 0};
  float nd[4] = {0, 0, 0, 0}, nv[4] = {0, 0, 0, 0}, na[4] = {0, 0, 0, 0}, npsi[4] = {0, 0, 0, 0};
  float npsid[4] = {0, 0, 0, 0}, nxis[4] = {0, 0, 0, 0}, nxi[4] = {0, 0, 0, 0}, Ptm[4] = {0, 0, 0, 0};
  float p[4] = {0, 0, 0, 0};
  // previous step's pressure (WS == 1) lives in LDS: it is touched once per step
  if (WS == 1) pold_lds[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
  float ow_pd = 0, ow_pv = 0, ow_pa = 0, ow_nd = 0, ow_nv = 0, ow_na = 0;

  float dt = sc.time_step;
  // dt of the previous step (WS == 1), kept in the spare LDS slot after the pressures
  float *dt_prev_lds = reinterpret_cast<float *>(pold_lds + 64);
  if (WS == 1 && lane == 0) *dt_prev_lds = dt;
  float t = 0.0f;
  int in_s = it.start;
  float sx0 = in_at(a, in_s), sx1 = in_at(a, (int64_t)in_s + 1);
  float c_fail = 0.f, c_dt = 0.f, c_steps = 0.f, c_passes = 0.f;
  int64_t n_steps = 0, n_passes = 0, n_sweeps = 0;
  int32_t fb = 0;
  float ctb = 0.f, clb = 0.f, cjb = 0.f;
  const float span = (float)(it.mesi - it.mosi);

  while (in_s < it.mesi) {
    int out_loop = 0, reset = 1, another = 1;
    while (another) {
      float vref[4], aref[4];
      UNROLL for (int j = 0; j < 4; ++j) { vref[j] = nv[j]; aref[j] = na[j]; }
      if (out_loop == 0 || reset) {
        out_loop = 0;
        UNROLL for (int j = 0; j < 4; ++j) {
          nxi[j] = pxis[j] * dt + pxi[j];
          nv[j] = pa[j] * dt + pv[j];
          nd[j] = ((pv[j] + nv[j]) * 0.5f) * dt + pd[j];
          npsi[j] = ppsid[j] * dt + ppsi[j];
        }
        ow_nd = ow_pv * dt + ow_pd;
        ow_nv = ow_pa * dt + ow_pv;
      } else {
        UNROLL for (int j = 0; j < 4; ++j) {
          nv[j] = ((pa[j] + na[j]) * 0.5f) * dt + pv[j];
          nd[j] = ((pv[j] + nv[j]) * 0.5f) * dt + pd[j];
          nxi[j] = ((pxis[j] + nxis[j]) * 0.5f) * dt + pxi[j];
          npsi[j] = ((ppsid[j] + npsid[j]) * 0.5f) * dt + ppsi[j];
        }
        ow_nd = ((ow_pv + ow_nv) * 0.5f) * dt + ow_pd;
        ow_nv = ((ow_pa + ow_na) * 0.5f) * dt + ow_pv;
      }
      float od[4], G[4], Y[4];
      const bool aux_pass = out_loop < 2;
      UNROLL for (int j = 0; j < 4; ++j) {
        od[j] = nxi[j] - nd[j];
        const float ohc_p = (gam[j] * Sohc[j]) * od[j];
        if (aux_pass) {
          float aux = cm_tanhf(al_over_as * npsi[j]);
          const float dL = neg_as * aux;
          aux = aux * aux - 1.0f;
          aux = al * aux;
          float n = (eta2 * nv[j] + eta1 * od[j]) - w_ohc * npsi[j];
          n = Stm[j] * dL + (n * Rtm[j]) * aux;
          const float ptmp = Stm[j] * nxi[j] + ohc_p;
          n = ptmp - gam[j] * n;
          float den = (gam[j] * aux) * eta2 - 1.0f;
          den = Rtm[j] * den;
          nxis[j] = n / den;
          Ptm[j] = Rtm[j] * nxis[j] + Stm[j] * nxi[j];
        }
        float g = nv[j] * nv[j];
        g = R[j] * (alpha_r * g + 1.0f);
        g = g * nv[j] + S[j] * nd[j];
        G[j] = -g - Ptm[j];
        Y[j] = (dxdx * G[j]) * Q[j];
      }
      const float delta = t / Ts;
      const float s = sx1 * delta + sx0 * (1.0f - delta);
      const float Y0 = dxa0 * (((neg_a2 * ow_nv) - a1 * ow_nd) + Gme * s);
      if (lane == 0) Y[0] = Y0;
      if (lane == 63) Y[3] = 0.0f;

      int J;
      if (TOL) {
        // sweep until |Y - A p|_2 <= tol |Y|_2, checking every check_every sweeps
        const float yy = tree_sum(Y[0] * Y[0], Y[1] * Y[1], Y[2] * Y[2], Y[3] * Y[3]);
        const float tau2 = sc.tol_rel * sc.tol_rel;
        if (WS == 1 && out_loop == 0) {
          // predictor-pass guess: keep the linear extrapolation of the last
          // two steps' pressures when its residual is smaller
          // (in place: p is parked in LDS while the extrapolation is scored)
          const float rho = dt / *dt_prev_lds;
          const float rr0 = resid2(Y, mf, low23, p);
          const float4 po = pold_lds[lane];
          pold_lds[lane] = make_float4(p[0], p[1], p[2], p[3]);
          p[0] = p[0] + rho * (p[0] - po.x);
          p[1] = p[1] + rho * (p[1] - po.y);
          p[2] = p[2] + rho * (p[2] - po.z);
          p[3] = p[3] + rho * (p[3] - po.w);
          const float rre = resid2(Y, mf, low23, p);
          if (!(rre < rr0)) {
            const float4 pk = pold_lds[lane];
            p[0] = pk.x; p[1] = pk.y; p[2] = pk.z; p[3] = pk.w;
          }
          if (lane == 0) *dt_prev_lds = dt;
        }
        J = 0;
        for (;;) {
          const float rr = resid2(Y, mf, low23, p);
          if (rr <= tau2 * yy || J >= sc.max_sweeps) break;
          const int kk = sc.check_every < sc.max_sweeps - J ? sc.check_every : sc.max_sweeps - J;
          jacobi_sweeps(Y, rsm, low23, p, kk);
          J += kk;
        }
      } else {
        J = aux_pass ? sc.j_slow : sc.j_fast;
        jacobi_sweeps(Y, rsm, low23, p, J);
      }
      n_sweeps += J;
      UNROLL for (int j = 0; j < 4; ++j) na[j] = (p[j] + G[j]) * rmass[j];
      {
        const float p0 = readlane(p[0], 0);
        const float NN1 = neg_a2 * ow_nv;
        const float NN2 = Gme * s + p0;
        const float NN3 = neg_a1 * ow_nd;
        ow_na = ((NN1 + NN2) + NN3) / sigma_ow;
      }
      UNROLL for (int j = 0; j < 4; ++j)
        npsid[j] = (eta1 * od[j] + eta2 * (nxis[j] - nv[j])) + neg_w * npsi[j];

      if (reset == 0) {
        float ex[4], ey[4], ez[4];
        UNROLL for (int j = 0; j < 4; ++j) {
          ex[j] = __builtin_fabsf(aref[j] - na[j]);
          ey[j] = __builtin_fabsf(nv[j] - vref[j]);
          ez[j] = ey[j] - fmaxf(mtt, nv[j] * dt);
        }
        float emax = tree_max(ex[0], ex[1], ex[2], ex[3]);
        float esum = tree_sum(ey[0], ey[1], ey[2], ey[3]);
        float fsum = tree_sum(ez[0], ez[1], ez[2], ez[3]);
        if (sc.decision_mode == 1) { emax = 0.f; esum = 0.f; fsum = 0.f; }
        float L;
        if (esum > mmspt) { L = emax * dt; L = L / esum; } else L = 0.0f;
        if (L > 6.0f) {
          dt = dt * 0.5f;
          if (dt < dt_min) dt = dt_min;
          another = 0; reset = 1;
        } else if (fsum > 0.0f) {
          another = 1; reset = 0;
        } else if (L < 0.5f) {
          dt = dt * 2.0f;
          if (dt > dt_max) dt = dt_max;
          another = 0; reset = 1;
        } else {
          another = 1; reset = 0;
        }
      } else {
        another = 1; reset = 0;
      }
      out_loop++;
      if (out_loop >= sc.max_passes) { another = 0; reset = 1; c_fail = 1.f; }
    }
    c_steps += 1.0f;
    c_dt += dt;
    c_passes += (float)out_loop;
    n_steps++;
    n_passes += out_loop;
    if (t > tso) {
      t -= tso;
      sx0 = sx1;
      sx1 = in_at(a, (int64_t)in_s + 2);
      if (in_s + 1 >= it.mosi) {
        const float div_times = c_dt / c_steps;
        const float div_j = c_passes / c_steps;
        if (in_s >= it.write_start && in_s < a.rows) {
          if (a.speeds)
            *reinterpret_cast<float4 *>(a.speeds + (int64_t)in_s * CM_SECTIONS + base) =
                make_float4(nv[0], nv[1], nv[2], nv[3]);
          if (lane == 0) {
            if (a.failed_nodes) a.failed_nodes[in_s] = (int32_t)c_fail;
            if (a.conv_time) a.conv_time[in_s] = div_times;
            if (a.conv_jacobi) a.conv_jacobi[in_s] = div_j;
            if (a.conv_loops) a.conv_loops[in_s] = c_steps;
          }
        }
        if (c_fail > 0.f) fb += 1;
        ctb += div_times / span;
        cjb += div_j / span;
        clb += c_steps / span;
      }
      c_fail = 0.f; c_dt = 0.f; c_steps = 0.f; c_passes = 0.f;
      in_s++;
    }
    UNROLL for (int j = 0; j < 4; ++j) {
      pd[j] = nd[j]; pv[j] = nv[j]; pa[j] = na[j]; ppsi[j] = npsi[j];
      ppsid[j] = npsid[j]; pxis[j] = nxis[j]; pxi[j] = nxi[j];
    }
    ow_pd = ow_nd; ow_pv = ow_nv; ow_pa = ow_na;
    t += dt;
  }
  if (lane == 0) {
    if (a.failed_blocks) a.failed_blocks[item_index] = fb;
    if (a.conv_time_blocks) a.conv_time_blocks[item_index] = ctb;
    if (a.conv_loops_blocks) a.conv_loops_blocks[item_index] = clb;
    if (a.conv_jacobi_blocks) a.conv_jacobi_blocks[item_index] = cjb;
    if (a.stats) {
      a.stats[item_index].steps = n_steps;
      a.stats[item_index].passes = n_passes;
      a.stats[item_index].sweeps = n_sweeps;
    }
  }
}

template <int WAVES, bool TOL, int WS>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(3))) void bm_kernel(BmLaunch a) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int idx = blockIdx.x * WAVES + wave;
  __shared__ float4 pold_lds[WS ? WAVES * 65 : 1];
  if (idx < a.n_items) run_item<TOL, WS>(a, idx, lane, pold_lds + (WS ? wave * 65 : 0));
}


}  // namespace cm

// Launcher used by the C-ABI layer.
extern "C" int cm_launch_bm(const cm::BmLaunch *a, int waves_per_block, int grid_blocks, hipStream_t stream) {
  if (a->use_queue) return -1;   // the work-queue launch is switched off
  dim3 grid(grid_blocks);
  const bool tol = a->sc.jacobi_mode == 1;
  const int ws = tol ? a->sc.warm_start : 0;
  if (ws < 0 || ws > 1) return -1;
#define CM_BM_CASE(W, T, S) \
  case (W) * 4 + (T) * 2 + (S): \
    hipLaunchKernelGGL((cm::bm_kernel<W, T, S>), grid, dim3(64 * (W)), 0, stream, *a); \
    break;
  switch (waves_per_block * 4 + (tol ? 2 : 0) + ws) {
    CM_BM_CASE(1, false, 0)
    CM_BM_CASE(1, true, 0)
    CM_BM_CASE(1, true, 1)
    CM_BM_CASE(2, false, 0)
    CM_BM_CASE(2, true, 0)
    CM_BM_CASE(2, true, 1)
    CM_BM_CASE(4, false, 0)
    CM_BM_CASE(4, true, 0)
    CM_BM_CASE(4, true, 1)
    default: return -1;
  }
#undef CM_BM_CASE
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
