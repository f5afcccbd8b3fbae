// Device functions of the BM/OHC velocity solver with the semantics of the
// reference's kernel #17 (include/cochlea.cuh:13573-14245, launched from
// src/cochlea.cu:1436-1493), shared by the fused kernels (bm_kernel.hip,
// bm_split_kernel.hip) and the learned warm-start kernel.
//
// gfx950 layout: one wave64 integrates one time block; lane l holds the four
// consecutive sections 4l..4l+3 in registers.  The Jacobi pressure sweep
// exchanges the two lane-boundary pressures with DPP wave shifts, the three
// step-control reductions run as DPP butterflies plus four readlanes, and
// accept/reject is a lane-uniform decision.  No LDS, no barriers.  The same
// functions also run a block over a workgroup of 2 or 4 waves (2 or 1
// sections per lane), exchanging wave-boundary values through LDS (Lay<N>
// below); results are identical in every layout.
//
// Arithmetic contract (shared with the CPU oracle): FP32, every operation
// IEEE-rounded, no contraction (built with -ffp-contract=off), IEEE division
// where the reference uses __fdividef, tanh by cm_tanhf below, and the fixed
// reduction tree (s0+s1)+(s2+s3) -> xor 1,2,4,8 -> (r0+r16)+(r32+r48).
#ifndef CM_BM_DEVICE_H
#define CM_BM_DEVICE_H

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

__device__ __forceinline__ float uniform(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// ------------------------------------------------ section layouts -------
// A block's 256 sections over W = 4 / N waves of one workgroup: lane l of wave
// w holds the N consecutive sections N (64 w + l) .. N (64 w + l) + N - 1.
//
// N = 4 is one wave per block: neighbour pressures move by DPP wave shifts,
// reductions run as DPP butterflies, nothing goes through LDS and there is no
// barrier.  N = 2 and N = 1 spread a block over 2 or 4 waves, so a wave needs
// about a half or a quarter of the registers and more blocks fit on a SIMD.
// The wave-boundary pressures, section 0's pressure and each wave's partial
// reductions then pass through a small LDS area behind a workgroup barrier.
// The area has two halves used in turn, so one barrier per exchange suffices:
// a wave writes a half again only after the next exchange's barrier, which
// every wave reaches after its reads of that half.  Totals leave an exchange
// through readfirstlane, so every branch that depends on them is a scalar
// branch, and every wave of a block forms the same totals bit for bit (the
// tree below) and takes the same branches: the barriers always match.
template <int N>
struct Lay {
  static_assert(N == 1 || N == 2 || N == 4, "1, 2 or 4 sections per lane");
  static constexpr int W = 4 / N;
  int lane;
  int wave;    // wave of the block (0 when W == 1)
  float *xa;   // exchange area: 2 halves x W waves x 4 floats (W > 1)
  int ph;      // half used by the next exchange

  __device__ bool sec0() const { return (W == 1 || wave == 0) && lane == 0; }
  __device__ bool last() const { return (W == 1 || wave == W - 1) && lane == 63; }
  __device__ int base() const { return N * (64 * (W == 1 ? 0 : wave) + lane); }
  // Slots that can hold section 254 or 255, whose lower coupling l_i is 0
  // (l_i = 1 elsewhere): slots 2 and 3 when N = 4, every slot otherwise.
  static constexpr bool has_low(int j) { return N == 4 ? j >= 2 : true; }
  __device__ float low() const {
    if (N == 4) return lane == 63 ? 0.0f : 1.0f;
    return base() + N - 1 >= 254 ? 0.0f : 1.0f;
  }

  // Pressures of the sections just before and after this lane's slots (0 past
  // either end of the block), from every lane's last and first slot.
  __device__ __forceinline__ void nbrs(float last_v, float first_v, float &left, float &right) {
    left = wave_shr1(last_v);
    right = wave_shl1(first_v);
    if constexpr (W > 1) {
      float *h = xa + ph * 4 * W;
      if (lane == 0) h[4 * wave] = first_v;
      if (lane == 63) h[4 * wave + 1] = last_v;
      __syncthreads();
      const float from_prev = wave > 0 ? h[4 * (wave - 1) + 1] : 0.0f;
      const float from_next = wave < W - 1 ? h[4 * (wave + 1)] : 0.0f;
      if (lane == 0) left = from_prev;
      if (lane == 63) right = from_next;
      ph ^= 1;
    }
  }

  // This wave's part of a reduction (N < 4), the same in every lane.  The
  // one-wave tree is (s0 + s1) + (s2 + s3) over each lane's four sections,
  // then xor 1, 2, 4, 8 up to the 64-section sums R_k and (R0 + R1) + (R2 + R3).
  // Here a group of 4 sections spans 2 lanes (N = 2) or a quad (N = 1); each
  // step below adds the same two partial sums as the matching one-wave step
  // (IEEE addition is commutative), ending with R_{2w} + R_{2w+1} (N = 2) or
  // R_w (N = 1).
  __device__ __forceinline__ float part_sum(const float (&x)[N]) const {
    float s = x[0];
    if constexpr (N == 2) s = x[0] + x[1];
    s = s + quad_xor1(s);
    s = s + quad_xor2(s);
    s = s + row_half_mirror(s);   // xor 4 once quads are uniform
    s = s + row_mirror(s);        // xor 8 once octets are uniform
    s = s + row_bcast15(s);       // rows 1 and 3 hold their 16-lane pair's sum
    return readlane(s, 31) + readlane(s, 63);
  }
  __device__ __forceinline__ float part_max(const float (&x)[N]) const {
    float s = x[0];
    if constexpr (N == 2) s = fmaxf(x[0], x[1]);
    s = fmaxf(s, quad_xor1(s));
    s = fmaxf(s, quad_xor2(s));
    s = fmaxf(s, row_half_mirror(s));
    s = fmaxf(s, row_mirror(s));
    s = fmaxf(s, row_bcast15(s));
    return fmaxf(readlane(s, 31), readlane(s, 63));
  }

  // Block totals from the K wave parts in v (bit k of is_max: a maximum, else
  // a sum) through one exchange: W = 2 adds (R0 + R1) + (R2 + R3) as
  // part_0 + part_1, W = 4 as (R0 + R1) + (R2 + R3).
  template <int K>
  __device__ __forceinline__ void combine(float (&v)[K], unsigned is_max) {
    static_assert(K <= 4, "4 values per wave and half");
    float *h = xa + ph * 4 * W;
    if (lane == 0) {
      UNROLL for (int k = 0; k < K; ++k) h[4 * wave + k] = v[k];
    }
    __syncthreads();
    UNROLL for (int k = 0; k < K; ++k) {
      float r[W];
      UNROLL for (int w = 0; w < W; ++w) r[w] = h[4 * w + k];
      float t;
      if constexpr (W == 2) {
        t = (is_max >> k) & 1u ? fmaxf(r[0], r[1]) : r[0] + r[1];
      } else {
        t = (is_max >> k) & 1u ? fmaxf(fmaxf(r[0], r[1]), fmaxf(r[2], r[3])) : (r[0] + r[1]) + (r[2] + r[3]);
      }
      v[k] = uniform(t);
    }
    ph ^= 1;
  }

  __device__ __forceinline__ float sum(const float (&x)[N]) {
    if constexpr (N == 4) {
      return tree_sum(x[0], x[1], x[2], x[3]);
    } else {
      float v[1] = {part_sum(x)};
      combine<1>(v, 0u);
      return v[0];
    }
  }
  // max of ex, sums of ey and ez: the three step-control reductions
  __device__ __forceinline__ void reduce3(const float (&ex)[N], const float (&ey)[N], const float (&ez)[N],
                                          float &emax, float &esum, float &fsum) {
    if constexpr (N == 4) {
      emax = tree_max(ex[0], ex[1], ex[2], ex[3]);
      esum = tree_sum(ey[0], ey[1], ey[2], ey[3]);
      fsum = tree_sum(ez[0], ez[1], ez[2], ez[3]);
    } else {
      float v[3] = {part_max(ex), part_sum(ey), part_sum(ez)};
      combine<3>(v, 1u);
      emax = v[0]; esum = v[1]; fsum = v[2];
    }
  }
  // section 0's value (slot 0 of the lane that holds it) in every lane
  __device__ __forceinline__ float sec0_value(float x0) {
    if constexpr (W == 1) {
      return readlane(x0, 0);
    } else {
      float *h = xa + ph * 4 * W;
      if (wave == 0 && lane == 0) h[0] = x0;
      __syncthreads();
      const float v = uniform(h[0]);
      ph ^= 1;
      return v;
    }
  }
};

template <int N>
__device__ __forceinline__ void ld_slots(const float *p, float (&o)[N]) {
  if constexpr (N == 4) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else if constexpr (N == 2) {
    const float2 v = *reinterpret_cast<const float2 *>(p);
    o[0] = v.x; o[1] = v.y;
  } else {
    o[0] = *p;
  }
}
template <int N>
__device__ __forceinline__ void st_slots(float *p, const float (&v)[N]) {
  if constexpr (N == 4)
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  else if constexpr (N == 2)
    *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
  else
    *p = v[0];
}

// ------------------------------------------------- one time block -------
// Per-lane constants of an item: section parameters of this lane's N sections.
template <int N>
struct ItemConstN {
  float rsm[N], rmass[N], gam[N], R[N], S[N], Q[N], Sohc[N], Stm[N], Rtm[N], mf[N];
  float low;   // lower-coupling multiplier l_i of the Lay<N>::has_low slots
};

// Scalars shared by all lanes (uniform).
struct ItemScal {
  float w_ohc, alpha_r, Ts, al, Gme, a1, sigma_ow, eta1, eta2, dt_min, dt_max, tso, mtt, mmspt;
  float al_over_as, dxdx, dxa0, neg_a1, neg_a2, neg_w, neg_as, span;
};

// Integration state of an item (include/cochlea.cuh:13620-13700 names in comments).
template <int N>
struct ItemStateN {
  float pd[N], pv[N], pa[N], ppsi[N], ppsid[N], pxis[N], pxi[N];    // committed
  float nd[N], nv[N], na[N], npsi[N], npsid[N], nxis[N], nxi[N], Ptm[N];
  float p[N];                                                          // resident pressure
  float ow_pd, ow_pv, ow_pa, ow_nd, ow_nv, ow_na;                      // oval window
  float dt, t, sx0, sx1;
  int in_s;
  float c_fail, c_dt, c_steps, c_passes;                               // sx[11], sx[12], sx[13], sx[15]
  int64_t n_steps, n_passes, n_sweeps;
  int32_t fb;
  float ctb, clb, cjb;
};
using ItemConst = ItemConstN<4>;
using ItemState = ItemStateN<4>;

// The functions below take the per-lane section constants as K: ItemConstN<N>
// (every row in registers) or a type with the same members whose rarely read
// rows are pointers into LDS (the learned kernel's WsConst).

// a_i = l_i p_{i-1} + p_{i+1} of slot j (include/cochlea.cuh:13962-13982).  l_i
// is 0 or 1, so l_i p_{i-1} is exact and fma(l_i, p_{i-1}, p_{i+1}) rounds as
// the separate multiply and add do; the fma is used where both operands are
// in the lane (one instruction), the multiply and add where one comes from a
// DPP shift (which folds into the add).
template <int N>
__device__ __forceinline__ float coupling(int j, float low, float pl, float pr) {
  if (!Lay<N>::has_low(j)) return pl + pr;
  if (j > 0 && j < N - 1) return __builtin_fmaf(low, pl, pr);
  return low * pl + pr;
}

// One Jacobi sweep of the kernel convention, p_i <- (Y_i - a_i) / m_i with the
// reciprocal diagonal.
template <int N, class K>
__device__ __forceinline__ void jacobi_sweep(Lay<N> &L, const float (&Y)[N], const K &k,
                                             const float (&p)[N], float (&q)[N]) {
  float left, right;
  L.nbrs(p[N - 1], p[0], left, right);
  UNROLL for (int j = 0; j < N; ++j) {
    const float pl = j == 0 ? left : p[j - 1];
    const float pr = j == N - 1 ? right : p[j + 1];
    q[j] = (Y[j] - coupling<N>(j, k.low, pl, pr)) * k.rsm[j];
  }
}
// J is wave-uniform: the count goes to a scalar register.  U (1, 2 or 4) is
// the unroll of the sweep loop, a launch knob of the autotuner: with U >= 2
// sweeps come in pairs so the two iterates alternate between register sets
// instead of being copied back each sweep; U = 4 runs two pairs per loop trip.
// Every U performs the same sweeps in the same order, so results are identical.
template <int U = 2, int N, class K>
__device__ __forceinline__ void jacobi_sweeps(Lay<N> &L, const float (&Y)[N], const K &k, float (&p)[N],
                                              int J) {
  static_assert(U == 1 || U == 2 || U == 4, "unroll 1, 2 or 4");
  J = __builtin_amdgcn_readfirstlane(J);
  float q[N];
  if (U == 1) {
    for (int i = 0; i < J; ++i) {
      jacobi_sweep(L, Y, k, p, q);
      UNROLL for (int j = 0; j < N; ++j) p[j] = q[j];
    }
    return;
  }
  int i = 0;
  if (U == 4) {
    for (; i + 4 <= J; i += 4) {
      jacobi_sweep(L, Y, k, p, q);
      jacobi_sweep(L, Y, k, q, p);
      jacobi_sweep(L, Y, k, p, q);
      jacobi_sweep(L, Y, k, q, p);
    }
  }
  for (; i + 2 <= J; i += 2) {
    jacobi_sweep(L, Y, k, p, q);
    jacobi_sweep(L, Y, k, q, p);
  }
  if (J & 1) {
    jacobi_sweep(L, Y, k, p, q);
    UNROLL for (int j = 0; j < N; ++j) p[j] = q[j];
  }
}

// Residual Y - A p of the kernel convention, per lane, and its squared 2-norm
// with the fixed tree.
template <int N, class K>
__device__ __forceinline__ void resid_vec(Lay<N> &L, const float (&Y)[N], const K &k, const float (&p)[N],
                                          float (&r)[N]) {
  float left, right;
  L.nbrs(p[N - 1], p[0], left, right);
  UNROLL for (int j = 0; j < N; ++j) {
    const float pl = j == 0 ? left : p[j - 1];
    const float pr = j == N - 1 ? right : p[j + 1];
    r[j] = Y[j] - (coupling<N>(j, k.low, pl, pr) + k.mf[j] * p[j]);
  }
}
template <int N, class K>
__device__ __forceinline__ float resid2(Lay<N> &L, const float (&Y)[N], const K &k, const float (&p)[N]) {
  float r[N];
  resid_vec(L, Y, k, p, r);
  UNROLL for (int j = 0; j < N; ++j) r[j] = r[j] * r[j];
  return L.sum(r);
}
template <int N>
__device__ __forceinline__ float sum_sq(Lay<N> &L, const float (&x)[N]) {
  float q[N];
  UNROLL for (int j = 0; j < N; ++j) q[j] = x[j] * x[j];
  return L.sum(q);
}

template <bool TOL, int N>
__device__ __forceinline__ void load_consts(const BmLaunch &a, const cm_bm_item &it, const Lay<N> &L,
                                            ItemConstN<N> &k) {
  const int base = L.base();
  float mm[N];
  ld_slots<N>(a.sec + CM_SEC_MASS * CM_SECTIONS + base, mm);
  ld_slots<N>(a.sec + CM_SEC_RSM * CM_SECTIONS + base, k.rsm);
  ld_slots<N>(a.sec + CM_SEC_R * CM_SECTIONS + base, k.R);
  ld_slots<N>(a.sec + CM_SEC_S * CM_SECTIONS + base, k.S);
  ld_slots<N>(a.sec + CM_SEC_Q * CM_SECTIONS + base, k.Q);
  ld_slots<N>(a.sec + CM_SEC_SOHC * CM_SECTIONS + base, k.Sohc);
  ld_slots<N>(a.sec + CM_SEC_STM * CM_SECTIONS + base, k.Stm);
  ld_slots<N>(a.sec + CM_SEC_RTM * CM_SECTIONS + base, k.Rtm);
  ld_slots<N>(a.gamma + (int64_t)it.set * CM_SECTIONS + base, k.gam);
  UNROLL for (int j = 0; j < N; ++j) k.rmass[j] = 1.0f / mm[j];
  if (TOL) {
    ld_slots<N>(a.sec + CM_SEC_M * CM_SECTIONS + base, k.mf);
  } else {
    UNROLL for (int j = 0; j < N; ++j) k.mf[j] = 0.0f;
  }
  k.low = L.low();
}

template <bool TOL, int N>
__device__ __forceinline__ void item_setup(const BmLaunch &a, const cm_bm_item &it, const Lay<N> &L,
                                           ItemConstN<N> &k, ItemScal &c, ItemStateN<N> &s) {
  const cm_bm_scalars &sc = a.sc;
  load_consts<TOL>(a, it, L, k);

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

  UNROLL for (int j = 0; j < N; ++j) {
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
template <int N>
__device__ __forceinline__ void pass_begin(ItemStateN<N> &s, bool predictor) {
  const float dt = s.dt;
  if (predictor) {
    UNROLL for (int j = 0; j < N; ++j) {
      s.nxi[j] = s.pxis[j] * dt + s.pxi[j];
      s.nv[j] = s.pa[j] * dt + s.pv[j];
      s.nd[j] = ((s.pv[j] + s.nv[j]) * 0.5f) * dt + s.pd[j];
      s.npsi[j] = s.ppsid[j] * dt + s.ppsi[j];
    }
    s.ow_nd = s.ow_pv * dt + s.ow_pd;
    s.ow_nv = s.ow_pa * dt + s.ow_pv;
  } else {
    UNROLL for (int j = 0; j < N; ++j) {
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
template <int N, class K>
__device__ __forceinline__ float pass_rhs(const K &k, const ItemScal &c, ItemStateN<N> &s, bool aux_pass,
                                          const Lay<N> &L, float (&od)[N], float (&G)[N], float (&Y)[N]) {
  UNROLL for (int j = 0; j < N; ++j) {
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
  if (L.sec0()) Y[0] = Y0;
  if (L.last()) Y[N - 1] = 0.0f;
  return sf;
}

// Jacobi until |Y - A p|_2 <= tol |Y|_2 (checked every check_every sweeps,
// capped at max_sweeps); returns the sweeps done.  rr is |Y - A p|_2^2 of the
// incoming p (the guess selection has already computed it).
template <int U = 2, int N, class K>
__device__ __forceinline__ int jacobi_tol(Lay<N> &L, const cm_bm_scalars &sc, const K &k,
                                          const float (&Y)[N], float yy, float (&p)[N], float rr) {
  const float tau2 = sc.tol_rel * sc.tol_rel;
  int J = 0;
  for (;;) {
    if (rr <= tau2 * yy || J >= sc.max_sweeps) break;
    const int kk = sc.check_every < sc.max_sweeps - J ? sc.check_every : sc.max_sweeps - J;
    jacobi_sweeps<U>(L, Y, k, p, kk);
    J += kk;
    rr = resid2(L, Y, k, p);
  }
  return J;
}

// Acceleration, oval-window acceleration and psi derivative, :13989-14013.
template <int N, class K>
__device__ __forceinline__ void pass_finish(Lay<N> &L, const K &k, const ItemScal &c, ItemStateN<N> &s,
                                            const float (&od)[N], const float (&G)[N], float sf) {
  UNROLL for (int j = 0; j < N; ++j) s.na[j] = (s.p[j] + G[j]) * k.rmass[j];
  {
    const float p0 = L.sec0_value(s.p[0]);
    const float NN1 = c.neg_a2 * s.ow_nv;
    const float NN2 = c.Gme * sf + p0;
    const float NN3 = c.neg_a1 * s.ow_nd;
    s.ow_na = ((NN1 + NN2) + NN3) / c.sigma_ow;
  }
  UNROLL for (int j = 0; j < N; ++j)
    s.npsid[j] = (c.eta1 * od[j] + c.eta2 * (s.nxis[j] - s.nv[j])) + c.neg_w * s.npsi[j];
}

// Step control (:14021-14088) and the pass counter; Q1 lane-0 semantics.
template <int N>
__device__ __forceinline__ void pass_control(Lay<N> &L, const cm_bm_scalars &sc, const ItemScal &c,
                                             ItemStateN<N> &s, const float (&vref)[N], const float (&aref)[N],
                                             int &out_loop, int &reset, int &another) {
  if (reset == 0) {
    float ex[N], ey[N], ez[N];
    UNROLL for (int j = 0; j < N; ++j) {
      ex[j] = __builtin_fabsf(aref[j] - s.na[j]);
      ey[j] = __builtin_fabsf(s.nv[j] - vref[j]);
      ez[j] = ey[j] - fmaxf(c.mtt, s.nv[j] * s.dt);
    }
    float emax, esum, fsum;
    L.reduce3(ex, ey, ez, emax, esum, fsum);
    if (sc.decision_mode == 1) { emax = 0.f; esum = 0.f; fsum = 0.f; }
    float Lr;
    if (esum > c.mmspt) { Lr = emax * s.dt; Lr = Lr / esum; } else Lr = 0.0f;
    if (Lr > 6.0f) {
      s.dt = s.dt * 0.5f;
      if (s.dt < c.dt_min) s.dt = c.dt_min;
      another = 0; reset = 1;
    } else if (fsum > 0.0f) {
      another = 1; reset = 0;
    } else if (Lr < 0.5f) {
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
template <int N>
__device__ __forceinline__ void step_commit(const BmLaunch &a, const cm_bm_item &it, const ItemScal &c,
                                            ItemStateN<N> &s, int out_loop, const Lay<N> &L) {
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
        if (a.speeds) {
          // (one float4 per lane when N = 4, written out in full: the register
          // allocation of the learned kernel is tight enough that the generic
          // store costs it a spill)
          if constexpr (N == 4)
            *reinterpret_cast<float4 *>(a.speeds + (int64_t)s.in_s * CM_SECTIONS + 4 * L.lane) =
                make_float4(s.nv[0], s.nv[1], s.nv[2], s.nv[3]);
          else
            st_slots<N>(a.speeds + (int64_t)s.in_s * CM_SECTIONS + L.base(), s.nv);
        }
        if (L.sec0()) {
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
  UNROLL for (int j = 0; j < N; ++j) {
    s.pd[j] = s.nd[j]; s.pv[j] = s.nv[j]; s.pa[j] = s.na[j]; s.ppsi[j] = s.npsi[j];
    s.ppsid[j] = s.npsid[j]; s.pxis[j] = s.nxis[j]; s.pxi[j] = s.nxi[j];
  }
  s.ow_pd = s.ow_nd; s.ow_pv = s.ow_nv; s.ow_pa = s.ow_na;
  s.t += s.dt;
}

template <int N>
__device__ __forceinline__ void item_final(const BmLaunch &a, const ItemStateN<N> &s, int item_index,
                                           const Lay<N> &L) {
  if (L.sec0()) {
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
// step's pressure and dt_prev live in LDS (pold[lane], first N floats, and
// pold[64].x, one copy per wave).  Returns the squared residual of the plain
// guess; *rre gets the extrapolation's, and p holds the extrapolation on
// return when keep_lin, else the choice.
template <int N, class K>
__device__ __forceinline__ float extrapolate_guess(Lay<N> &L, const K &k, const float (&Y)[N], float dt,
                                                   float4 *pold, float (&p)[N], float &rre, bool keep_lin) {
  float *dt_prev = reinterpret_cast<float *>(pold + 64);
  float *mine = reinterpret_cast<float *>(pold + L.lane);
  const float rho = dt / *dt_prev;
  const float rr0 = resid2(L, Y, k, p);
  float po[N];
  ld_slots<N>(mine, po);
  st_slots<N>(mine, p);
  UNROLL for (int j = 0; j < N; ++j) p[j] = p[j] + rho * (p[j] - po[j]);
  rre = resid2(L, Y, k, p);
  if (!keep_lin && !(rre < rr0)) ld_slots<N>(mine, p);
  if (L.lane == 0) *dt_prev = dt;
  return rr0;
}

// One time block, include/cochlea.cuh:13620-14232.  TOL selects the
// sweep-to-tolerance Jacobi; otherwise the reference's fixed 20/2 sweeps.
// WS = 1 adds the extrapolated predictor-pass guess (tolerance mode only).
// U is the Jacobi sweep unroll, N the sections per lane.
template <int N, bool TOL, int WS, int U>
__device__ void run_item(const BmLaunch &a, int item_index, Lay<N> &L, float4 *pold) {
  const cm_bm_item it = a.items[item_index];
  ItemConstN<N> k;
  ItemScal c;
  ItemStateN<N> s;
  item_setup<TOL>(a, it, L, k, c, s);
  if (WS == 1) {
    pold[L.lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (L.lane == 0) reinterpret_cast<float *>(pold + 64)[0] = s.dt;
  }

  while (s.in_s < it.mesi) {
    int out_loop = 0, reset = 1, another = 1;
    while (another) {
      float vref[N], aref[N];
      UNROLL for (int j = 0; j < N; ++j) { vref[j] = s.nv[j]; aref[j] = s.na[j]; }
      const bool predictor = out_loop == 0 || reset;
      if (predictor) out_loop = 0;
      pass_begin(s, predictor);
      float od[N], G[N], Y[N];
      const float sf = pass_rhs(k, c, s, out_loop < 2, L, od, G, Y);
      int J;
      if (TOL) {
        const float yy = sum_sq(L, Y);
        float rr;
        if (WS == 1 && out_loop == 0) {
          float rre;
          const float rr0 = extrapolate_guess(L, k, Y, s.dt, pold, s.p, rre, false);
          rr = rre < rr0 ? rre : rr0;   // residual of the guess kept
        } else {
          rr = resid2(L, Y, k, s.p);
        }
        J = jacobi_tol<U>(L, a.sc, k, Y, yy, s.p, rr);
      } else {
        J = out_loop < 2 ? a.sc.j_slow : a.sc.j_fast;
        jacobi_sweeps<U>(L, Y, k, s.p, J);
      }
      s.n_sweeps += J;
      pass_finish(L, k, c, s, od, G, sf);
      pass_control(L, a.sc, c, s, vref, aref, out_loop, reset, another);
    }
    step_commit(a, it, c, s, out_loop, L);
  }
  item_final(a, s, item_index, L);
}

}  // namespace cm

#endif
