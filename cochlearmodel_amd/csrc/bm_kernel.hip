// BM/OHC velocity kernels with one wave per time block (4 sections per lane):
// the fused kernel of the reference method and its tolerance-mode variants,
// and the learned warm-start kernel (lockstep groups with MFMA inference).
// The solver's device functions and arithmetic contract are in bm_device.h.
#include <hip/hip_runtime.h>

#include "bm_device.h"

namespace cm {

// One wave per block (4 sections per lane), WAVES blocks per workgroup.
template <int WAVES, bool TOL, int WS, int U>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(3))) void bm_kernel(BmLaunch a) {
  const int lane = threadIdx.x & 63;
  // wave-uniform by construction: readfirstlane lets the compiler keep the
  // item's fields and the whole step control in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int idx = blockIdx.x * WAVES + wave;
  __shared__ float4 pold_lds[WS ? WAVES * 65 : 1];
  Lay<4> L{lane, 0, nullptr, 0};
  if (idx < a.n_items)
    run_item<4, TOL, WS, U>(a, __builtin_amdgcn_readfirstlane(a.order ? a.order[idx] : idx), L,
                            pold_lds + (WS ? wave * 65 : 0));
}

// ------------------------------------------- learned warm start ---------
// A workgroup of G waves integrates G time blocks in lockstep, one step per
// round.  At each step's predictor pass every wave writes its features to
// LDS and the G waves evaluate the 769 -> 512 -> 512 -> 256 GELU MLP together
// on MFMA (v_mfma_f32_16x16x32_bf16, rows = blocks, padded to 16), with the
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
  static_assert(G <= NT, "every wave needs a first tile group");
  // tile counts a wave can have: 3 only when NT > 2G, 2 only when NT > G (the
  // others are not compiled, so the ring and accumulators are sized for what
  // this G can use)
  constexpr bool may3 = NT > 2 * G, may2 = NT > G;
  ws_u32x4 ring[WS_RING];
  const int n0 = may3 && wave + 2 * G < NT ? 3 : (may2 && wave + G < NT ? 2 : 1);
  if (may3 && n0 == 3)
    ws_fill<G, KT, NT, (may3 ? 3 : 1)>(ring, W, lane, wave);
  else if (may2 && n0 == 2)
    ws_fill<G, KT, NT, (may2 ? 2 : 1)>(ring, W, lane, wave);
  else
    ws_fill<G, KT, NT, 1>(ring, W, lane, wave);
  __syncthreads();
  if (may3 && n0 == 3)
    ws_run<G, KT, NT, GELU, (may3 ? 3 : 1)>(ring, A, sa, W, bias, out, so, lane, wave);
  else if (may2 && n0 == 2)
    ws_run<G, KT, NT, GELU, (may2 ? 2 : 1)>(ring, A, sa, W, bias, out, so, lane, wave);
  else
    ws_run<G, KT, NT, GELU, 1>(ring, A, sa, W, bias, out, so, lane, wave);
  if constexpr (NT > 3 * G) {
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
}

// RMS over the 256 sections, for the network's input scales (the hardware
// square root, 1 ulp, is enough there)
__device__ __forceinline__ float rms256(float x0, float x1, float x2, float x3) {
  return __builtin_amdgcn_sqrtf(tree_sum(x0 * x0, x1 * x1, x2 * x2, x3 * x3) * (1.0f / 256.0f)) + 1e-30f;
}

// Section constants of a lockstep wave.  Every block of a launch has the same
// section parameters, so the group keeps one table of them in LDS, one float4
// per lane and row, filled once by wave 0; only the per-set gamma differs per
// wave.  Every row, gamma included, is read from LDS where a pass uses it
// (the compiler keeps a row in registers only through the loops that read
// it, so nothing constant is held across the network evaluation).  Nothing
// is re-read from global memory and 1/mass is not recomputed per round.
struct WsConst {
  float low;
  const CM_LDS float *R, *S, *Q, *Sohc, *Stm, *Rtm, *rmass, *rsm, *mf, *gam;
};
constexpr int WS_TAB_ROWS = 9;   // R, S, Q, S_ohc, S_tm, R_tm, 1/mass, rsm, m

template <int G>
__global__ __launch_bounds__(64 * G) void bm_ws_kernel(BmLaunch a, WsArgs w) {
  __shared__ __align__(16) uint16_t Xs[16 * WS_XS];
  __shared__ __align__(16) uint16_t H1s[16 * WS_HS];
  __shared__ __align__(16) uint16_t H2s[16 * WS_HS];      // also the output tile (16 x WS_OS floats)
  __shared__ float4 pold_lds[G * 65];
  __shared__ float4 keep_lds[G * 3 * 64];                  // predictor G, nxis, Ptm across the network
  __shared__ float4 tab_lds[WS_TAB_ROWS * 64];
  __shared__ float4 gam_lds[G * 64];                       // gamma of each wave's parameter set
  __shared__ int active_lds[2][G];                         // double-buffered by round parity

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int slot = blockIdx.x * G + wave;
  const bool valid = slot < a.n_items;
  const int idx = __builtin_amdgcn_readfirstlane(valid ? (a.order ? a.order[slot] : slot) : slot);   // item of this wave
  float4 *pold = pold_lds + wave * 65;
  float4 *keep = keep_lds + wave * 3 * 64;   // G, nxis, Ptm of the predictor pass
  Lay<4> L{lane, 0, nullptr, 0};             // one wave per block

  // zero the feature tile once (rows >= G stay zero)
  for (int e = threadIdx.x; e < 16 * WS_XS; e += 64 * G) Xs[e] = 0;

  cm_bm_item it;
  WsConst k;
  ItemScal c;
  ItemState s;
  if (valid) {
    // the record is the same for every lane: its fields go to scalar registers
    // (a vector load would leave them, and the addresses formed from them, in
    // VGPRs that the groups of 16 cannot spare)
    const cm_bm_item r = a.items[idx];
    it.start = __builtin_amdgcn_readfirstlane(r.start);
    it.mosi = __builtin_amdgcn_readfirstlane(r.mosi);
    it.mesi = __builtin_amdgcn_readfirstlane(r.mesi);
    it.write_start = __builtin_amdgcn_readfirstlane(r.write_start);
    it.tol = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.tol)));
    it.m1tol = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(r.m1tol)));
    it.set = __builtin_amdgcn_readfirstlane(r.set);
    it.pad = 0;
  } else {
    it.start = 0; it.mosi = 0; it.mesi = 0; it.write_start = 0; it.tol = 0.f; it.m1tol = 0.f; it.set = 0;
    it.pad = 0;
  }
  {
    ItemConst k0;
    item_setup<true>(a, it, L, k0, c, s);
    k.low = k0.low;
    if (wave == 0) {
      tab_lds[0 * 64 + lane] = make_float4(k0.R[0], k0.R[1], k0.R[2], k0.R[3]);
      tab_lds[1 * 64 + lane] = make_float4(k0.S[0], k0.S[1], k0.S[2], k0.S[3]);
      tab_lds[2 * 64 + lane] = make_float4(k0.Q[0], k0.Q[1], k0.Q[2], k0.Q[3]);
      tab_lds[3 * 64 + lane] = make_float4(k0.Sohc[0], k0.Sohc[1], k0.Sohc[2], k0.Sohc[3]);
      tab_lds[4 * 64 + lane] = make_float4(k0.Stm[0], k0.Stm[1], k0.Stm[2], k0.Stm[3]);
      tab_lds[5 * 64 + lane] = make_float4(k0.Rtm[0], k0.Rtm[1], k0.Rtm[2], k0.Rtm[3]);
      tab_lds[6 * 64 + lane] = make_float4(k0.rmass[0], k0.rmass[1], k0.rmass[2], k0.rmass[3]);
      tab_lds[7 * 64 + lane] = make_float4(k0.rsm[0], k0.rsm[1], k0.rsm[2], k0.rsm[3]);
      tab_lds[8 * 64 + lane] = make_float4(k0.mf[0], k0.mf[1], k0.mf[2], k0.mf[3]);
    }
    gam_lds[wave * 64 + lane] = make_float4(k0.gam[0], k0.gam[1], k0.gam[2], k0.gam[3]);
  }
  {
    const CM_LDS float *t = (const CM_LDS float *)(tab_lds) + 4 * lane;
    k.R = t; k.S = t + 256; k.Q = t + 512; k.Sohc = t + 768; k.Stm = t + 1024; k.Rtm = t + 1280;
    k.rmass = t + 1536; k.rsm = t + 1792; k.mf = t + 2048;
    k.gam = (const CM_LDS float *)(gam_lds + wave * 64 + lane);
  }
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
      const float sf = pass_rhs(k, c, s, true, L, od, G4, Y);
      keep[lane] = make_float4(G4[0], G4[1], G4[2], G4[3]);
      keep[64 + lane] = make_float4(s.nxis[0], s.nxis[1], s.nxis[2], s.nxis[3]);
      keep[128 + lane] = make_float4(s.Ptm[0], s.Ptm[1], s.Ptm[2], s.Ptm[3]);
      if (lane == 0) reinterpret_cast<float *>(pold + 64)[1] = sf;
      rr0 = extrapolate_guess(L, k, Y, s.dt, pold, s.p, rre, true);   // s.p = p_lin
      rr_scale = __builtin_amdgcn_sqrtf(rre * inv_256) + 1e-30f;
      if (net_round || w.capture) {
      float r[4];
      resid_vec(L, Y, k, s.p, r);
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
    if (active) {
      // the predictor state again (same inputs, same values), G, nxis and Ptm
      // from LDS, and od and Y formed from them as pass_rhs formed them (same
      // operations on the same values): cheaper than holding any of it in
      // registers across the network
      float od[4], G4[4], Y[4];
      pass_begin(s, true);
      const float sf = reinterpret_cast<const float *>(pold + 64)[1];
      {
        const float4 g = keep[lane], x = keep[64 + lane], t = keep[128 + lane];
        G4[0] = g.x; G4[1] = g.y; G4[2] = g.z; G4[3] = g.w;
        s.nxis[0] = x.x; s.nxis[1] = x.y; s.nxis[2] = x.z; s.nxis[3] = x.w;
        s.Ptm[0] = t.x; s.Ptm[1] = t.y; s.Ptm[2] = t.z; s.Ptm[3] = t.w;
        for (int j = 0; j < 4; ++j) {
          od[j] = s.nxi[j] - s.nd[j];
          Y[j] = (c.dxdx * G4[j]) * k.Q[j];
        }
        const float Y0 = c.dxa0 * (((c.neg_a2 * s.ow_nv) - c.a1 * s.ow_nd) + c.Gme * sf);
        if (lane == 0) Y[0] = Y0;
        if (lane == 63) Y[3] = 0.0f;
      }
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
        const float rrh = resid2(L, Y, k, ph);
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
        const int J = jacobi_tol(L, a.sc, k, Y, yy, s.p, rr_sel);
        s.n_sweeps += J;
        pass_finish(L, k, c, s, od, G4, sf);
        float vref[4], aref[4];   // unused: step control is skipped on the predictor pass
        UNROLL for (int j = 0; j < 4; ++j) { vref[j] = 0.f; aref[j] = 0.f; }
        pass_control(L, a.sc, c, s, vref, aref, out_loop, reset, another);
      }
      // vref/aref of the first corrector pass are the predictor pass's values
      while (another) {
        float vref[4], aref[4];
        UNROLL for (int j = 0; j < 4; ++j) { vref[j] = s.nv[j]; aref[j] = s.na[j]; }
        const bool predictor = out_loop == 0 || reset;
        if (predictor) out_loop = 0;
        pass_begin(s, predictor);
        float od2[4], G2[4], Y2[4];
        const float sf2 = pass_rhs(k, c, s, out_loop < 2, L, od2, G2, Y2);
        const float yy2 = tree_sum(Y2[0] * Y2[0], Y2[1] * Y2[1], Y2[2] * Y2[2], Y2[3] * Y2[3]);
        const int J = jacobi_tol(L, a.sc, k, Y2, yy2, s.p, resid2(L, Y2, k, s.p));
        s.n_sweeps += J;
        pass_finish(L, k, c, s, od2, G2, sf2);
        pass_control(L, a.sc, c, s, vref, aref, out_loop, reset, another);
      }
      step_commit(a, it, c, s, out_loop, L);
    }
  }
  if (valid) item_final(a, s, idx, L);
}

}  // namespace cm

// Launcher used by the C-ABI layer.  sections_per_lane 4: one wave per block,
// waves_per_block (1, 2 or 4) blocks per workgroup of grid_blocks; 2 or 1: one
// block per workgroup of 2 or 4 waves, one workgroup per block (grid_blocks is
// not used).  Jacobi unroll 1, 2 or 4.
extern "C" int cm_launch_bm(const cm::BmLaunch *a, int waves_per_block, int unroll, int sections_per_lane,
                            int grid_blocks, hipStream_t stream) {
  if (a->use_queue) return -1;   // the work-queue launch is switched off
  const bool tol = a->sc.jacobi_mode == 1;
  const int ws = tol ? a->sc.warm_start : 0;
  if (ws < 0 || ws > 1) return -1;
  const int u = unroll == 1 ? 0 : (unroll == 2 ? 1 : (unroll == 4 ? 2 : -1));
  const int w = waves_per_block == 1 ? 0 : (waves_per_block == 2 ? 1 : (waves_per_block == 4 ? 2 : -1));
  if (u < 0 || w < 0) return -1;
  const int mode = tol ? 1 + ws : 0;   // 0 fixed sweeps, 1 tolerance, 2 tolerance + extrapolation
  if (sections_per_lane == 2 || sections_per_lane == 1)
    return cm_launch_bm_split(a, sections_per_lane, unroll, stream);
  if (sections_per_lane != 4) return -1;
  const dim3 grid(grid_blocks);
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

// Lockstep groups of G in {8, 12, 16} blocks.  A round costs the solver work
// of the busiest SIMD (a group's waves are spread over the CU's 4 SIMDs: 2, 3
// or 4 per SIMD) plus the network's weight stream, which is the same for every
// G; so bigger groups stream fewer bytes per block-step but take longer per
// round.  Measured per wave of groups over the chip (the launch time over
// ceil(n / (G x CUs)) on 20480 held-out blocks): G = 8 costs 0.95 and G = 16
// 1.2 times G = 12.  The host picks the G with the smallest
// ceil(n / (G x CUs)) x cost (ties: the smaller G), so small launches fill the
// CUs with small groups and large ones use 16; cm_ws_set_group fixes G
// instead.  Results do not depend on G.
// Returns G (> 0) or a negative error.
extern "C" int cm_launch_bm_ws(const cm::BmLaunch *a, const cm::WsArgs *w, int n_cus, hipStream_t stream) {
  if (a->sc.jacobi_mode != 1 || w->net_every < 1) return -1;
  const int n = a->n_items;
  const int cus = n_cus > 0 ? n_cus : 1;
  const int64_t c8 = 19 * (int64_t)((n + 8 * cus - 1) / (8 * cus));
  const int64_t c12 = 20 * (int64_t)((n + 12 * cus - 1) / (12 * cus));
  const int64_t c16 = 24 * (int64_t)((n + 16 * cus - 1) / (16 * cus));
  const int G = w->group ? w->group : ((c8 <= c12 && c8 <= c16) ? 8 : (c12 <= c16 ? 12 : 16));
  if (G != 8 && G != 12 && G != 16) return -1;
  const dim3 grid((n + G - 1) / G);
  if (G == 8)
    hipLaunchKernelGGL((cm::bm_ws_kernel<8>), grid, dim3(64 * 8), 0, stream, *a, *w);
  else if (G == 12)
    hipLaunchKernelGGL((cm::bm_ws_kernel<12>), grid, dim3(64 * 12), 0, stream, *a, *w);
  else
    hipLaunchKernelGGL((cm::bm_ws_kernel<16>), grid, dim3(64 * 16), 0, stream, *a, *w);
  return hipGetLastError() == hipSuccess ? G : -2;
}
