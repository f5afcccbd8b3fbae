// IHC and auditory-nerve rates (lambda) of the reference's RunIHCKernel
// (src/cochlea.cu:3573-3644): AC filter of BM velocity (FIR in
// CudaUnifiedCalcIHC :3431-3506 or IIR in CudaIIRFilter[Decoupled]
// :2840-2908), dS = (BM - AC)^2, box-car DC FIR, PRE_IHC = eta_AC*AC +
// eta_DC*DC, IHC = log10(max(10^IHC_Vector * PRE_IHC, 0) + 1e-20) and
// lambda_k = min(SAT, max(A_k * IHC, spont_k)) with section 0 at spont
// (CudaCalcLambda :3365-3408).
//
// Data layout: rows x 256 sections, section fastest; a workgroup is the 256
// sections of a chunk of rows, so every load and store is a coalesced 1 KB
// row.  Work is described by an interval table: each interval is a run of
// rows whose FIR window follows min((k+1) mod per, taps) with the
// reference's per = L before the final regular division and the analysed
// length after it (src/cochlea.cu:2719-2728).  FIR sums accumulate in FP64
// in tap order, as the reference does, so the CPU oracle agrees bit for bit.
#include <hip/hip_runtime.h>

#include "cm_internal.h"

namespace cm {

__device__ __forceinline__ int fir_window(int64_t k, const cm_ihc_interval &iv, int taps) {
  const int64_t sg = iv.frdp > 1 ? k / (iv.frdp - 1) : 1;
  const int64_t per = sg * iv.tla + (1 - sg) * iv.L;
  if (per <= 0) return 0;
  const int64_t w = (k + 1) % per;
  return (int)(w < taps ? w : taps);
}

constexpr int kRows = 16;   // outputs per thread (register block) = taps per chunk

// Row loaders of the FIR helpers: the value of one section at a row.
struct GlobalRows {   // rows x 256 in global memory, section s
  const float *__restrict__ src;
  int s;
  __device__ __forceinline__ float operator()(int64_t row) const { return src[row * CM_SECTIONS + s]; }
};
constexpr int kRing = 128;   // rows of the dS ring of the fused kernel (a power of two)
struct RingRows {     // this thread's column of an LDS ring, interval-local rows
  const float *ring;
  int stride, col;
  __device__ __forceinline__ float operator()(int64_t row) const {
    return ring[(int)(row & (kRing - 1)) * stride + col];
  }
};

// General register-blocked FIR over the n <= kRows rows row_first.. of one
// section: inputs stream from the newest output row down, so each output
// accumulates fma(b[0],x[k]), fma(b[1],x[k-1]), ... in tap order; a tap
// outside an output's window contributes an exact zero.  No row past the
// block or before row 0 is read.
template <class Rows>
__device__ __forceinline__ void fir_block_general(const Rows &src, int64_t row_first, int n,
                                                  const double *__restrict__ b, int taps, const int (&w)[kRows],
                                                  double (&acc)[kRows]) {
#pragma unroll
  for (int i = 0; i < kRows; ++i) acc[i] = 0.0;
  const int top = n - 1;
  const int span = n + taps - 1;
  for (int jj = 0; jj < span; ++jj) {
    const int64_t row = row_first + top - jj;
    if (row < 0) break;
    const double x = (double)src(row);
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
      const int t = jj - top + i;
      const bool in = t >= 0 && t < w[i];
      const double c = in ? b[t] : 0.0;
      acc[i] = __builtin_fma(c, x, acc[i]);
    }
  }
}

// Steady-state FIR (every output uses all taps): taps in chunks of kRows with
// a 2*kRows-1 row register window; coefficients are uniform per tap and the
// tap array is zero-padded to a whole number of chunks.  Output i takes tap
// t from row row_first + i - t, accumulated in ascending t.  Loads are
// clamped to [0, row_first + n - 1]; clamped rows only meet zero taps or
// outputs past n.  (Requesting the next chunk's rows ahead of the FMAs costs
// the two-pass kernels a wave per SIMD of occupancy: 114 -> 132 VGPRs, 9.1 ->
// 19.2 ms on 2.26 M rows.)
template <class Rows>
__device__ __forceinline__ void fir_block_steady(const Rows &src, int64_t row_first, int n,
                                                 const double *__restrict__ bpad, int chunks, double (&acc)[kRows]) {
#pragma unroll
  for (int i = 0; i < kRows; ++i) acc[i] = 0.0;
  const int64_t hi = row_first + n - 1;
  double X[2 * kRows - 1];   // X[m] = row row_first - 16c - 15 + m
#pragma unroll
  for (int m = kRows - 1; m < 2 * kRows - 1; ++m) {
    int64_t r = row_first - (kRows - 1) + m;
    r = r > hi ? hi : r;
    X[m] = (double)src(r);
  }
  for (int c = 0; c < chunks; ++c) {
#pragma unroll
    for (int m = 0; m < kRows; ++m) {
      int64_t r = row_first - (int64_t)kRows * c - (kRows - 1) + m;
      r = r < 0 ? 0 : r;
      X[m] = (double)src(r);
    }
    const double *bc = bpad + kRows * c;
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      const double coef = bc[u];
#pragma unroll
      for (int i = 0; i < kRows; ++i) acc[i] = __builtin_fma(coef, X[i - u + kRows - 1], acc[i]);
    }
#pragma unroll
    for (int m = 2 * kRows - 2; m >= kRows; --m) X[m] = X[m - kRows];
  }
}

// XCD-aware tile order.  Workgroups are dealt round-robin to the 8 XCDs of
// the MI355X, each with its own L2; consecutive row tiles share the FIR's tap
// window, so the linear workgroup id is remapped for every XCD to walk one
// contiguous range of tiles and find the window rows in its own L2.  Returns
// (tile x, interval y); a bijection on [0, gx * gy).
__device__ __forceinline__ void xcd_tile(int &tx, int &ty) {
  constexpr unsigned kXcd = 8;
  const unsigned gx = gridDim.x, total = gx * gridDim.y;
  const unsigned lid = blockIdx.x + blockIdx.y * gx;
  const unsigned x = lid % kXcd, k = lid / kXcd, per = total / kXcd, rem = total % kXcd;
  const unsigned logical = x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
  tx = (int)(logical % gx);
  ty = (int)(logical / gx);
}

// Window of kRows outputs starting at interval row r0; returns 1 when every
// output of the block uses all taps.
__device__ __forceinline__ int block_windows(const cm_ihc_interval &iv, int64_t r0, int n, int taps, int (&w)[kRows]) {
  int steady = 1;
#pragma unroll
  for (int i = 0; i < kRows; ++i) {
    w[i] = i < n ? fir_window(iv.kbase + r0 + i, iv, taps) : 0;
    if (i < n && w[i] != taps) steady = 0;
  }
  return steady;
}

// AC (FIR, or read back when an IIR pass already wrote it) and dS.
__global__ __launch_bounds__(256) void ac_ds_kernel(const float *__restrict__ bm, const cm_ihc_interval *__restrict__ ivs,
                                                    const double *__restrict__ b, const double *__restrict__ bpad,
                                                    int nb, float coefA, double coefB, int ac_ready,
                                                    float *__restrict__ AC, float *__restrict__ dS) {
  int tx, ty;
  xcd_tile(tx, ty);
  const cm_ihc_interval iv = ivs[ty];
  const int s = threadIdx.x;
  const int64_t r0 = (int64_t)tx * kRows;
  if (r0 >= iv.len) return;
  const int n = iv.len - r0 < kRows ? (int)(iv.len - r0) : kRows;
  float ac[kRows];
  if (ac_ready) {
#pragma unroll
    for (int i = 0; i < kRows; ++i) ac[i] = i < n ? AC[(iv.row0 + r0 + i) * CM_SECTIONS + s] : 0.0f;
  } else {
    int w[kRows];
    double acc[kRows];
    const GlobalRows rows{bm, s};
    if (block_windows(iv, r0, n, nb, w))
      fir_block_steady(rows, iv.row0 + r0, n, bpad, (nb + kRows - 1) / kRows, acc);
    else
      fir_block_general(rows, iv.row0 + r0, n, b, nb, w, acc);
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
      ac[i] = (float)acc[i];
      if (i < n) AC[(iv.row0 + r0 + i) * CM_SECTIONS + s] = ac[i];
    }
  }
#pragma unroll
  for (int i = 0; i < kRows; ++i) {
    if (i < n) {
      const int64_t o = (iv.row0 + r0 + i) * CM_SECTIONS + s;
      const float mid = __builtin_fmaf(bm[o], coefA, (float)((double)ac[i] * coefB));
      dS[o] = mid * mid;
    }
  }
}

// IIR AC filter, one thread per section, serial over the interval
// (DeviceCudaIIRFilter<float,float>, src/cochlea.cu:2742-2772).
__global__ __launch_bounds__(256) void ac_iir_kernel(const float *__restrict__ bm, const cm_ihc_interval *__restrict__ ivs,
                                                     const double *__restrict__ b, const double *__restrict__ a,
                                                     int taps, float *__restrict__ AC) {
  const cm_ihc_interval iv = ivs[blockIdx.x];
  const int s = threadIdx.x;
  const int64_t per = iv.len;
  for (int64_t k = 0; k < iv.len; ++k) {
    const int64_t row = iv.row0 + k;
    const int ob = per > 0 ? (int)(((k + 1) % per) < taps ? ((k + 1) % per) : taps) : 0;
    float y = 0.0f;
    for (int i = 0; i < ob; ++i) y = (float)((double)y + b[i] * (double)bm[(row - i) * CM_SECTIONS + s]);
    for (int i = 0; i < ob - 1; ++i) y = (float)((double)y - a[i + 1] * (double)AC[(row - i - 1) * CM_SECTIONS + s]);
    AC[row * CM_SECTIONS + s] = y;
  }
}

struct LambdaArgs {
  const float *AC;
  const float *dS;
  const cm_ihc_interval *ivs;
  const double *dc;
  const double *dcpad;    // dc zero-padded to a multiple of kRows taps
  int ndc;
  float etaAC;
  double etaDC;
  const double *damage;   // sets x 256
  const float *A;         // 3 x 256
  float spont[3];
  double sat;
  int64_t rows;           // lambda plane stride in rows
  float *lambda;          // 3 x rows x 256
  float *stages;          // optional 5 x rows x 256 (AC, dS, DC, PRE_IHC, IHC)
};

// DC box-car FIR, PRE_IHC, IHC and the three lambdas.
__global__ __launch_bounds__(256) void dc_lambda_kernel(LambdaArgs a) {
  int tx, ty;
  xcd_tile(tx, ty);
  const cm_ihc_interval iv = a.ivs[ty];
  const int s = threadIdx.x;
  const int64_t r0 = (int64_t)tx * kRows;
  if (r0 >= iv.len) return;
  const int n = iv.len - r0 < kRows ? (int)(iv.len - r0) : kRows;
  const double dmg = a.damage[(int64_t)iv.set * CM_SECTIONS + s];
  const double zero_factor = s != 0 ? 1.0 : 0.0;
  double A3[3], sp[3], zo[3];
  for (int j = 0; j < 3; ++j) {
    A3[j] = (double)a.A[j * CM_SECTIONS + s];
    sp[j] = (double)a.spont[j];
    zo[j] = s != 0 ? 0.0 : sp[j];
  }
  int w[kRows];
  double acc[kRows];
  const GlobalRows rows{a.dS, s};
  if (block_windows(iv, r0, n, a.ndc, w))
    fir_block_steady(rows, iv.row0 + r0, n, a.dcpad, (a.ndc + kRows - 1) / kRows, acc);
  else
    fir_block_general(rows, iv.row0 + r0, n, a.dc, a.ndc, w, acc);
  const int64_t plane = a.rows * CM_SECTIONS;
#pragma unroll
  for (int i = 0; i < kRows; ++i) {
    if (i >= n) continue;
    const int64_t o = (iv.row0 + r0 + i) * CM_SECTIONS + s;
    const float dc = (float)acc[i];
    const float ac = a.AC[o];
    const float pre = __builtin_fmaf(ac, a.etaAC, (float)((double)dc * a.etaDC));
    const float ihc = (float)log10(fmax((double)(dmg * (double)pre), 0.0) + 1e-20);
    for (int j = 0; j < 3; ++j) {
      const double base = fmin(a.sat, fmax(A3[j] * (double)ihc, sp[j]));
      a.lambda[j * plane + o] = (float)fma(zero_factor, base, zo[j]);
    }
    if (a.stages) {
      a.stages[0 * plane + o] = ac;
      a.stages[1 * plane + o] = a.dS[o];
      a.stages[2 * plane + o] = dc;
      a.stages[3 * plane + o] = pre;
      a.stages[4 * plane + o] = ihc;
    }
  }
}

// PRE_IHC, IHC and the three lambdas of one node (CudaCalcIHCComposite /
// CudaCalcLambda order of operations).
struct LambdaNode {
  double dmg, zero_factor, A3[3], sp[3], zo[3];
  __device__ __forceinline__ void init(const LambdaArgs &a, int set, int s) {
    dmg = a.damage[(int64_t)set * CM_SECTIONS + s];
    zero_factor = s != 0 ? 1.0 : 0.0;
    for (int j = 0; j < 3; ++j) {
      A3[j] = (double)a.A[j * CM_SECTIONS + s];
      sp[j] = (double)a.spont[j];
      zo[j] = s != 0 ? 0.0 : sp[j];
    }
  }
  __device__ __forceinline__ void store(const LambdaArgs &a, int64_t o, int64_t plane, float ac, float ds,
                                        float dc) const {
    const float pre = __builtin_fmaf(ac, a.etaAC, (float)((double)dc * a.etaDC));
    const float ihc = (float)log10(fmax((double)(dmg * (double)pre), 0.0) + 1e-20);
    for (int j = 0; j < 3; ++j) {
      const double base = fmin(a.sat, fmax(A3[j] * (double)ihc, sp[j]));
      a.lambda[j * plane + o] = (float)fma(zero_factor, base, zo[j]);
    }
    if (a.stages) {
      a.stages[0 * plane + o] = ac;
      a.stages[1 * plane + o] = ds;
      a.stages[2 * plane + o] = dc;
      a.stages[3 * plane + o] = pre;
      a.stages[4 * plane + o] = ihc;
    }
  }
};

// The whole chain in one pass for a FIR AC filter: BM -> AC FIR -> dS -> DC
// box-car FIR -> PRE_IHC -> IHC -> 3 lambdas, with no AC / dS round trip
// through HBM and no workspace.  A workgroup is SW sections of one tile of
// kTile rows of an interval; each thread walks its section down the tile in
// blocks of kRows rows.  dS goes into this thread's column of an LDS ring of
// kRing rows, from which the DC FIR of the same and later blocks reads it;
// no thread reads another's column, so the ring needs no barrier.  The DC
// window reaches up to kHalo rows back, so a tile first computes dS for the
// kHalo rows before it (within its interval: with kbase 0 no window crosses
// the interval start).  The arithmetic is the two-kernel path's, so the
// results are bit-identical.
constexpr int kTile = 1024;                // output rows per workgroup
constexpr int kHalo = 96;                  // dS rows before a tile (DC taps <= 97)
template <int SW>
__global__ __launch_bounds__(SW) void ihc_fused_kernel(const float *__restrict__ bm, const double *__restrict__ b,
                                                      const double *__restrict__ bpad, int nb, float coefA,
                                                      double coefB, LambdaArgs a) {
  __shared__ float ring[kRing * SW];
  constexpr int kGroups = CM_SECTIONS / SW;
  int tx, ty;
  xcd_tile(tx, ty);
  const cm_ihc_interval iv = a.ivs[ty];
  const int tid = threadIdx.x;
  const int s = (tx % kGroups) * SW + tid;
  const int64_t r_begin = (int64_t)(tx / kGroups) * kTile;
  if (r_begin >= iv.len) return;
  const int64_t r_end = r_begin + kTile < iv.len ? r_begin + kTile : iv.len;
  // rows before the first DC input are never read unless with a zero tap;
  // they must hold finite values
  for (int r = 0; r < kRing; ++r) ring[r * SW + tid] = 0.0f;
  LambdaNode node;
  node.init(a, iv.set, s);
  const GlobalRows bmrows{bm, s};
  const RingRows dsrows{ring, SW, tid};
  const int64_t plane = a.rows * CM_SECTIONS;
  const int ac_chunks = (nb + kRows - 1) / kRows, dc_chunks = (a.ndc + kRows - 1) / kRows;
  const int64_t start = r_begin > kHalo ? r_begin - kHalo : 0;
  for (int64_t r0 = start; r0 < r_end; r0 += kRows) {
    const int n = r_end - r0 < kRows ? (int)(r_end - r0) : kRows;
    int w[kRows];
    double acc[kRows];
    float ac[kRows];
    if (block_windows(iv, r0, n, nb, w))
      fir_block_steady(bmrows, iv.row0 + r0, n, bpad, ac_chunks, acc);
    else
      fir_block_general(bmrows, iv.row0 + r0, n, b, nb, w, acc);
#pragma unroll
    for (int i = 0; i < kRows; ++i) ac[i] = (float)acc[i];
    float ds[kRows];
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
      if (i < n) {
        const float mid = __builtin_fmaf(bm[(iv.row0 + r0 + i) * CM_SECTIONS + s], coefA,
                                         (float)((double)ac[i] * coefB));
        ds[i] = mid * mid;
        ring[(int)((r0 + i) & (kRing - 1)) * SW + tid] = ds[i];
      }
    }
    if (r0 + kRows <= r_begin) continue;      // halo block: dS only
    if (block_windows(iv, r0, n, a.ndc, w))
      fir_block_steady(dsrows, r0, n, a.dcpad, dc_chunks, acc);
    else
      fir_block_general(dsrows, r0, n, a.dc, a.ndc, w, acc);
#pragma unroll
    for (int i = 0; i < kRows; ++i)
      if (i < n) node.store(a, (iv.row0 + r0 + i) * CM_SECTIONS + s, plane, ac[i], ds[i], (float)acc[i]);
  }
}

}  // namespace cm

// Runs the whole IHC/lambda chain for one interval table.  d_scratch (2 x
// rows x 256 floats: AC, dS) selects the two-pass kernels; NULL the one-pass
// kernel (FIR AC filter, kbase 0 intervals, DC taps <= 97).
extern "C" int cm_launch_ihc_lambda(const float *d_bm, int64_t rows, const cm_ihc_interval *d_ivs,
                                    const cm_ihc_interval *h_ivs, int n_ivs, const cm_ihc_interval *d_iir_ivs,
                                    int n_iir_ivs, const double *d_ac_b, const double *d_ac_bpad,
                                    const double *d_ac_a, int n_ac, const double *d_dc, const double *d_dcpad, int n_dc, const double *d_damage, const float *d_A,
                                    const float *spont3, float lambda_sat, float eta_ac, float eta_dc, float scale_bm,
                                    float *d_scratch, float *d_lambda, float *d_stages, hipStream_t stream) {
  using namespace cm;
  if (n_ivs <= 0) return 0;
  int64_t max_len = 0;
  for (int i = 0; i < n_ivs; ++i) max_len = h_ivs[i].len > max_len ? h_ivs[i].len : max_len;
  const int64_t gx = (max_len + kRows - 1) / kRows;
  if (gx > 0x7fffffff || n_ivs > 65535) return -3;
  const int iir = d_ac_a != nullptr;
  LambdaArgs a;
  a.ivs = d_ivs;
  a.dc = d_dc;
  a.dcpad = d_dcpad;
  a.ndc = n_dc;
  a.etaAC = eta_ac;
  a.etaDC = (double)eta_dc;
  a.damage = d_damage;
  a.A = d_A;
  for (int j = 0; j < 3; ++j) a.spont[j] = spont3[j];
  a.sat = (double)lambda_sat;
  a.rows = rows;
  a.lambda = d_lambda;
  a.stages = d_stages;
  // Two passes through the AC/dS workspace when the caller gives one (faster:
  // 9.1 against 20.6 ms on 2.26 M rows; the one-pass kernel holds 1.25 waves
  // per SIMD, its dS ring filling the LDS); one pass, with no workspace and
  // 16 instead of 32 B of HBM traffic per node, when it does not.  The one-pass
  // kernel needs every window to start at its interval's first row (kbase 0)
  // and the DC window to fit the tile halo; an IIR AC filter still needs the
  // workspace for AC.
  const bool fused = d_scratch == nullptr;
  if (fused) {
    if (n_dc > kHalo + 1 || iir) return -3;
    for (int i = 0; i < n_ivs; ++i)
      if (h_ivs[i].kbase != 0) return -3;
    constexpr int SW = 64;
    const int64_t tiles = (max_len + kTile - 1) / kTile * (CM_SECTIONS / SW);
    if (tiles > 0x7fffffff) return -3;
    hipLaunchKernelGGL((ihc_fused_kernel<SW>), dim3((unsigned)tiles, n_ivs), dim3(SW), 0, stream, d_bm, d_ac_b,
                       d_ac_bpad, n_ac, scale_bm, -1.0, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  float *AC = d_scratch;
  float *dS = d_scratch + rows * CM_SECTIONS;
  if (iir) {
    hipLaunchKernelGGL(ac_iir_kernel, dim3(n_iir_ivs), dim3(256), 0, stream, d_bm, d_iir_ivs, d_ac_b, d_ac_a, n_ac,
                       AC);
  }
  hipLaunchKernelGGL(ac_ds_kernel, dim3((unsigned)gx, n_ivs), dim3(256), 0, stream, d_bm, d_ivs, d_ac_b, d_ac_bpad,
                     n_ac, scale_bm, -1.0, iir, AC, dS);
  a.AC = AC;
  a.dS = dS;
  hipLaunchKernelGGL(dc_lambda_kernel, dim3((unsigned)gx, n_ivs), dim3(256), 0, stream, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Spontaneous-rate fix-up of lambda loaded for a run that resumes from the
// lambda stage (CUDAFIXJND_Lambda, src/cochlea.cu:4512-4524): per fibre group
// k, buffer = max(spont_k, lambda) when kept, lambda = max(0, lambda - spont_k).
// Elementwise over [3][rows][256]; 16-byte accesses, grid-stride.
namespace cm {
__global__ __launch_bounds__(256) void fix_lambda_kernel(float4 *__restrict__ lam, float4 *__restrict__ buf,
                                                        int64_t plane4, float s0, float s1, float s2) {
  const int64_t n4 = 3 * plane4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const int k = (int)(i / plane4);
    const float s = k == 0 ? s0 : (k == 1 ? s1 : s2);
    const float4 v = lam[i];
    if (buf) buf[i] = make_float4(fmaxf(s, v.x), fmaxf(s, v.y), fmaxf(s, v.z), fmaxf(s, v.w));
    lam[i] = make_float4(fmaxf(0.0f, v.x - s), fmaxf(0.0f, v.y - s), fmaxf(0.0f, v.z - s), fmaxf(0.0f, v.w - s));
  }
}
}  // namespace cm

// lambda and buffer: [3][rows][256] floats (16-byte aligned device memory).
extern "C" int cm_launch_fix_lambda(float *d_lambda, float *d_buffer, int64_t rows, const float *spont3,
                                    hipStream_t stream) {
  if (rows <= 0) return 0;
  const int64_t plane4 = rows * (CM_SECTIONS / 4);
  int64_t blocks = (3 * plane4 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(cm::fix_lambda_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                     reinterpret_cast<float4 *>(d_lambda), reinterpret_cast<float4 *>(d_buffer), plane4, spont3[0],
                     spont3[1], spont3[2]);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
