// Learned warm start: kernels of the pressure-guess MLP (769 -> 512 -> 512 ->
// 256, GELU), BF16 operands on MFMA with FP32 accumulation, FP32 master
// weights.  No reference counterpart.
//
// GEMM: C[M x N] (fp32) = op(A) * op(B) (+ bias[N]) with op = identity or
// transpose, A/B bf16 row-major with leading dimensions.  A workgroup of 4
// waves owns a 64 x 64 tile of C (each wave a 32 x 32 quarter = 2 x 2
// v_mfma_f32_16x16x16_bf16 tiles) and walks K in steps of 16 through LDS.
// Out-of-range elements load as zero, so any M, N, K works.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "cm_internal.h"

namespace cm {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
  // round to nearest even; NaN kept quiet
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

constexpr int TM = 64, TN = 64, TK = 16;

__global__ __launch_bounds__(256) void gemm_bf16_kernel(int transA, int transB, int M, int N, int K,
                                                        const uint16_t *__restrict__ A, int lda,
                                                        const uint16_t *__restrict__ B, int ldb, float *__restrict__ C,
                                                        int ldc, const float *__restrict__ bias, float beta) {
  __shared__ uint16_t sA[TM][TK + 2];   // sA[m][k]
  __shared__ uint16_t sB[TK][TN + 2];   // sB[k][n]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += TK) {
    // load A tile: 64 x 16 elements, 4 per thread
    for (int e = tid; e < TM * TK; e += 256) {
      const int mm = e / TK, kk = e % TK;
      const int gm = m0 + mm, gk = k0 + kk;
      uint16_t v = 0;
      if (gm < M && gk < K) v = transA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
      sA[mm][kk] = v;
    }
    for (int e = tid; e < TK * TN; e += 256) {
      const int kk = e / TN, nn = e % TN;
      const int gk = k0 + kk, gn = n0 + nn;
      uint16_t v = 0;
      if (gk < K && gn < N) v = transB ? B[(int64_t)gn * ldb + gk] : B[(int64_t)gk * ldb + gn];
      sB[kk][nn] = v;
    }
    __syncthreads();
    // lane l: A rows (l%16), k = 4*(l/16)..+3 ; B cols (l%16), same k
    const int r = lane & 15, kq = (lane >> 4) * 4;
    bf16x4 a[2], b[2];
    for (int i = 0; i < 2; ++i)
      for (int q = 0; q < 4; ++q) a[i][q] = (short)sA[wm + 16 * i + r][kq + q];
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < 4; ++q) b[j][q] = (short)sB[kq + q][wn + 16 * j + r];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // lane l holds C rows 4*(l/16)..+3, column l%16 of each 16 x 16 tile
  const int cn = lane & 15, cm = (lane >> 4) * 4;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < 4; ++q) {
        const int gm = m0 + wm + 16 * i + cm + q, gn = n0 + wn + 16 * j + cn;
        if (gm < M && gn < N) {
          float v = acc[i][j][q];
          if (bias) v += bias[gn];
          float *dst = C + (int64_t)gm * ldc + gn;
          *dst = beta != 0.f ? beta * *dst + v : v;
        }
      }
}

// GELU (tanh form) with the solver's tanh: g = 0.5 x (1 + tanh(c (x + 0.044715 x^3))).
__device__ __forceinline__ float gelu_tanh_arg(float x) { return 0.7978845608f * (x + 0.044715f * x * x * x); }


__global__ void gelu_fwd_kernel(const float *__restrict__ Z, uint16_t *__restrict__ Gbf, float *__restrict__ Gf,
                                int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = Z[i];
  const float g = 0.5f * x * (1.0f + tanhf(gelu_tanh_arg(x)));
  if (Gbf) Gbf[i] = f32_to_bf16(g);
  if (Gf) Gf[i] = g;
}

// dZ = dG * gelu'(Z), written as bf16 (next GEMM operand) and fp32 (bias grad).
__global__ void gelu_bwd_kernel(const float *__restrict__ dG, const float *__restrict__ Z, uint16_t *__restrict__ dZbf,
                                float *__restrict__ dZf, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = Z[i];
  const float t = tanhf(gelu_tanh_arg(x));
  const float dt = (1.0f - t * t) * 0.7978845608f * (1.0f + 3.0f * 0.044715f * x * x);
  const float d = dG[i] * (0.5f * (1.0f + t) + 0.5f * x * dt);
  if (dZbf) dZbf[i] = f32_to_bf16(d);
  if (dZf) dZf[i] = d;
}

__global__ void to_bf16_kernel(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f32_to_bf16(x[i]);
}

// column sums of dZ [M x N] -> db [N] (bias gradient)
__global__ void colsum_kernel(const float *__restrict__ dZ, float *__restrict__ db, int M, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dZ[(int64_t)m * N + n];
  db[n] = s;
}

// MSE loss gradient: dY = (2/(M*N)) (Y - T), loss partials per row.
__global__ void mse_grad_kernel(const float *__restrict__ Y, const float *__restrict__ T, float *__restrict__ dY,
                                uint16_t *__restrict__ dYbf, float *__restrict__ row_loss, int M, int N) {
  const int m = blockIdx.x;
  float part = 0.f;
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    const int64_t i = (int64_t)m * N + n;
    const float e = Y[i] - T[i];
    const float g = 2.0f * e / ((float)M * (float)N);
    dY[i] = g;
    if (dYbf) dYbf[i] = f32_to_bf16(g);
    part += e * e;
  }
  __shared__ float sh[256];
  sh[threadIdx.x] = part;
  __syncthreads();
  for (int t = blockDim.x >> 1; t >= 1; t >>= 1) {
    if ((int)threadIdx.x < t) sh[threadIdx.x] += sh[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) row_loss[m] = sh[0];
}

// Adam on FP32 master weights; refreshes the BF16 copy.
__global__ void adam_kernel(float *__restrict__ w, uint16_t *__restrict__ wbf, const float *__restrict__ g,
                            float *__restrict__ m1, float *__restrict__ m2, int64_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float upd = lr * (a / bc1) / (sqrtf(b / bc2) + eps);
  const float nw = w[i] - upd;
  w[i] = nw;
  if (wbf) wbf[i] = f32_to_bf16(nw);
}

}  // namespace cm

using namespace cm;

extern "C" {

int cm_gemm_bf16(int transA, int transB, int M, int N, int K, const void *A, int lda, const void *B, int ldb,
                 float *C, int ldc, const float *bias, float beta) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) {
    set_error("cm_gemm_bf16: bad arguments");
    return -3;
  }
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, stream(), transA, transB, M, N, K, (const uint16_t *)A,
                     lda, (const uint16_t *)B, ldb, C, ldc, bias, beta);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cm_gelu_fwd(const float *Z, void *Gbf, float *Gf, int64_t n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gelu_fwd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), Z, (uint16_t *)Gbf,
                     Gf, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cm_gelu_bwd(const float *dG, const float *Z, void *dZbf, float *dZf, int64_t n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gelu_bwd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), dG, Z,
                     (uint16_t *)dZbf, dZf, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cm_to_bf16(const float *x, void *y, int64_t n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), x, (uint16_t *)y, n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cm_colsum(const float *dZ, float *db, int M, int N) {
  hipLaunchKernelGGL(colsum_kernel, dim3((N + 255) / 256), dim3(256), 0, stream(), dZ, db, M, N);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cm_mse_grad(const float *Y, const float *T, float *dY, void *dYbf, float *row_loss, int M, int N) {
  hipLaunchKernelGGL(mse_grad_kernel, dim3(M), dim3(256), 0, stream(), Y, T, dY, (uint16_t *)dYbf, row_loss, M, N);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int cm_adam(float *w, void *wbf, const float *g, float *m1, float *m2, int64_t n, float lr, float b1, float b2,
            float eps, int step, float wd) {
  const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), w, (uint16_t *)wbf, g,
                     m1, m2, n, lr, b1, b2, eps, bc1, bc2, wd);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
