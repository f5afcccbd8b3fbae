// Learned warm start of the Jacobi pressure solve: model upload (BF16 weights
// packed in MFMA fragment order), capture of training records from the solver
// and exact training targets.  No reference counterpart.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cm_internal.h"

namespace cm {

namespace {

struct WsHolder {
  uint4 *w[3] = {nullptr, nullptr, nullptr};
  float *b[3] = {nullptr, nullptr, nullptr};
  float c_out = 0.f;
  bool loaded = false;
  float *records = nullptr;   // capture target (caller-owned device memory)
  int64_t n_records = 0;
  int stride = 0, per_item = 0;
  int net_every = 1;          // network on every net_every-th lockstep round
  int group = 0;              // lockstep group size 8, 12 or 16; 0: from the launch size
  double *ainv = nullptr;     // 256 x 256 inverse of the kernel-convention matrix (targets)
};
WsHolder g_ws;

// layer shapes: K (inputs), N (outputs); K padded to a multiple of 16
constexpr int kIn[3] = {CM_WS_IN, 512, 512};
constexpr int kOut[3] = {512, 512, 256};

uint16_t host_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// W row-major [K][N] -> [KT/2][NT][64 lanes] x 8 bf16, the B operand of one
// K = 32 MFMA per k-tile pair: lane l of pair j, tile t holds
// W[32 j + 8 (l / 16) + e][16 t + l % 16] for e = 0..7; zero past K.
std::vector<uint16_t> pack_layer(const float *W, int K, int N) {
  const int KT = (K + 15) / 16, KP = (KT + 1) / 2, NT = N / 16;
  std::vector<uint16_t> out((size_t)KP * NT * 64 * 8, 0);
  for (int kp = 0; kp < KP; ++kp)
    for (int nt = 0; nt < NT; ++nt)
      for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 8; ++e) {
          const int k = 32 * kp + 8 * (l / 16) + e, n = 16 * nt + l % 16;
          out[(((size_t)kp * NT + nt) * 64 + l) * 8 + e] = k < K ? host_bf16(W[(size_t)k * N + n]) : 0;
        }
  return out;
}

void free_model() {
  for (int i = 0; i < 3; ++i) {
    if (g_ws.w[i]) (void)hipFree(g_ws.w[i]);
    if (g_ws.b[i]) (void)hipFree(g_ws.b[i]);
    g_ws.w[i] = nullptr;
    g_ws.b[i] = nullptr;
  }
  g_ws.loaded = false;
}

// p* = A^{-1} Y in FP64 for each record; target = (p* - p_lin) / (rms(r_lin) c_out)
__global__ void ws_targets_kernel(const float *__restrict__ rec, int64_t n, const double *__restrict__ ainv,
                                  float c_out, float *__restrict__ out) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n * 256) return;
  const int64_t r = gid >> 8;
  const int i = (int)(gid & 255);
  const float *R = rec + r * CM_WS_RECORD;
  const double *row = ainv + (size_t)i * 256;
  double acc = 0.0;
  for (int k = 0; k < 256; ++k) acc = __builtin_fma(row[k], (double)R[CM_WS_REC_Y + k], acc);
  const double scale = (double)R[769] * (double)c_out;
  out[r * 256 + i] = (float)((acc - (double)R[CM_WS_REC_PLIN + i]) / scale);
}

}  // namespace

bool ws_wants_group_kernel(const cm_bm_scalars &sc) {
  return sc.jacobi_mode == 1 && (sc.warm_start == 2 || (sc.warm_start == 1 && g_ws.records != nullptr));
}


int ws_args(const cm_bm_scalars &sc, int n_items, WsArgs *out) {
  std::memset(out, 0, sizeof(*out));
  if (sc.warm_start == 2) {
    if (!g_ws.loaded) {
      set_error("warm_start=2 needs a model: call cm_ws_load first");
      return -1;
    }
    out->w1 = g_ws.w[0];
    out->w2 = g_ws.w[1];
    out->w3 = g_ws.w[2];
    out->b1 = g_ws.b[0];
    out->b2 = g_ws.b[1];
    out->b3 = g_ws.b[2];
    out->c_out = g_ws.c_out;
    out->use_net = 1;
  }
  out->net_every = g_ws.net_every;
  out->group = g_ws.group;
  if (g_ws.records) {
    if ((int64_t)n_items * g_ws.per_item > g_ws.n_records) {
      set_error("capture buffer holds %lld records, %d items x %d slots need more", (long long)g_ws.n_records,
                n_items, g_ws.per_item);
      return -1;
    }
    out->capture = 1;
    out->records = g_ws.records;
    out->stride = g_ws.stride;
    out->per_item = g_ws.per_item;
  }
  return 0;
}

}  // namespace cm

using namespace cm;

extern "C" {

int cm_ws_load(const float *W1, const float *b1, const float *W2, const float *b2, const float *W3, const float *b3,
               float c_out) {
  const float *W[3] = {W1, W2, W3};
  const float *B[3] = {b1, b2, b3};
  for (int i = 0; i < 3; ++i)
    if (!W[i] || !B[i]) {
      set_error("cm_ws_load: null weight or bias");
      return -1;
    }
  if (!(c_out > 0.f) || !std::isfinite(c_out)) {
    set_error("cm_ws_load: c_out must be positive and finite");
    return -3;
  }
  free_model();
  for (int i = 0; i < 3; ++i) {
    const std::vector<uint16_t> pk = pack_layer(W[i], kIn[i], kOut[i]);
    if (check(hipMalloc((void **)&g_ws.w[i], pk.size() * sizeof(uint16_t)), "hipMalloc ws weights") ||
        check(hipMalloc((void **)&g_ws.b[i], kOut[i] * sizeof(float)), "hipMalloc ws bias") ||
        check(hipMemcpy(g_ws.w[i], pk.data(), pk.size() * sizeof(uint16_t), hipMemcpyHostToDevice), "upload ws") ||
        check(hipMemcpy(g_ws.b[i], B[i], kOut[i] * sizeof(float), hipMemcpyHostToDevice), "upload ws bias")) {
      free_model();
      return -1;
    }
  }
  g_ws.c_out = c_out;
  g_ws.loaded = true;
  return 0;
}

int cm_ws_load_file(const char *path) {
  // float32: W1 (769 x 512), b1, W2 (512 x 512), b2, W3 (512 x 256), b3, c_out
  const size_t n[6] = {(size_t)kIn[0] * kOut[0], (size_t)kOut[0], (size_t)kIn[1] * kOut[1], (size_t)kOut[1],
                       (size_t)kIn[2] * kOut[2], (size_t)kOut[2]};
  size_t total = 1;
  for (size_t x : n) total += x;
  std::vector<float> buf(total);
  FILE *f = path ? std::fopen(path, "rb") : nullptr;
  if (!f) {
    set_error("cm_ws_load_file: cannot open %s", path ? path : "(null)");
    return -1;
  }
  const size_t got = std::fread(buf.data(), sizeof(float), total, f);
  const bool extra = std::fgetc(f) != EOF;
  std::fclose(f);
  if (got != total || extra) {
    set_error("cm_ws_load_file: %s holds %zu floats%s, expected %zu", path, got, extra ? " or more" : "", total);
    return -3;
  }
  const float *p = buf.data();
  const float *parts[6];
  for (int i = 0; i < 6; ++i) {
    parts[i] = p;
    p += n[i];
  }
  return cm_ws_load(parts[0], parts[1], parts[2], parts[3], parts[4], parts[5], *p);
}

// Evaluate the warm-start network on every k-th lockstep round only; the other
// rounds take the better of the previous and the extrapolated pressure, as
// warm_start = 1 does.  k = 1 (default) uses the network on every step.
int cm_ws_set_group(int g) {
  if (g != 0 && g != 8 && g != 12 && g != 16) {
    set_error("cm_ws_set_group: 0 (from the launch size), 8, 12 or 16 (got %d)", g);
    return -3;
  }
  g_ws.group = g;
  return 0;
}

int cm_ws_set_every(int k) {
  if (k < 1 || k > 1000000) {
    set_error("cm_ws_set_every: k must be >= 1");
    return -3;
  }
  g_ws.net_every = k;
  return 0;
}

int cm_ws_set_capture(float *d_records, int64_t n_records, int stride, int per_item) {
  if (!d_records) {
    g_ws.records = nullptr;
    g_ws.n_records = 0;
    return 0;
  }
  if (stride < 1 || per_item < 1 || n_records < 1) {
    set_error("cm_ws_set_capture: stride, per_item and n_records must be positive");
    return -3;
  }
  g_ws.records = d_records;
  g_ws.n_records = n_records;
  g_ws.stride = stride;
  g_ws.per_item = per_item;
  return 0;
}

int cm_ws_targets(const float *d_sec, const float *d_records, int64_t n, float c_out, float *d_targets) {
  if (!d_sec || !d_records || !d_targets || n < 0 || !(c_out > 0.f)) {
    set_error("cm_ws_targets: bad arguments");
    return -3;
  }
  if (n == 0) return 0;
  // A from the m row of the section table (kernel convention of the solver):
  // row i: l_i p_{i-1} + m_i p_i + p_{i+1}, l_i = 1 below section 254
  std::vector<float> m(CM_SECTIONS);
  if (check(hipMemcpy(m.data(), d_sec + CM_SEC_M * CM_SECTIONS, CM_SECTIONS * sizeof(float), hipMemcpyDeviceToHost),
            "download m row"))
    return -1;
  const int N = CM_SECTIONS;
  std::vector<double> inv((size_t)N * N), cp(N), dp(N), col(N);
  for (int j = 0; j < N; ++j) {   // column j of A^{-1}: Thomas on e_j (diagonal dominance, no pivoting)
    for (int i = 0; i < N; ++i) {
      const double lo = (i > 0 && i < N - 2) ? 1.0 : 0.0;
      const double up = i < N - 1 ? 1.0 : 0.0;
      const double rhs = i == j ? 1.0 : 0.0;
      const double den = (double)m[i] - (i > 0 ? lo * cp[i - 1] : 0.0);
      cp[i] = up / den;
      dp[i] = (rhs - (i > 0 ? lo * dp[i - 1] : 0.0)) / den;
    }
    col[N - 1] = dp[N - 1];
    for (int i = N - 2; i >= 0; --i) col[i] = dp[i] - cp[i] * col[i + 1];
    for (int i = 0; i < N; ++i) inv[(size_t)i * N + j] = col[i];
  }
  if (!g_ws.ainv && check(hipMalloc((void **)&g_ws.ainv, (size_t)N * N * sizeof(double)), "hipMalloc A^-1"))
    return -1;
  if (check(hipMemcpy(g_ws.ainv, inv.data(), (size_t)N * N * sizeof(double), hipMemcpyHostToDevice), "upload A^-1"))
    return -1;
  const int64_t threads = n * 256;
  hipLaunchKernelGGL(ws_targets_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream(), d_records,
                     n, g_ws.ainv, c_out, d_targets);
  return check(hipGetLastError(), "ws_targets_kernel");
}

int cm_ws_release(void) {
  free_model();
  if (g_ws.ainv) (void)hipFree(g_ws.ainv);
  g_ws.ainv = nullptr;
  g_ws.records = nullptr;
  g_ws.n_records = 0;
  return 0;
}

}  // extern "C"
