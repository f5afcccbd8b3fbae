// JND reductions of CudaCalculateJND (src/cochlea.cu:3703-4065) in one
// kernel: per-interval mean rate (GlobalCalculateMeanRate :3741-3771) and the
// Fisher AI / rate terms reduced over the 256 sections (GlobalCalculateJND
// :3773-3944).  One workgroup = 256 sections of one (interval, fiber group);
// time sums run down the section-fastest lambda planes, so every step is a
// coalesced 1 KB read.  The section reduction keeps the reference's halving
// tree (stride 128, 64, ..., 1) so the CPU oracle agrees bit for bit.
#include <hip/hip_runtime.h>

#include "cm_internal.h"

namespace cm {



__device__ __forceinline__ float halving_tree(float *sh, float v, int s) {
  sh[s] = v;
  for (int t = CM_SECTIONS >> 1; t >= 1; t >>= 1) {
    __syncthreads();
    if (s < t) sh[s] = sh[s] + sh[s + t];
  }
  __syncthreads();
  const float r = sh[0];
  __syncthreads();
  return r;
}

// One pass over the interval body of (interval, fiber group) and of its
// reference: both mean rates (FP64 sums in time order, as
// GlobalCalculateMeanRate takes them) and the AI Fisher sum, then the rate
// term and the two section reductions.  The reference's mean rate is summed
// again by every interval that uses it, in the same order, so it is the value
// its own workgroup writes; fusing the mean-rate pass saves one read of every
// interval body.
__global__ __launch_bounds__(256) void jnd_kernel(JndArgs a) {
  __shared__ float sh[CM_SECTIONS];
  const int iv = blockIdx.x, li = blockIdx.y, s = threadIdx.x;
  const int ref = a.ref_index[iv];
  const float dAv = a.dA[iv];
  const int64_t c = ((int64_t)li * a.rows + (int64_t)iv * a.interval_nodes + a.head_nodes) * CM_SECTIONS + s;
  const int64_t r = ((int64_t)li * a.rows + (int64_t)ref * a.interval_nodes + a.head_nodes) * CM_SECTIONS + s;
  const int L = a.length_nodes;
  const int stage = a.buffer ? a.backup_stage : 0;
  // per-node stages 3 (dLambda) and 4 (preFisherAI) at the body's own lambda nodes
  float *node_stage = (stage == 3 || stage == 4) ? a.buffer + c : nullptr;
  const int64_t ms = ((int64_t)li * a.n_intervals + iv) * CM_SECTIONS + s;   // mean-rate layout
  float ai = 0.f, rms = 0.f;
  if (a.calc_ai || a.calc_rms) {
    double acc_c = a.lambda[c], acc_r = a.lambda[r];
    float accv = 0.0f;
    if (a.calc_ai) {
      const float refl = a.lambda[r], cal = a.lambda[c];
      const float dl = dAv > 0 ? (cal - refl) / dAv : 0.0f;
      const float pre = (float)((double)(dl * dl) / (a.Fs * (double)refl));
      accv += pre;
      if (node_stage) node_stage[0] = stage == 3 ? dl : pre;
#pragma unroll 4
      for (int t = 1; t < L; ++t) {
        const float refl = a.lambda[r + (int64_t)t * CM_SECTIONS];
        const float cal = a.lambda[c + (int64_t)t * CM_SECTIONS];
        acc_c = acc_c + cal;
        acc_r = acc_r + refl;
        const float dl = dAv > 0 ? (cal - refl) / dAv : 0.0f;
        const float pre = (float)((double)(dl * dl) / (a.Fs * (double)refl));
        accv += pre;
        if (node_stage) node_stage[(int64_t)t * CM_SECTIONS] = stage == 3 ? dl : pre;
      }
    } else {
#pragma unroll 8
      for (int t = 1; t < L; ++t) {
        acc_c = acc_c + a.lambda[c + (int64_t)t * CM_SECTIONS];
        acc_r = acc_r + a.lambda[r + (int64_t)t * CM_SECTIONS];
      }
    }
    const float mr = (float)acc_c / (float)L;
    const float mref = (float)acc_r / (float)L;
    a.mean_rate[ms] = mr;
    const float dMR = (mr - mref) / dAv;
    if (stage == 1 && a.calc_rms) {   // mean over the sections of this interval's mean rate
      const float sum = halving_tree(sh, mr, s);
      if (s == 0) a.buffer[(int64_t)li * a.n_intervals + iv] = sum / (float)CM_SECTIONS;
    }
    if (stage == 2) a.buffer[ms] = dMR;
    if (a.calc_ai) {
      const float rs = 1.0f / __builtin_sqrtf(accv);
      ai = (float)((double)(a.T / a.Tlength) * a.nIHC[s] / (double)(rs * rs));
      if (stage == 5) a.buffer[ms] = ai;
    }
    if (a.calc_rms) {
      const float crlb = 1.0f / __builtin_sqrtf(((a.T / mref) * dMR) * dMR);
      rms = (float)(a.nIHC[s] / (double)(crlb * crlb));
      if (stage == 6) a.buffer[ms] = crlb;
    }
  }
  const int64_t fi = (int64_t)li * a.out_stride + a.out_index[iv];
  if (a.calc_ai) {
    const float sum = halving_tree(sh, ai, s);
    if (s == 0) a.FisherAISum[fi] = sum * a.W[li];
  }
  if (a.calc_rms) {
    const float sum = halving_tree(sh, rms, s);
    if (s == 0) a.F_RA[fi] = sum * a.W[li];
  }
}

// cudaCalculateDA (src/cochlea.cu:3726-3739): dA = sqrt(max x^2) over the
// interval body of a non-generated input.
__global__ void calc_dA_kernel(const float *input, float *dA, int n, int interval_nodes, int head_nodes,
                               int length_nodes, int offset_start) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc = 0.0f;
  const int64_t start = (int64_t)offset_start + (int64_t)i * interval_nodes + head_nodes;
  for (int64_t k = start; k < start + length_nodes; ++k) {
    float c = input[k];
    c = c * c;
    if (acc < c) acc = c;
  }
  dA[i] = __builtin_sqrtf(acc);
}

}  // namespace cm

extern "C" int cm_launch_jnd(const cm::JndArgs *a, hipStream_t stream) {
  using namespace cm;
  if (a->n_intervals <= 0) return 0;
  if (a->n_intervals > 0x7fffffff) return -3;
  hipLaunchKernelGGL(jnd_kernel, dim3(a->n_intervals, 3), dim3(256), 0, stream, *a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int cm_launch_calc_dA(const float *input, float *dA, int n, int interval_nodes, int head_nodes,
                                 int length_nodes, int offset_start, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(cm::calc_dA_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, input, dA, n, interval_nodes,
                     head_nodes, length_nodes, offset_start);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
