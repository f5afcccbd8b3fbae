// Input synthesis from JND profiles (src/cochlea.cu:4070-4150):
// x = dA*cos(2*pi*f*(n-head)/Fs)/sig_norm + Wn*(WN[n]-wn_dc)/wn_norm per
// profile (or dA*(Signal[n]-sig_dc)/sig_norm for a given signal), placed
// back to back with IntervalLength samples each, plus the optional
// hearing-aid FIR / IIR transfer functions.
#include <hip/hip_runtime.h>

#include "cm_internal.h"

namespace cm {

__global__ void generate_input_kernel(const cm_jnd_profile *__restrict__ prof, const float *__restrict__ WN,
                                      const float *__restrict__ Signal, float *__restrict__ out, double wn_dc,
                                      double wn_norm, int signal_mode, double sig_dc, double sig_norm,
                                      int start_profile, int n_profiles, int interval_len, int head, float Fs) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n_profiles * interval_len;
  if (idx >= total) return;
  const int p = (int)(idx / interval_len);
  const int n = (int)(idx - (int64_t)p * interval_len);
  const cm_jnd_profile pr = prof[start_profile + p];
  const double time = (double)(n - head) / (double)Fs;
  const float two_pi = 2 * 3.141592653589793f;
  double sum;
  if (signal_mode) {
    sum = pr.dA * ((double)Signal[n] - sig_dc) / sig_norm;
  } else {
    const double cut = (double)two_pi * pr.frequency * time;
    sum = pr.dA * cos(cut) / sig_norm;
  }
  sum = sum + pr.Wn * ((double)WN[n] - wn_dc) / wn_norm;
  out[idx] = (float)sum;
}

// CudaProcessSignalTroughHearingAID (src/cochlea.cu:4105-4125).
__global__ void hearing_aid_fir_kernel(const float *in, float *out, int n_profiles, int interval_len,
                                       const float *h, int hn) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_profiles * interval_len) return;
  const int pos = (int)(idx % interval_len);
  const int back = pos < hn ? pos : hn;
  float acc = 0.0f;
  for (int i = 0; i < back; ++i) acc = __builtin_fmaf(h[i], in[idx - i], acc);
  out[idx] = acc;
}

// CudaProcessSignalTroughHearingAIDIIR (src/cochlea.cu:4127-4150).
__global__ void hearing_aid_iir_kernel(const float *in, float *out, int n_profiles, int interval_len,
                                       const float *h, int hn) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_profiles) return;
  const int64_t s0 = (int64_t)p * interval_len;
  for (int64_t i = s0; i < s0 + interval_len; ++i) {
    const int back = (int)((i - s0 + 1) < hn ? (i - s0 + 1) : hn);
    float acc = in[i];
    for (int j = 1; j < back; ++j) acc = __builtin_fmaf(h[j], in[i - j], acc);
    out[i] = acc;
  }
}

}  // namespace cm

extern "C" int cm_launch_generate_input(const cm_jnd_profile *d_prof, const float *d_WN, const float *d_Signal,
                                        float *d_out, double wn_dc, double wn_norm, int signal_mode, double sig_dc,
                                        double sig_norm, int start_profile, int n_profiles, int interval_len,
                                        int head, float Fs, hipStream_t stream) {
  const int64_t total = (int64_t)n_profiles * interval_len;
  if (total <= 0) return 0;
  const int64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffff) return -3;
  hipLaunchKernelGGL(cm::generate_input_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_prof, d_WN, d_Signal,
                     d_out, wn_dc, wn_norm, signal_mode, sig_dc, sig_norm, start_profile, n_profiles, interval_len,
                     head, Fs);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int cm_launch_hearing_aid(const float *d_in, float *d_tmp, float *d_out, int n_profiles, int interval_len,
                                     const float *d_fir, int fir_n, const float *d_iir, int iir_n,
                                     hipStream_t stream) {
  const int64_t total = (int64_t)n_profiles * interval_len;
  if (total <= 0) return 0;
  hipLaunchKernelGGL(cm::hearing_aid_fir_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, d_in,
                     d_tmp, n_profiles, interval_len, d_fir, fir_n);
  if (d_iir && iir_n > 0) {
    hipLaunchKernelGGL(cm::hearing_aid_iir_kernel, dim3((n_profiles + 63) / 64), dim3(64), 0, stream, d_tmp, d_out,
                       n_profiles, interval_len, d_iir, iir_n);
  } else {
    if (hipMemcpyAsync(d_out, d_tmp, total * sizeof(float), hipMemcpyDeviceToDevice, stream) != hipSuccess) return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
