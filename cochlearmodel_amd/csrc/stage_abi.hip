// C-ABI of the post-BM stages and of input generation:
//   IHCNewKernel / RunIHCKernel / BMOHCKernel_Copy_Lambda ...  src/cochlea.cu:2417-2628, 3573-3644, 834-837
//   CudaCalculateJND                                           src/cochlea.cu:3946-4065
//   calcJNDFinal (host, FP64)                                  src/AudioGramCreator.cpp:454-508
//   InputProfilesArrayInitialization / uploadProfiles /
//   generateInputFromProfile                                   src/cochlea.cu:2391-2415, 4160-4510
// plus batched entries over device pointers (cm_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cm_internal.h"

namespace cm {

template <typename T> struct Buf {
  T *ptr = nullptr;
  size_t cap = 0;
  int ensure(size_t n) {
    if (n <= cap && ptr) return 0;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    if (n == 0) n = 1;
    if (hipMalloc((void **)&ptr, n * sizeof(T)) != hipSuccess) {
      set_error("hipMalloc of %zu bytes failed", n * sizeof(T));
      return -1;
    }
    cap = n;
    return 0;
  }
  int upload(const T *src, size_t n) {
    if (ensure(n)) return -1;
    return check(hipMemcpy(ptr, src, n * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

// cudaLambdaHolderData / cudaJNDHolder / cudaSignalHolder / cudaModelAihc of
// src/cochlea.cu:63-631.
struct StageHolder {
  Buf<double> ac_b, ac_a, dc, damage, nIHC, ac_bpad, dcpad;
  Buf<float> A, lambda, scratch, buffer1, stages;
  Buf<cm_ihc_interval> ivs, iir_ivs;
  Buf<cm_jnd_profile> profiles;
  Buf<float> WN, Signal, dA, mean_rate, F_RA, FAI, hear_fir, hear_iir, tmp;
  Buf<int32_t> ref_index, out_index;
  std::vector<cm_jnd_profile> host_profiles;
  int n_ac = 0, n_dc = 0, is_iir = 0;
  float spont[3] = {60.f, 3.f, 0.1f};
  float W[3] = {0.61f, 0.23f, 0.16f};
  float sat = 500.f, eta_ac = 1.f, eta_dc = 100.f, scale = 1.f;
  int calc_time = 0, write_time = 0, interval_nodes = 0, time_blocks = 1, decouple = 0;
  int aihc_loaded = 0;
  // lambda nodes the copy buffer holds for the current run (0: not filled,
  // lambda copies then read the lambda planes themselves)
  size_t buffer_valid = 0;
  double wn_dc = 0, wn_norm = 1, sig_dc = 0, sig_norm = 1;
};
StageHolder g_st;

// Device scratch of cm_ihc_lambda_run_noalloc: the interval table and the
// zero-padded AC and DC tap tables, staged in a fixed array of this code
// object instead of allocated memory.  Every GPU holds its own copy for the
// life of the process; host code reaches it only through the per-device
// address from hipGetSymbolAddress, never through the array's name.
constexpr int kNoallocIntervals = 4096;
constexpr int kNoallocAcTaps = 512;    // a multiple of 16 (the FIR chunk)
constexpr int kNoallocDcTaps = 112;    // 97 used at most, padded to 16
constexpr size_t kNoallocIvBytes = (size_t)kNoallocIntervals * sizeof(cm_ihc_interval);
constexpr size_t kNoallocBytes =
    kNoallocIvBytes + (size_t)kNoallocAcTaps * sizeof(double) + (size_t)kNoallocDcTaps * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kMaxDevices = 64;
static unsigned char *g_scratch_dev[kMaxDevices];   // per-device address of `scratch`

// The current device's address of the scratch array, looked up once per device.
static unsigned char *noalloc_scratch() {
  int dev = 0;
  if (check(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("cm_ihc_lambda_run_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
    return nullptr;
  }
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    if (check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) return nullptr;
    g_scratch_dev[dev] = static_cast<unsigned char *>(p);
  }
  return g_scratch_dev[dev];
}

// dst[i] = src[i] for i < n, 0 up to npad (a fill kernel rather than a memset)
__global__ void pad_taps_kernel(double *dst, const double *src, int n, int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npad) dst[i] = i < n ? src[i] : 0.0;
}

}  // namespace cm

using namespace cm;

extern "C" {

// include/cochlea_common.h:238-274 -> src/cochlea.cu:2417-2598.
// Nerves_Clusters = {spont[3], (unused)[3], W[3]}.
int IHCNewKernel(const double *IHC_Damage_Factor, const float *Nerves_Clusters, const double *dc_filter,
                 int order_of_dc_filter, const double *ac_b_filter, const double *ac_a_filter, int is_ac_iir_filter,
                 int order_of_ac_filter, int cochlea_sections, int time_blocks, double SPLRefVal,
                 const float *backup_speeds, int backup_speeds_length, int calcTime, int writeTime, int allocateTime,
                 int intervalTimeNodes, int max_backup_nodes_len, int lambda_offset, float Lambda_SAT, float eta_AC,
                 float eta_DC, int first_time, int first_time_for_param_set, int loadedFromHD,
                 int disable_advanced_memory_handling, int review_memory_handling, int asMemoryHandlingOnly,
                 float scaleBMVelocityForLambdaCalculation, int CalculateJNDOnGPU, int maxJNDIntervals,
                 int overlapNodes, int Decouple_Filter, int Show_Run_Time) {
  (void)SPLRefVal; (void)backup_speeds; (void)backup_speeds_length; (void)max_backup_nodes_len;
  (void)first_time; (void)loadedFromHD; (void)disable_advanced_memory_handling; (void)review_memory_handling;
  (void)CalculateJNDOnGPU; (void)maxJNDIntervals; (void)overlapNodes; (void)Show_Run_Time;
  if (cochlea_sections != CM_SECTIONS) {
    set_error("IHCNewKernel: %d sections, only %d supported", cochlea_sections, CM_SECTIONS);
    return -3;
  }
  if (lambda_offset != 0) {
    set_error("IHCNewKernel: lambda_offset must be 0 (AudioGramCreator::lambdaOffset)");
    return -3;
  }
  if (order_of_dc_filter <= 0 || order_of_ac_filter <= 0 || !dc_filter || !ac_b_filter ||
      (is_ac_iir_filter && !ac_a_filter)) {
    set_error("IHCNewKernel: missing AC/DC filter");
    return -3;
  }
  if (calcTime <= 0 || allocateTime < calcTime) {
    set_error("IHCNewKernel: calcTime=%d allocateTime=%d", calcTime, allocateTime);
    return -3;
  }
  g_st.calc_time = calcTime;
  g_st.write_time = writeTime;
  g_st.interval_nodes = intervalTimeNodes;
  g_st.time_blocks = time_blocks;
  g_st.decouple = Decouple_Filter;
  g_st.sat = Lambda_SAT;
  g_st.eta_ac = eta_AC;
  g_st.eta_dc = eta_DC;
  g_st.scale = scaleBMVelocityForLambdaCalculation;
  if (first_time_for_param_set || !g_st.dc.ptr) {
    // the AC b coefficients are scaled by scaleBMVelocityForLambdaCalculation (:2477-2481)
    std::vector<double> b(ac_b_filter, ac_b_filter + order_of_ac_filter);
    for (auto &x : b) x = scaleBMVelocityForLambdaCalculation * x;
    if (g_st.ac_b.upload(b.data(), b.size()) || g_st.dc.upload(dc_filter, order_of_dc_filter)) return -1;
    g_st.is_iir = is_ac_iir_filter ? 1 : 0;
    if (g_st.is_iir && g_st.ac_a.upload(ac_a_filter, order_of_ac_filter)) return -1;
    if (g_st.damage.upload(IHC_Damage_Factor, CM_SECTIONS)) return -1;
    g_st.n_ac = order_of_ac_filter;
    g_st.n_dc = order_of_dc_filter;
    if (Nerves_Clusters) {
      for (int j = 0; j < 3; ++j) {
        g_st.spont[j] = Nerves_Clusters[j];
        g_st.W[j] = Nerves_Clusters[6 + j];
      }
    }
  }
  const size_t nodes = (size_t)allocateTime * CM_SECTIONS;
  if (g_st.lambda.ensure(3 * nodes) || g_st.scratch.ensure(2 * nodes)) return -1;
  if (asMemoryHandlingOnly) return 0;
  return 0;
}

// model_Aihc upload (src/cochlea.cu:111-120, 449-455): 3 x 256 floats.
int loadAihc(const float *Aihc) {
  if (g_st.aihc_loaded) return 0;
  if (g_st.A.upload(Aihc, 3 * CM_SECTIONS)) return -1;
  g_st.aihc_loaded = 1;
  return 0;
}
int enableloadAihc(void) {
  g_st.aihc_loaded = 0;
  return 0;
}

// src/cochlea.cu:2599-2628: the lambda copy buffer for saving.
int allocateBuffer(int size_in_nodes, int Show_Run_Time) {
  (void)Show_Run_Time;
  return g_st.buffer1.ensure((size_t)size_in_nodes);
}
int releaseBuffer(int Show_Run_Time) {
  (void)Show_Run_Time;
  g_st.buffer1.release();
  g_st.buffer_valid = 0;
  return 0;
}

// Zero-padded copy of a device tap array (a whole number of 16-tap chunks).
static int pad_taps(Buf<double> &dst, const double *d_src, int n) {
  const size_t np = (size_t)((n + 15) / 16) * 16;
  if (dst.ensure(np)) return -1;
  if (check(hipMemsetAsync(dst.ptr, 0, np * sizeof(double), stream()), "pad taps")) return -1;
  return check(hipMemcpyAsync(dst.ptr, d_src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, stream()),
               "pad taps");
}

static void build_intervals(int rows, std::vector<cm_ihc_interval> &fir, std::vector<cm_ihc_interval> &iir) {
  const int DF = g_st.decouple;
  const int L = DF > 0 ? g_st.interval_nodes * DF : rows;
  const int frdp = g_st.interval_nodes * g_st.time_blocks;
  fir.assign(1, cm_ihc_interval{0, rows, 0, L, rows, frdp, 0});
  iir.clear();
  if (DF > 0) {
    const int nb = std::max(1, g_st.time_blocks / DF);
    for (int b = 0; b < nb; ++b) {
      int len = L;
      if (b + 1 == nb) len += rows - nb * L;
      iir.push_back(cm_ihc_interval{(int64_t)b * L, len, 0, len, len, frdp, 0});
    }
  } else {
    iir.push_back(cm_ihc_interval{0, rows, 0, rows, rows, frdp, 0});
  }
}

// include/cochlea_common.h:276 -> src/cochlea.cu:3573-3644.  backup_stage
// 9..14 copies the AC / DC / dS / IHC / PRE_IHC stage to host_backup.
int RunIHCKernel(float *host_backup, int Show_Run_Time, int save_lambda, int backup_stage,
                 int Decouple_Unified_IHC_Factor) {
  (void)Show_Run_Time; (void)Decouple_Unified_IHC_Factor;
  const int rows = g_st.calc_time;
  if (!g_st.A.ptr || !g_st.dc.ptr) {
    set_error("RunIHCKernel: IHCNewKernel / loadAihc not called");
    return -3;
  }
  if ((size_t)rows * CM_SECTIONS > bm_result_nodes()) {
    set_error("RunIHCKernel: BM buffer holds %zu nodes, %d rows requested", bm_result_nodes(), rows);
    return -3;
  }
  std::vector<cm_ihc_interval> fir, iir;
  build_intervals(rows, fir, iir);
  if (g_st.ivs.upload(fir.data(), fir.size()) || g_st.iir_ivs.upload(iir.data(), iir.size())) return -1;
  const size_t plane = (size_t)rows * CM_SECTIONS;
  const int want_stages = backup_stage >= 9 && backup_stage <= 14 && host_backup;
  if (want_stages && g_st.stages.ensure(5 * plane)) return -1;
  if (g_st.lambda.ensure(3 * plane) || g_st.scratch.ensure(2 * plane)) return -1;
  if (pad_taps(g_st.ac_bpad, g_st.ac_b.ptr, g_st.n_ac) || pad_taps(g_st.dcpad, g_st.dc.ptr, g_st.n_dc)) return -1;
  int rc = cm_launch_ihc_lambda(bm_speeds_ptr(), rows, g_st.ivs.ptr, fir.data(), (int)fir.size(), g_st.iir_ivs.ptr,
                                (int)iir.size(), g_st.ac_b.ptr, g_st.ac_bpad.ptr, g_st.is_iir ? g_st.ac_a.ptr : nullptr,
                                g_st.n_ac, g_st.dc.ptr, g_st.dcpad.ptr, g_st.n_dc, g_st.damage.ptr, g_st.A.ptr, g_st.spont, g_st.sat,
                                g_st.eta_ac, g_st.eta_dc, g_st.scale, g_st.scratch.ptr, g_st.lambda.ptr,
                                want_stages ? g_st.stages.ptr : nullptr, stream());
  if (rc) {
    set_error("RunIHCKernel: launch failed (%d)", rc);
    return -1;
  }
  g_st.buffer_valid = 0;
  if (save_lambda) {
    if (g_st.buffer1.ensure(3 * plane)) return -1;
    if (check(hipMemcpyAsync(g_st.buffer1.ptr, g_st.lambda.ptr, 3 * plane * sizeof(float), hipMemcpyDeviceToDevice,
                             stream()), "save lambda"))
      return -1;
    g_st.buffer_valid = 3 * plane;
  }
  if (want_stages) {
    // AC=9, DC=10, dS=11/Shigh=12, IHC=13, PRE_IHC=14 (src/params.cpp:165-187)
    const int idx = backup_stage == 9 ? 0 : backup_stage == 10 ? 2 : (backup_stage == 11 || backup_stage == 12) ? 1
                    : backup_stage == 13 ? 4 : 3;
    if (check(hipMemcpy(host_backup, g_st.stages.ptr + idx * plane, plane * sizeof(float), hipMemcpyDeviceToHost),
              "stage backup"))
      return -1;
  }
  return check(hipStreamSynchronize(stream()), "RunIHCKernel");
}

// src/cochlea.cu:834-837: copies from the saved-lambda buffer.  When this run
// filled it (RunIHCKernel with save_lambda, or the lambda-stage fix-up with a
// buffer update) the buffer is read, as in the reference; otherwise the
// lambda planes themselves.
int BMOHCKernel_Copy_Lambda(float *target, size_t lambdaNodes, int offset) {
  const bool from_buffer = g_st.buffer_valid > 0;
  const float *src = from_buffer ? g_st.buffer1.ptr : g_st.lambda.ptr;
  const size_t cap = from_buffer ? g_st.buffer_valid : g_st.lambda.cap;
  if (offset < 0 || (size_t)offset + lambdaNodes > cap) {
    set_error("BMOHCKernel_Copy_Lambda: %d+%zu exceeds %zu", offset, lambdaNodes, cap);
    return -3;
  }
  return check(hipMemcpy(target, src + offset, lambdaNodes * sizeof(float), hipMemcpyDeviceToHost), "copy lambda");
}

// ReverseKernel_Copy_Results_Template<JNDFloat> (src/cochlea.cu:3691-3693):
// host rows [start_time_node, start_time_node + time_nodes) of cpu_src to the
// start of cuda_target, which must be one of this library's stage buffers
// (getCudaLambda, getCudaBuffer) or the BM velocity buffer.
int ReverseKernel_Copy_Results_Float(const float *cpu_src, float *cuda_target, size_t start_time_node,
                                     size_t time_nodes, int sections) {
  if (!cpu_src || !cuda_target || sections <= 0) {
    set_error("ReverseKernel_Copy_Results_Float: bad argument");
    return -3;
  }
  const size_t n = time_nodes * (size_t)sections;
  size_t cap = 0;
  if (cuda_target == g_st.lambda.ptr) cap = g_st.lambda.cap;
  else if (cuda_target == g_st.buffer1.ptr) cap = g_st.buffer1.cap;
  else if (cuda_target == bm_speeds_ptr()) cap = bm_result_nodes();
  else {
    set_error("ReverseKernel_Copy_Results_Float: target is not a buffer of this library");
    return -3;
  }
  if (n > cap) {
    set_error("ReverseKernel_Copy_Results_Float: %zu nodes exceed the target's %zu", n, cap);
    return -3;
  }
  return check(hipMemcpy(cuda_target, cpu_src + start_time_node * (size_t)sections, n * sizeof(float),
                         hipMemcpyHostToDevice), "reverse copy");
}

// updateCUDALambdaArray<JNDFloat> (src/cochlea.cu:4526-4540) on the lambda
// planes of calc_time_nodes rows: lambda = max(0, lambda - spont_k) and, when
// cuda_buffer_update, the saved-lambda buffer = max(spont_k, lambda).
int updateCUDALambdaArray(float *lambda_array, float *cuda_buffer, size_t calc_time_nodes, int sections,
                          int Show_Run_Time, int Show_Device_Data, int cuda_buffer_update) {
  (void)Show_Run_Time; (void)Show_Device_Data;
  const size_t n = 3 * calc_time_nodes * CM_SECTIONS;
  if (sections != CM_SECTIONS || lambda_array != g_st.lambda.ptr || n > g_st.lambda.cap) {
    set_error("updateCUDALambdaArray: needs the lambda planes of this library (%d sections, %zu rows)", CM_SECTIONS,
              calc_time_nodes);
    return -3;
  }
  float *buf = nullptr;
  if (cuda_buffer_update) {
    if (cuda_buffer && cuda_buffer != g_st.buffer1.ptr) {
      set_error("updateCUDALambdaArray: buffer is not this library's saved-lambda buffer");
      return -3;
    }
    if (g_st.buffer1.ensure(n)) return -1;
    buf = g_st.buffer1.ptr;
  }
  if (cm_launch_fix_lambda(g_st.lambda.ptr, buf, (int64_t)calc_time_nodes, g_st.spont, stream())) {
    set_error("updateCUDALambdaArray: launch failed");
    return -1;
  }
  g_st.buffer_valid = cuda_buffer_update ? n : 0;
  return check(hipStreamSynchronize(stream()), "updateCUDALambdaArray");
}

// Batched lambda fix-up over device pointers: [3][rows][256] (buffer may be NULL).
int cm_fix_lambda_run(float *d_lambda, float *d_buffer, int64_t rows, const float *spont3) {
  if (!d_lambda || !spont3 || rows < 0) {
    set_error("cm_fix_lambda_run: bad argument");
    return -3;
  }
  if (((uintptr_t)d_lambda | (uintptr_t)d_buffer) & 15) {
    set_error("cm_fix_lambda_run: buffers must be 16-byte aligned");
    return -3;
  }
  return cm_launch_fix_lambda(d_lambda, d_buffer, rows, spont3, stream()) ? -1 : 0;
}

void *getCudaLambda(void) { return g_st.lambda.ptr; }
void *getCudaBuffer(void) { return g_st.buffer1.ptr; }
void *getCudaMeanRate(void) { return g_st.mean_rate.ptr; }

// Batched IHC/lambda over device pointers: an interval table (one interval per
// signal, or the reference layout), per-set IHC damage (sets x 256 doubles).
// d_scratch: 2 x rows x 256 floats.  d_stages may be NULL.
int cm_ihc_lambda_run(const float *d_bm, int64_t rows, const cm_ihc_interval *h_ivs, int n_ivs,
                      const cm_ihc_interval *h_iir_ivs, int n_iir_ivs, const double *d_ac_b, const double *d_ac_a,
                      int n_ac, const double *d_dc, int n_dc, const double *d_damage, const float *d_A,
                      const float *spont3, float lambda_sat, float eta_ac, float eta_dc, float scale_bm,
                      float *d_scratch, float *d_lambda, float *d_stages) {
  if (n_ivs < 0 || n_iir_ivs < 0 || rows < 0) {
    set_error("cm_ihc_lambda_run: negative count (%d intervals, %d IIR intervals, %lld rows)", n_ivs, n_iir_ivs,
              (long long)rows);
    return -3;
  }
  if (n_ivs == 0) return 0;   // an empty batch
  if (!d_bm || !h_ivs || !d_ac_b || !d_dc || !d_damage || !d_A || !spont3 || !d_lambda) {
    set_error("cm_ihc_lambda_run: null argument");
    return -3;
  }
  if (!d_scratch) {   // the one-pass kernel: FIR AC, windows from each interval's first row, DC in the halo
    if (d_ac_a) {
      set_error("cm_ihc_lambda_run: an IIR AC filter needs the workspace");
      return -3;
    }
    if (n_dc > 97) {
      set_error("cm_ihc_lambda_run: %d DC taps need the workspace (one pass holds 97)", n_dc);
      return -3;
    }
    for (int i = 0; i < n_ivs; ++i)
      if (h_ivs[i].kbase != 0) {
        set_error("cm_ihc_lambda_run: interval %d continues a window (kbase %d): it needs the workspace", i,
                  h_ivs[i].kbase);
        return -3;
      }
  }
  for (int i = 0; i < n_ivs; ++i)
    if (h_ivs[i].row0 < 0 || h_ivs[i].row0 + h_ivs[i].len > rows || h_ivs[i].len < 0) {
      set_error("cm_ihc_lambda_run: interval %d outside [0, %lld)", i, (long long)rows);
      return -3;
    }
  if (g_st.ivs.upload(h_ivs, n_ivs)) return -1;
  if (d_ac_a && (!h_iir_ivs || n_iir_ivs <= 0)) {
    set_error("cm_ihc_lambda_run: IIR filter needs its interval table");
    return -3;
  }
  // the IIR pass runs each of its intervals from rest with period len (the
  // kbase / L / tla / frdp fields are not read): rows [row0, row0 + len)
  for (int i = 0; d_ac_a && i < n_iir_ivs; ++i)
    if (h_iir_ivs[i].row0 < 0 || h_iir_ivs[i].len < 0 || h_iir_ivs[i].row0 + h_iir_ivs[i].len > rows) {
      set_error("cm_ihc_lambda_run: IIR interval %d outside [0, %lld)", i, (long long)rows);
      return -3;
    }
  if (d_ac_a && g_st.iir_ivs.upload(h_iir_ivs, n_iir_ivs)) return -1;
  if (pad_taps(g_st.ac_bpad, d_ac_b, n_ac) || pad_taps(g_st.dcpad, d_dc, n_dc)) return -1;
  const int rc = cm_launch_ihc_lambda(d_bm, rows, g_st.ivs.ptr, h_ivs, n_ivs, g_st.iir_ivs.ptr, n_iir_ivs, d_ac_b,
                                      g_st.ac_bpad.ptr, d_ac_a, n_ac, d_dc, g_st.dcpad.ptr, n_dc, d_damage, d_A, spont3, lambda_sat, eta_ac, eta_dc,
                                      scale_bm, d_scratch, d_lambda, d_stages, stream());
  if (rc) set_error("cm_ihc_lambda_run: launch failed (%d)", rc);
  return rc ? -1 : 0;
}

// cm_ihc_lambda_run without allocating: the one-pass kernel (no AC/dS
// workspace) with its tables staged in the device scratch array.  FIR AC
// filters only, at most kNoallocIntervals intervals (kbase 0), kNoallocAcTaps
// AC taps and 97 DC taps.  Calls that overlap on one GPU (other streams or
// threads) race on the scratch array.
int cm_ihc_lambda_run_noalloc(const float *d_bm, int64_t rows, const cm_ihc_interval *h_ivs, int n_ivs,
                              const double *d_ac_b, int n_ac, const double *d_dc, int n_dc, const double *d_damage,
                              const float *d_A, const float *spont3, float lambda_sat, float eta_ac, float eta_dc,
                              float scale_bm, float *d_lambda, float *d_stages) {
  if (n_ivs < 0 || rows < 0) {
    set_error("cm_ihc_lambda_run_noalloc: negative count (%d intervals, %lld rows)", n_ivs, (long long)rows);
    return -3;
  }
  if (n_ivs == 0) return 0;   // an empty batch
  if (!d_bm || !h_ivs || !d_ac_b || !d_dc || !d_damage || !d_A || !spont3 || !d_lambda) {
    set_error("cm_ihc_lambda_run_noalloc: null argument");
    return -3;
  }
  if (n_ivs > kNoallocIntervals || n_ac < 1 || n_ac > kNoallocAcTaps || n_dc < 1 || n_dc > 97) {
    set_error("cm_ihc_lambda_run_noalloc: %d intervals, %d AC and %d DC taps; at most %d, %d and 97", n_ivs, n_ac,
              n_dc, kNoallocIntervals, kNoallocAcTaps);
    return -3;
  }
  for (int i = 0; i < n_ivs; ++i) {
    if (h_ivs[i].row0 < 0 || h_ivs[i].len < 0 || h_ivs[i].row0 + h_ivs[i].len > rows) {
      set_error("cm_ihc_lambda_run_noalloc: interval %d outside [0, %lld)", i, (long long)rows);
      return -3;
    }
    if (h_ivs[i].kbase != 0) {
      set_error("cm_ihc_lambda_run_noalloc: interval %d continues a window (kbase %d)", i, h_ivs[i].kbase);
      return -3;
    }
  }
  unsigned char *base = noalloc_scratch();
  if (!base) return -1;
  cm_ihc_interval *d_ivs = reinterpret_cast<cm_ihc_interval *>(base);
  double *ac_pad = reinterpret_cast<double *>(base + kNoallocIvBytes);
  double *dc_pad = ac_pad + kNoallocAcTaps;
  const int ac_np = (n_ac + 15) / 16 * 16, dc_np = (n_dc + 15) / 16 * 16;
  if (check(hipMemcpyAsync(d_ivs, h_ivs, (size_t)n_ivs * sizeof(cm_ihc_interval), hipMemcpyHostToDevice, stream()),
            "interval table"))
    return -1;
  hipLaunchKernelGGL(pad_taps_kernel, dim3((ac_np + 255) / 256), dim3(256), 0, stream(), ac_pad, d_ac_b, n_ac, ac_np);
  hipLaunchKernelGGL(pad_taps_kernel, dim3(1), dim3(256), 0, stream(), dc_pad, d_dc, n_dc, dc_np);
  if (check(hipGetLastError(), "pad taps")) return -1;
  const int rc = cm_launch_ihc_lambda(d_bm, rows, d_ivs, h_ivs, n_ivs, nullptr, 0, d_ac_b, ac_pad, nullptr, n_ac, d_dc,
                                      dc_pad, n_dc, d_damage, d_A, spont3, lambda_sat, eta_ac, eta_dc, scale_bm,
                                      nullptr, d_lambda, d_stages, stream());
  if (rc) set_error("cm_ihc_lambda_run_noalloc: launch failed (%d)", rc);
  return rc ? -1 : 0;
}

// ---------------------------------------------------------------- JND ---
// Local interval -> reference interval of GlobalCalculateJND (:3828-3855).
// serial[g] of a calculated interval indexes to_ref (n_calc entries).
static int reference_map(int JNDIntervals, int handled, const int *refs, int nrefs, const int *serial,
                         const int *to_ref, int n_calc, std::vector<int32_t> &out) {
  out.resize(JNDIntervals);
  for (int i = 0; i < JNDIntervals; ++i) {
    const int g = i + handled;
    bool is_ref = false;
    for (int r = 0; r < nrefs; ++r)
      if (refs[r] == g) is_ref = true;
    int gref = g;
    if (!is_ref && nrefs > 0) {
      if (serial[g] < 0 || serial[g] >= n_calc) {
        set_error("CudaCalculateJND: serial position %d of interval %d outside %d calculated intervals", serial[g],
                  g, n_calc);
        return -3;
      }
      gref = to_ref[serial[g]];
    }
    const int local = gref - handled;
    if (local < 0 || local >= JNDIntervals) {
      set_error("CudaCalculateJND: interval %d maps to reference %d outside this run", g, gref);
      return -3;
    }
    out[i] = local;
  }
  return 0;
}

// include/cochlea_common.h:292-328 -> src/cochlea.cu:3946-4065.
int CudaCalculateJND(int calculate_ai, int calculate_rms, int mean_size, int fisher_size, double SPLRefVal, double Fs,
                     double scaleBMVelocityForLambdaCalculation, const double *nIHC,
                     const int *JND_Calculated_Intervals, int numOFJNDCalculated, const int *JND_Refrence_Intervals,
                     int numOFJNDReferences, int handeledIntervalsJND, int JNDIntervalsFull, int JNDIntervals,
                     int JNDIntervalHeadNodes, int overlapNodes, int JNDIntervalNodes, int lengthOffset,
                     const int *JND_Serial_Intervals_Positions, const int *JND_Interval_To_Reference, float *F_RA,
                     float *FisherAISum, int writeMatrixSize, int calculatedMatrixSize,
                     double JND_Delta_Alpha_Length_Factor, int JND_USE_Spont_Base_Reference_Lambda,
                     int Show_Run_Time, int calcdA, int show_generated_input_params_cuda, int backup_stage) {
  (void)mean_size; (void)SPLRefVal; (void)scaleBMVelocityForLambdaCalculation; (void)JND_Calculated_Intervals;
  (void)writeMatrixSize; (void)JND_USE_Spont_Base_Reference_Lambda; (void)Show_Run_Time;
  (void)show_generated_input_params_cuda;
  if (JNDIntervals <= 0) return 0;
  if (!g_st.lambda.ptr || calculatedMatrixSize <= 0 || calculatedMatrixSize % CM_SECTIONS) {
    set_error("CudaCalculateJND: no lambda buffer / bad calculatedMatrixSize");
    return -3;
  }
  const int64_t rows = calculatedMatrixSize / CM_SECTIONS;
  // the lambda planes have a stride of `rows` rows: all three must lie in the
  // buffer RunIHCKernel filled
  if ((size_t)3 * rows * CM_SECTIONS > g_st.lambda.cap) {
    set_error("CudaCalculateJND: 3 planes of %lld rows exceed the lambda buffer (%zu nodes)", (long long)rows,
              g_st.lambda.cap);
    return -3;
  }
  if (JNDIntervalNodes <= 0 || (int64_t)JNDIntervals * JNDIntervalNodes > rows) {
    set_error("CudaCalculateJND: %d intervals of %d nodes exceed %lld rows", JNDIntervals, JNDIntervalNodes,
              (long long)rows);
    return -3;
  }
  // each interval body [head, head + length) must lie inside its interval
  if (JNDIntervalHeadNodes < 0 || lengthOffset <= 0 || JNDIntervalHeadNodes + lengthOffset > JNDIntervalNodes) {
    set_error("CudaCalculateJND: body [%d, %d) outside an interval of %d nodes", JNDIntervalHeadNodes,
              JNDIntervalHeadNodes + lengthOffset, JNDIntervalNodes);
    return -3;
  }
  // this run's columns handled .. handled + JNDIntervals - 1 of the Fisher arrays
  if (handeledIntervalsJND < 0 || JNDIntervalsFull < handeledIntervalsJND + JNDIntervals) {
    set_error("CudaCalculateJND: intervals %d + %d exceed JNDIntervalsFull %d", handeledIntervalsJND, JNDIntervals,
              JNDIntervalsFull);
    return -3;
  }
  if (fisher_size < 3 * JNDIntervalsFull) {
    set_error("CudaCalculateJND: fisher_size %d < 3 x %d", fisher_size, JNDIntervalsFull);
    return -3;
  }
  if (numOFJNDReferences < 0 || (numOFJNDReferences > 0 && (!JND_Refrence_Intervals ||
                                                             !JND_Serial_Intervals_Positions ||
                                                             !JND_Interval_To_Reference))) {
    set_error("CudaCalculateJND: reference tables missing");
    return -3;
  }
  if (calcdA && (size_t)overlapNodes + (size_t)(JNDIntervals - 1) * JNDIntervalNodes + JNDIntervalHeadNodes +
                        lengthOffset > bm_input_cap()) {
    set_error("CudaCalculateJND: dA reads past the BM input buffer");
    return -3;
  }
  if (overlapNodes < 0) {
    set_error("CudaCalculateJND: overlapNodes < 0");
    return -3;
  }
  // stage dumps go to the copy buffer (getCudaBuffer), sized for the stage
  size_t stage_nodes = 0;
  if (backup_stage == 1) stage_nodes = (size_t)3 * JNDIntervals;
  else if (backup_stage == 2 || backup_stage == 5 || backup_stage == 6)
    stage_nodes = (size_t)3 * JNDIntervals * CM_SECTIONS;
  else if (backup_stage == 3 || backup_stage == 4)
    stage_nodes = (size_t)3 * rows * CM_SECTIONS;
  if (stage_nodes > g_st.buffer1.cap) {
    if (g_st.buffer1.ensure(stage_nodes)) return -1;
    g_st.buffer_valid = 0;   // a new buffer holds no saved lambda
  }
  std::vector<int32_t> ref;
  int rc = reference_map(JNDIntervals, handeledIntervalsJND, JND_Refrence_Intervals, numOFJNDReferences,
                         JND_Serial_Intervals_Positions, JND_Interval_To_Reference, numOFJNDCalculated, ref);
  if (rc) return rc;
  std::vector<int32_t> outi(JNDIntervals);
  for (int i = 0; i < JNDIntervals; ++i) outi[i] = i + handeledIntervalsJND;
  if (g_st.ref_index.upload(ref.data(), JNDIntervals) || g_st.out_index.upload(outi.data(), JNDIntervals)) return -1;
  if (g_st.nIHC.upload(nIHC, CM_SECTIONS)) return -1;
  if (g_st.dA.ensure(JNDIntervals) || g_st.mean_rate.ensure((size_t)3 * JNDIntervals * CM_SECTIONS) ||
      g_st.F_RA.ensure((size_t)3 * JNDIntervalsFull) || g_st.FAI.ensure((size_t)3 * JNDIntervalsFull))
    return -1;
  if (calcdA) {
    if (cm_launch_calc_dA(bm_input_ptr(), g_st.dA.ptr, JNDIntervals, JNDIntervalNodes, JNDIntervalHeadNodes,
                          lengthOffset, overlapNodes, stream()))
      return -1;
  } else {
    if ((int)g_st.host_profiles.size() < handeledIntervalsJND + JNDIntervals) {
      set_error("CudaCalculateJND: profiles not uploaded");
      return -3;
    }
    std::vector<float> dA(JNDIntervals);
    for (int i = 0; i < JNDIntervals; ++i) dA[i] = (float)g_st.host_profiles[handeledIntervalsJND + i].dA;
    if (check(hipMemcpy(g_st.dA.ptr, dA.data(), JNDIntervals * sizeof(float), hipMemcpyHostToDevice), "dA")) return -1;
  }
  JndArgs a;
  a.lambda = g_st.lambda.ptr;
  a.rows = rows;
  a.n_intervals = JNDIntervals;
  a.interval_nodes = JNDIntervalNodes;
  a.head_nodes = JNDIntervalHeadNodes;
  a.length_nodes = lengthOffset;
  a.dA = g_st.dA.ptr;
  a.ref_index = g_st.ref_index.ptr;
  a.out_index = g_st.out_index.ptr;
  a.out_stride = JNDIntervalsFull;
  a.nIHC = g_st.nIHC.ptr;
  for (int j = 0; j < 3; ++j) a.W[j] = g_st.W[j];
  a.Fs = Fs;
  a.T = (float)(JND_Delta_Alpha_Length_Factor / Fs);
  a.Tlength = (float)((double)(float)lengthOffset / Fs);
  a.calc_ai = calculate_ai;
  a.calc_rms = calculate_rms;
  a.mean_rate = g_st.mean_rate.ptr;
  a.F_RA = g_st.F_RA.ptr;
  a.FisherAISum = g_st.FAI.ptr;
  a.backup_stage = stage_nodes ? backup_stage : 0;
  a.buffer = stage_nodes ? g_st.buffer1.ptr : nullptr;
  if (cm_launch_jnd(&a, stream())) {
    set_error("CudaCalculateJND: launch failed");
    return -1;
  }
  if (check(hipStreamSynchronize(stream()), "JND")) return -1;
  // the reference copies whole fisher arrays back; only this run's columns change
  std::vector<float> tmp((size_t)3 * JNDIntervalsFull);
  if (F_RA && calculate_rms) {
    if (check(hipMemcpy(tmp.data(), g_st.F_RA.ptr, tmp.size() * sizeof(float), hipMemcpyDeviceToHost), "F_RA"))
      return -1;
    for (int l = 0; l < 3; ++l)
      for (int i = 0; i < JNDIntervals; ++i)
        F_RA[l * JNDIntervalsFull + outi[i]] = tmp[l * JNDIntervalsFull + outi[i]];
  }
  if (FisherAISum && calculate_ai) {
    if (check(hipMemcpy(tmp.data(), g_st.FAI.ptr, tmp.size() * sizeof(float), hipMemcpyDeviceToHost), "FAI"))
      return -1;
    for (int l = 0; l < 3; ++l)
      for (int i = 0; i < JNDIntervals; ++i)
        FisherAISum[l * JNDIntervalsFull + outi[i]] = tmp[l * JNDIntervalsFull + outi[i]];
  }
  return 0;
}

// Batched JND over device pointers (lambda planes of `rows` rows).
int cm_jnd_run(const float *d_lambda, int64_t rows, int n_intervals, int interval_nodes, int head_nodes,
               int length_nodes, const float *d_dA, const int32_t *d_ref_index, const int32_t *d_out_index,
               int out_stride, const double *d_nIHC, const float *W3, double Fs, double delta_alpha_factor,
               int calc_ai, int calc_rms, float *d_mean_rate, float *d_F_RA, float *d_FisherAISum) {
  if (!d_lambda || !d_dA || !d_ref_index || !d_out_index || !d_nIHC || !W3 || !d_mean_rate ||
      (calc_rms && !d_F_RA) || (calc_ai && !d_FisherAISum)) {
    set_error("cm_jnd_run: null argument");
    return -3;
  }
  if (n_intervals < 0 || interval_nodes < 0 || head_nodes < 0 || length_nodes < 1 || out_stride < 0 || rows < 0 ||
      !(Fs > 0.0)) {
    set_error("cm_jnd_run: negative or empty size (%d intervals of %d nodes, head %d, length %d, stride %d, Fs %g)",
              n_intervals, interval_nodes, head_nodes, length_nodes, out_stride, Fs);
    return -3;
  }
  if ((int64_t)n_intervals * interval_nodes > rows || (int64_t)head_nodes + length_nodes > interval_nodes) {
    set_error("cm_jnd_run: intervals exceed the lambda planes");
    return -3;
  }
  JndArgs a;
  a.lambda = d_lambda;
  a.rows = rows;
  a.n_intervals = n_intervals;
  a.interval_nodes = interval_nodes;
  a.head_nodes = head_nodes;
  a.length_nodes = length_nodes;
  a.dA = d_dA;
  a.ref_index = d_ref_index;
  a.out_index = d_out_index;
  a.out_stride = out_stride;
  a.nIHC = d_nIHC;
  for (int j = 0; j < 3; ++j) a.W[j] = W3[j];
  a.Fs = Fs;
  a.T = (float)(delta_alpha_factor / Fs);
  a.Tlength = (float)((double)(float)length_nodes / Fs);
  a.calc_ai = calc_ai;
  a.calc_rms = calc_rms;
  a.mean_rate = d_mean_rate;
  a.F_RA = d_F_RA;
  a.FisherAISum = d_FisherAISum;
  a.backup_stage = 0;
  a.buffer = nullptr;
  return cm_launch_jnd(&a, stream()) ? -1 : 0;
}

// calcJNDFinal (src/AudioGramCreator.cpp:454-508) with
// PA2SPLForFinalJNDBounded (include/params.h:331): per non-reference interval,
// JND = 20 log10((sum_k F)^-1/2 / SPLRef), capped at max_db.  Outputs are
// indexed by JND_Serial_Intervals_Positions[g].
int cm_jnd_final(const float *F_RA, const float *FisherAISum, int JNDIntervalsFull, int JNDIntervals,
                 int handled, const int *refs, int nrefs, const int *serial, double SPLRefVal, double max_db,
                 int calc_ai, int calc_rms, double *AiJND, double *RateJND, float *FisherAI, float *JND_RA) {
  const double maxPA = std::pow(10.0, max_db / 20.0) * SPLRefVal;
  auto bounded = [&](double pa) {
    return (std::isinf(pa) || pa > maxPA) ? max_db : 20.0 * std::log10(pa / SPLRefVal);
  };
  for (int i = 0; i < JNDIntervals; ++i) {
    const int g = i + handled;
    bool is_ref = false;
    for (int r = 0; r < nrefs; ++r)
      if (refs[r] == g) is_ref = true;
    if (is_ref) continue;
    double fai = 0.0, frate = 0.0;
    for (int l = 0; l < 3; ++l) {
      if (FisherAISum) fai += FisherAISum[l * JNDIntervalsFull + g];
      if (F_RA) frate += F_RA[l * JNDIntervalsFull + g];
    }
    const int out = serial ? serial[g] : g;
    if (calc_ai) {
      const double s = 1.0 / std::sqrt(fai);
      if (FisherAI) FisherAI[out] = (float)s;
      if (AiJND) AiJND[out] = bounded(s);
    }
    if (calc_rms) {
      const double s = 1.0 / std::sqrt(frate);
      if (JND_RA) JND_RA[out] = (float)s;
      if (RateJND) RateJND[out] = bounded(s);
    }
  }
  return 0;
}

// Failed_Converged_Signals of calcJNDFinal (src/AudioGramCreator.cpp:459-467,
// 490-493): every calculated interval adds the failed-block flags of its
// blocks, Decouple_Filter blocks per interval (all blocks when it is 0), into
// failed_signals[serial[g]].  The caller zeroes failed_signals once; calls for
// successive chunks accumulate, as the reference's member vector does.  Block
// indexes past n_blocks count as 0 (the reference reads past its vector there).
int cm_jnd_failed_signals(const int32_t *failed_blocks, int n_blocks, int JNDIntervals, int handled, const int *refs,
                          int nrefs, const int *serial, int decouple_filter, int32_t *failed_signals) {
  if (!failed_blocks || !failed_signals || n_blocks < 0 || JNDIntervals < 0 || (nrefs > 0 && !refs)) {
    set_error("cm_jnd_failed_signals: bad argument");
    return -3;
  }
  const int64_t coupler = decouple_filter > 0 ? decouple_filter : n_blocks;
  for (int i = 0; i < JNDIntervals; ++i) {
    const int g = i + handled;
    bool is_ref = false;
    for (int r = 0; r < nrefs; ++r)
      if (refs[r] == g) is_ref = true;
    if (is_ref) continue;
    const int out = serial ? serial[g] : g;
    for (int64_t p = 0; p < coupler; ++p) {
      const int64_t b = p + (int64_t)i * coupler;
      if (b < n_blocks) failed_signals[out] += failed_blocks[b];
    }
  }
  return 0;
}

// ---------------------------------------------------------- input gen ---
// src/cochlea.cu:2391-2415.
int InputProfilesArrayInitialization(int maxJNDIntervals, int wn_length, int signal_length, int signal_mode,
                                     int Show_Generated_Input) {
  (void)signal_mode; (void)Show_Generated_Input;
  if (g_st.profiles.ensure(std::max(1, maxJNDIntervals))) return -1;
  if (wn_length > 0 && g_st.WN.ensure(wn_length)) return -1;
  if (signal_length > 0 && g_st.Signal.ensure(signal_length)) return -1;
  return 0;
}
int InputProfilesArrayTermination(void) {
  g_st.profiles.release();
  g_st.WN.release();
  g_st.Signal.release();
  return 0;
}

// src/cochlea.cu:4288-4297.
int uploadProfiles(const cm_jnd_profile *profiles, int numOfProfiles, int profilesLoaded) {
  if (profilesLoaded) return 0;
  g_st.host_profiles.assign(profiles, profiles + numOfProfiles);
  return g_st.profiles.upload(profiles, numOfProfiles);
}

// calculateDCANDNornalizationPostProcess / ...Factor / ...PureTone
// (src/cochlea.cu:4165-4254), on the host in FP64.
static void post_process(int type, double given_interval, int s0, int s1, float Fs, double &f) {
  if (type == 1) f = f / std::sqrt((double)(s1 - s0));
  else if (type == 2 && given_interval > 0) f = f / std::sqrt(Fs * given_interval);
  else if (type == 3) f = 1;
}
static void dc_and_norm(const float *src, int type, double given, double remove_dc, int e0, int e1, int s0, int s1,
                        float Fs, double &dc, double &norm) {
  dc = src[e0];
  for (int i = e0 + 1; i < e1; ++i) dc = dc + src[i];
  dc = remove_dc * dc / (double)(e1 - e0);
  norm = (double)src[s0] * src[s0];
  for (int i = s0 + 1; i < s1; ++i) norm = norm + ((src[i] - dc) * (src[i] - dc));
  norm = std::sqrt(norm);
  post_process(type, given, s0, s1, Fs, norm);
}

int cm_input_normalization(const float *WN, int wn_length, const float *Signal, int signal_length, int signal_mode,
                           int Normalize_Sigma_Type, int Normalize_Sigma_Type_Signal, double given_interval,
                           double remove_dc, int e0, int e1, int s0, int s1, float Fs, double *out4) {
  double wn_dc = 0, wn_norm = 1, sig_dc = 0, sig_norm = 1;
  if (WN && wn_length > 0) {
    if (e1 > wn_length || s1 > wn_length || e0 < 0 || s0 < 0 || e1 <= e0 || s1 <= s0) {
      set_error("cm_input_normalization: accumulation bounds outside the noise vector");
      return -3;
    }
    dc_and_norm(WN, Normalize_Sigma_Type, given_interval, remove_dc, e0, e1, s0, s1, Fs, wn_dc, wn_norm);
  }
  if (signal_mode) {
    if (!Signal || e1 > signal_length || s1 > signal_length) {
      set_error("cm_input_normalization: accumulation bounds outside the signal");
      return -3;
    }
    dc_and_norm(Signal, Normalize_Sigma_Type_Signal, given_interval, remove_dc, e0, e1, s0, s1, Fs, sig_dc, sig_norm);
  } else {
    sig_dc = 0.0;
    sig_norm = std::sqrt(0.5 * (double)(s1 - s0));
    post_process(Normalize_Sigma_Type_Signal, given_interval, s0, s1, Fs, sig_norm);
  }
  out4[0] = wn_dc;
  out4[1] = wn_norm;
  out4[2] = sig_dc;
  out4[3] = sig_norm;
  return 0;
}

// include/cochlea_common.h:373-413 -> src/cochlea.cu:4298-4510.  The input is
// written to the BM input buffer (getCudaInputSamples()); target_input
// receives a copy when Show_Generated_Input & 1.
int generateInputFromProfile(const cm_jnd_profile *profiles, const float *WN, int wn_length, const float *Signal,
                             int signal_length, int signal_mode, int Normalize_Sigma_Type,
                             int Normalize_Sigma_Type_Signal, double Normalize_Noise_Energy_To_Given_Interval,
                             double Remove_Generated_Noise_DC, int start_dc_expected_value_calculation,
                             int end_dc_expected_value_calculation, int start_dc_normalized_value_calculation,
                             int end_dc_normalized_value_calculation, int numOfProfiles, int profilesLoaded,
                             int from_profile_index, int calculatedProfiles, int overlapNodes, int IntervalLength,
                             int JND_Interval_Head, int JND_Interval_Actual_Length, float Fs,
                             int Show_Generated_Input, float *target_input, int Show_Generated_Configuration,
                             int is_first_time_for_parameters_set, float Max_M1_SP_Error_Parameter,
                             float Max_Tolerance_Parameter, int Relative_Error_Parameters, float M1_SP_Fix_Factor,
                             float Tolerance_Fix_Factor, int Blocks_Per_Interval, int Show_Run_Time,
                             const float *fir_transfer_function, int fir_transfer_function_length,
                             const float *iir_transfer_function, int iir_transfer_function_length) {
  (void)overlapNodes; (void)JND_Interval_Actual_Length; (void)Show_Generated_Configuration; (void)Show_Run_Time;
  if (from_profile_index + calculatedProfiles > numOfProfiles || calculatedProfiles < 0 || IntervalLength <= 0) {
    set_error("generateInputFromProfile: profile range outside the table");
    return -3;
  }
  if (wn_length < IntervalLength || (signal_mode && signal_length < IntervalLength)) {
    set_error("generateInputFromProfile: noise/signal shorter than one interval (%d)", IntervalLength);
    return -3;
  }
  if (!profilesLoaded) {
    if (uploadProfiles(profiles, numOfProfiles, 0)) return -1;
    if (g_st.WN.upload(WN, wn_length)) return -1;
    if (signal_mode && g_st.Signal.upload(Signal, signal_length)) return -1;
  }
  if (end_dc_normalized_value_calculation == 0) end_dc_normalized_value_calculation = IntervalLength;
  if (end_dc_expected_value_calculation == 0) end_dc_expected_value_calculation = IntervalLength;
  if (is_first_time_for_parameters_set) {
    double n4[4];
    int rc = cm_input_normalization(WN, wn_length, Signal, signal_length, signal_mode, Normalize_Sigma_Type,
                                    Normalize_Sigma_Type_Signal, Normalize_Noise_Energy_To_Given_Interval,
                                    Remove_Generated_Noise_DC, start_dc_expected_value_calculation,
                                    end_dc_expected_value_calculation, start_dc_normalized_value_calculation,
                                    end_dc_normalized_value_calculation, Fs, n4);
    if (rc) return rc;
    g_st.wn_dc = n4[0];
    g_st.wn_norm = n4[1];
    g_st.sig_dc = n4[2];
    g_st.sig_norm = n4[3];
  }
  int rc = setupToleranceProfile(profiles, is_first_time_for_parameters_set, Max_M1_SP_Error_Parameter,
                                 Max_Tolerance_Parameter, Relative_Error_Parameters, M1_SP_Fix_Factor,
                                 Tolerance_Fix_Factor, Blocks_Per_Interval, from_profile_index, calculatedProfiles);
  if (rc) return rc;
  const size_t total = (size_t)calculatedProfiles * IntervalLength;
  if (bm_ensure_input(std::max(bm_input_cap(), total + 2))) return -1;
  float *dst = bm_input_ptr();
  rc = cm_launch_generate_input(g_st.profiles.ptr, g_st.WN.ptr, g_st.Signal.ptr, dst, g_st.wn_dc, g_st.wn_norm,
                                signal_mode, g_st.sig_dc, g_st.sig_norm, from_profile_index, calculatedProfiles,
                                IntervalLength, JND_Interval_Head, Fs, stream());
  if (rc) {
    set_error("generateInputFromProfile: launch failed");
    return -1;
  }
  // samples past the generated profiles are read by the last time block: zero them
  if (bm_input_cap() > total &&
      check(hipMemsetAsync(dst + total, 0, (bm_input_cap() - total) * sizeof(float), stream()), "zero input tail"))
    return -1;
  const bool fir_active = fir_transfer_function_length > 1 ||
                          (fir_transfer_function_length > 0 && fir_transfer_function[0] != 1);
  const bool iir_active = iir_transfer_function_length > 1 ||
                          (iir_transfer_function_length > 0 && iir_transfer_function[0] != 1);
  if (fir_active || iir_active) {
    if (g_st.hear_fir.upload(fir_transfer_function, std::max(1, fir_transfer_function_length)) ||
        g_st.tmp.ensure(total))
      return -1;
    if (iir_active && g_st.hear_iir.upload(iir_transfer_function, iir_transfer_function_length)) return -1;
    if (cm_launch_hearing_aid(dst, g_st.tmp.ptr, dst, calculatedProfiles, IntervalLength, g_st.hear_fir.ptr,
                              fir_transfer_function_length, iir_active ? g_st.hear_iir.ptr : nullptr,
                              iir_transfer_function_length, stream()))
      return -1;
  }
  if (check(hipStreamSynchronize(stream()), "generateInputFromProfile")) return -1;
  if ((Show_Generated_Input & 1) && target_input)
    return check(hipMemcpy(target_input, dst, total * sizeof(float), hipMemcpyDeviceToHost), "copy input");
  return 0;
}

// Batched input synthesis over device pointers.
int cm_generate_input_run(const cm_jnd_profile *d_profiles, const float *d_WN, const float *d_Signal, float *d_out,
                          double wn_dc, double wn_norm, int signal_mode, double sig_dc, double sig_norm,
                          int start_profile, int n_profiles, int interval_len, int head, float Fs) {
  if (!d_profiles || !d_WN || !d_out || (signal_mode && !d_Signal)) {
    set_error("cm_generate_input_run: null argument");
    return -3;
  }
  if (start_profile < 0 || n_profiles < 0 || interval_len < 0 || !(Fs > 0.f)) {
    set_error("cm_generate_input_run: negative size (profiles %d..+%d, interval %d) or Fs %g", start_profile,
              n_profiles, interval_len, (double)Fs);
    return -3;
  }
  return cm_launch_generate_input(d_profiles, d_WN, d_Signal, d_out, wn_dc, wn_norm, signal_mode, sig_dc, sig_norm,
                                  start_profile, n_profiles, interval_len, head, Fs, stream())
             ? -1
             : 0;
}

// src/cochlea.cu:2629-2634.
int IHCKernel_Free(void) {
  g_st.ac_b.release();
  g_st.ac_a.release();
  g_st.ac_bpad.release();
  g_st.dcpad.release();
  g_st.dc.release();
  g_st.damage.release();
  g_st.nIHC.release();
  g_st.A.release();
  g_st.lambda.release();
  g_st.scratch.release();
  g_st.buffer1.release();
  g_st.buffer_valid = 0;
  g_st.stages.release();
  g_st.ivs.release();
  g_st.iir_ivs.release();
  g_st.dA.release();
  g_st.mean_rate.release();
  g_st.F_RA.release();
  g_st.FAI.release();
  g_st.ref_index.release();
  g_st.out_index.release();
  g_st.hear_fir.release();
  g_st.hear_iir.release();
  g_st.tmp.release();
  g_st.aihc_loaded = 0;
  return 0;
}

}  // extern "C"
