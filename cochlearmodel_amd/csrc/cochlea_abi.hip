// C-ABI boundary: the entry points of the reference's include/cochlea_common.h
// (extern "C" blocks :86-185 and :194-418) with plain-C signatures.  C++
// objects of the reference signatures (Log&, cudaEvent_t&, cudaDeviceProp)
// are dropped; every function returns 0 on success or a negative status and
// leaves a message for cm_last_error().  Device memory is owned by file-scope
// grow-only holders, as in src/cochlea.cu:63-631.  Single host thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cm_internal.h"

namespace cm {

static char g_err[1024] = "";
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// Grow-only device buffer.
template <typename T> struct DevBuf {
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
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

// src/cochlea.cu:122-321 (cudaHolderData) and :322-447 (generated data).
struct BMHolder {
  DevBuf<float> input, speeds, sec, gamma;
  DevBuf<int32_t> failed_nodes, failed_blocks, queue;
  DevBuf<float> conv_time, conv_loops, conv_jacobi;
  DevBuf<float> conv_time_blocks, conv_loops_blocks, conv_jacobi_blocks;
  DevBuf<float> tol, m1tol;   // per-block tolerances (setupToleranceProfile)
  DevBuf<cm_bm_item> items, trial_items;
  DevBuf<cm_item_stats> stats;
  DevBuf<int32_t> order, trial_order;   // launch order (cm_set_bm_order) and autotune scratch
  int order_n = 0;
  size_t input_nodes = 0, result_nodes = 0, lambda_nodes = 0;
  int n_sets = 1;
  int tol_blocks = 0;
  std::vector<float> host_tol, host_m1tol;
  std::vector<cm_bm_item> host_items;
  int last_blocks = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_kernel_ms = 0.f;
  int waves_per_block = 1;
  int persistent = 0;
  int jacobi_unroll = 2;   // Jacobi sweep loop unroll of bm_kernel launches (cm_set_bm_jacobi_unroll)
  int sections_per_lane = 4;   // block layout of fused-kernel launches (cm_set_bm_sections_per_lane)
  // solver options for BMOHCNewKernel launches (cm_set_bm_solver); defaults = reference method
  int jacobi_mode = 0, check_every = 4, max_sweeps = 1000, warm_start = 0;
  float tol_rel = 1e-4f;
};
BMHolder g_bm;

hipStream_t g_stream = nullptr;
hipStream_t stream() { return g_stream; }

int check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

static int ensure_events() {
  if (!g_bm.ev0) {
    if (check(hipEventCreate(&g_bm.ev0), "hipEventCreate")) return -1;
    if (check(hipEventCreate(&g_bm.ev1), "hipEventCreate")) return -1;
  }
  return 0;
}

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 1;
  }
  return cus;
}

float *bm_speeds_ptr() { return g_bm.speeds.ptr; }
size_t bm_result_nodes() { return g_bm.speeds.cap; }
float *bm_input_ptr() { return g_bm.input.ptr; }
size_t bm_input_cap() { return g_bm.input.cap; }
int bm_ensure_input(size_t n) { return g_bm.input.ensure(n); }

// Host form of the block decomposition of src/cochlea.cu:1148-1176 plus the
// disjoint row ownership for per-row outputs (a row that two neighbouring
// blocks both write stays with the earlier block).
void block_decomposition(int TB, int bufP1, int overlap, int DF, std::vector<cm_bm_item> &items) {
  items.resize(TB);
  const int I8 = (bufP1 - overlap) / (TB + 1);
  for (int b = 0; b < TB; ++b) {
    const int transient_offset = (b > 0 && (DF == 0 || (DF != 1 && b % DF != 0)));
    const int preDecoupled = DF > 0 && (DF == 1 || ((b + 1) % DF == 0)) && b != TB - 1;
    const int postDecoupled = DF > 0 && (b % DF == 0);
    const int in0 = b * I8;
    int so = in0 + std::max(1 - transient_offset - postDecoupled, 0) * overlap;
    const int eo = in0 + I8 + (1 - preDecoupled) * overlap;
    if (b > 0) so = std::max(so, items[b - 1].mesi);
    int first = std::max(in0, so - 1);
    if (b > 0 && first < items[b - 1].mesi) first = items[b - 1].mesi;
    cm_bm_item &it = items[b];
    it.start = in0;
    it.mosi = so;
    it.mesi = eo;
    it.write_start = first;
    it.tol = 0.f;
    it.m1tol = 0.f;
    it.set = 0;
    it.pad = 0;
  }
}

int launch_bm_items(const cm_bm_scalars &sc, const float *d_input, int64_t input_len, const float *d_sec,
                    const float *d_gamma, const cm_bm_item *d_items, int n_items, float *d_speeds, int64_t rows,
                    int32_t *failed_nodes, float *ct, float *cl, float *cj, int32_t *fb, float *ctb, float *clb,
                    float *cjb, cm_item_stats *d_stats, int waves_per_block, int persistent, float *kernel_ms,
                    const int32_t *order_override, int unroll = 0, int sections_per_lane = 0) {
  BmLaunch a;
  a.sc = sc;
  a.input = d_input;
  a.input_len = input_len;
  a.sec = d_sec;
  a.gamma = d_gamma;
  a.items = d_items;
  a.n_items = n_items;
  a.speeds = d_speeds;
  a.rows = rows;
  a.failed_nodes = failed_nodes;
  a.conv_time = ct;
  a.conv_loops = cl;
  a.conv_jacobi = cj;
  a.failed_blocks = fb;
  a.conv_time_blocks = ctb;
  a.conv_loops_blocks = clb;
  a.conv_jacobi_blocks = cjb;
  a.stats = d_stats;
  a.order = nullptr;
  if (g_bm.order_n > 0) {
    if (g_bm.order_n != n_items) {
      set_error("launch order holds %d items, the launch has %d (cm_set_bm_order(NULL, 0) clears it)",
                g_bm.order_n, n_items);
      return -3;
    }
    a.order = g_bm.order.ptr;
  }
  if (order_override) a.order = order_override;
  a.use_queue = persistent ? 1 : 0;
  a.queue = nullptr;
  int grid;
  if (persistent) {
    if (g_bm.queue.ensure(1)) return -1;
    a.queue = g_bm.queue.ptr;
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int waves_per_cu = 4 * std::max(1, persistent);   // SIMDs x waves per SIMD
    grid = std::max(1, std::min((n_items + waves_per_block - 1) / waves_per_block,
                                cus * waves_per_cu / waves_per_block));
  } else {
    grid = (n_items + waves_per_block - 1) / waves_per_block;
  }
  if (n_items <= 0) return 0;
  if (ensure_events()) return -1;
  if (kernel_ms && check(hipEventRecord(g_bm.ev0, stream()), "hipEventRecord")) return -1;
  int rc;
  if (ws_wants_group_kernel(sc)) {
    WsArgs w;
    if (ws_args(sc, n_items, &w)) return -3;
    rc = cm_launch_bm_ws(&a, &w, device_cus(), stream());
    rc = rc > 0 ? 0 : rc;
  } else {
    rc = cm_launch_bm(&a, waves_per_block, unroll > 0 ? unroll : g_bm.jacobi_unroll,
                      sections_per_lane > 0 ? sections_per_lane : g_bm.sections_per_lane, grid, stream());
  }
  if (rc) {
    set_error("BM kernel launch failed (%d): %s", rc, hipGetErrorString(hipGetLastError()));
    return -1;
  }
  if (kernel_ms) {
    if (check(hipEventRecord(g_bm.ev1, stream()), "hipEventRecord")) return -1;
    if (check(hipEventSynchronize(g_bm.ev1), "hipEventSynchronize")) return -1;
    if (check(hipEventElapsedTime(kernel_ms, g_bm.ev0, g_bm.ev1), "hipEventElapsedTime")) return -1;
  }
  return 0;
}

}  // namespace cm

using namespace cm;

extern "C" {

const char *cm_last_error(void) { return g_err; }

int cm_set_stream(void *s) {
  g_stream = (hipStream_t)s;
  return 0;
}

int cm_device_synchronize(void) { return check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

// Diagnostics of src/cochlea.cu:708-746 with plain-C handles: the events are
// hipEvent_t values passed as void*, the prefix a C string.
int viewGPUStatus(int flags, const char *prefix) {
  if (!(flags & 16)) return 0;
  size_t free_b = 0, total_b = 0;
  if (check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo")) return -1;
  std::printf("%s : GPU Memory, Free(%d MB) / Total(%d MB)\n", prefix ? prefix : "", (int)(free_b >> 20),
              (int)(total_b >> 20));
  return 0;
}
int cudaEventsCreate(void **start, void **stop, int condition) {
  if (!condition) return 0;
  if (!start || !stop) {
    set_error("cudaEventsCreate: null handle");
    return -3;
  }
  hipEvent_t a, b;
  if (check(hipEventCreate(&a), "hipEventCreate") || check(hipEventCreate(&b), "hipEventCreate")) return -1;
  *start = a;
  *stop = b;
  return 0;
}
int cudaEventsStartTimer(void *start, void *stop, int condition) {
  (void)stop;
  return condition ? check(hipEventRecord((hipEvent_t)start, stream()), "hipEventRecord") : 0;
}
int cudaEventsStopTimer(void *start, void *stop, int condition) {
  (void)start;
  return condition ? check(hipEventRecord((hipEvent_t)stop, stream()), "hipEventRecord") : 0;
}
// returns the elapsed milliseconds (and prints them as the reference does), or
// -1 on an error
float cudaEventsStopQueryTimer(void *start, void *stop, int condition, const char *prefix) {
  if (!condition) return 0.0f;
  float ms = 0.0f;
  if (cudaEventsStopTimer(start, stop, condition) || check(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize") ||
      check(hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime"))
    return -1.0f;
  std::printf("%s : %.2f (msec) \n", prefix ? prefix : "", ms);
  return ms;
}
int cudaEventsDestroy(void *start, void *stop) {
  if (start) (void)hipEventDestroy((hipEvent_t)start);
  if (stop) (void)hipEventDestroy((hipEvent_t)stop);
  return 0;
}

float cm_last_kernel_ms(void) { return g_bm.last_kernel_ms; }

// Host-side longest-first order of a block table (stable for equal lengths).
static void lpt_order(const cm_bm_item *items, int n, std::vector<int32_t> &ord) {
  ord.resize(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
    return (items[x].mesi - items[x].start) > (items[y].mesi - items[y].start);
  });
}

int cm_set_bm_order(const int32_t *h_order, int n) {
  if (!h_order || n <= 0) {
    g_bm.order_n = 0;
    return 0;
  }
  std::vector<char> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    if (h_order[i] < 0 || h_order[i] >= n || seen[h_order[i]]) {
      set_error("cm_set_bm_order: not a permutation of 0..%d (entry %d)", n - 1, i);
      return -3;
    }
    seen[h_order[i]] = 1;
  }
  if (g_bm.order.ensure(n)) return -1;
  if (check(hipMemcpy(g_bm.order.ptr, h_order, n * sizeof(int32_t), hipMemcpyHostToDevice), "upload order"))
    return -1;
  g_bm.order_n = n;
  return 0;
}

// Times every launch configuration on a sample of the block table (every
// trial_stride-th block in longest-first order, so the sample keeps the length
// mix) and returns the fastest; with order_lpt the full longest-first order is
// installed for the following launches of this table.  Trial runs write their
// rows of d_speeds, which the real launch overwrites.
int cm_bm_autotune(const cm_bm_scalars *sc, const float *d_input, int64_t input_len, const float *d_sec,
                   const float *d_gamma, const cm_bm_item *h_items, int n_items, float *d_speeds, int64_t rows,
                   int trial_stride, cm_bm_tuning *out) {
  if (!sc || !d_input || !d_sec || !d_gamma || !h_items || n_items <= 0 || !d_speeds || !out || trial_stride < 1) {
    set_error("cm_bm_autotune: bad arguments");
    return -3;
  }
  std::vector<int32_t> full;
  lpt_order(h_items, n_items, full);
  std::vector<cm_bm_item> sample;
  for (int i = 0; i < n_items; i += trial_stride) sample.push_back(h_items[full[i]]);
  // natural order of the sample = the blocks' original order
  std::vector<int32_t> pick;
  for (int i = 0; i < n_items; i += trial_stride) pick.push_back(full[i]);
  std::vector<int32_t> nat(pick.size());
  for (size_t i = 0; i < nat.size(); ++i) nat[i] = (int32_t)i;
  std::stable_sort(nat.begin(), nat.end(), [&](int32_t x, int32_t y) { return pick[x] < pick[y]; });
  std::vector<int32_t> lpt(sample.size());
  for (size_t i = 0; i < lpt.size(); ++i) lpt[i] = (int32_t)i;   // sample is already longest first
  const int ns = (int)sample.size();
  if (g_bm.trial_items.ensure(ns) || g_bm.trial_order.ensure(2 * ns)) return -1;
  if (check(hipMemcpy(g_bm.trial_items.ptr, sample.data(), ns * sizeof(cm_bm_item), hipMemcpyHostToDevice),
            "upload trial items") ||
      check(hipMemcpy(g_bm.trial_order.ptr, nat.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice),
            "upload trial order") ||
      check(hipMemcpy(g_bm.trial_order.ptr + ns, lpt.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice),
            "upload trial order"))
    return -1;
  const int saved_n = g_bm.order_n;
  g_bm.order_n = 0;
  // trial c < 18 (4 sections per lane): c = 6 u + 2 w + o, waves per workgroup
  // wpb[w], order o (natural, longest first), Jacobi unroll unr[u]; trial
  // c = 18 + 6 v + 2 u + o: spl[v] sections per lane (one block per workgroup)
  const int wpb[3] = {1, 2, 4}, unr[3] = {2, 4, 1}, spl[2] = {2, 1};
  float best = 1e30f;
  int bi = 0;
  for (int c = -1; c < 30; ++c) {   // c = -1: untimed warm-up
    const int cc = c < 0 ? 0 : c;
    const bool split = cc >= 18;
    const int t = split ? cc - 18 : cc;
    const int o = t % 2;
    const int uu = split ? unr[(t % 6) / 2] : unr[t / 6];
    const int w = split ? 1 : wpb[(t % 6) / 2];
    const int n = split ? spl[t / 6] : 4;
    float ms = 0.f;
    const int rc = launch_bm_items(*sc, d_input, input_len, d_sec, d_gamma, g_bm.trial_items.ptr, ns, d_speeds, rows,
                                   nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                   w, 0, &ms, g_bm.trial_order.ptr + o * ns, uu, n);
    if (rc) {
      g_bm.order_n = saved_n;
      return rc;
    }
    if (c < 0) continue;
    if (split)
      out->ms_split[t] = ms;
    else
      out->ms[c] = ms;
    if (ms < best) {
      best = ms;
      bi = c;
    }
  }
  g_bm.order_n = saved_n;
  const bool bsplit = bi >= 18;
  const int bt = bsplit ? bi - 18 : bi;
  out->sections_per_lane = bsplit ? spl[bt / 6] : 4;
  out->waves_per_block = bsplit ? 1 : wpb[(bt % 6) / 2];
  out->order_lpt = bt % 2;
  out->n_trial_items = ns;
  out->jacobi_unroll = bsplit ? unr[(bt % 6) / 2] : unr[bt / 6];
  g_bm.jacobi_unroll = out->jacobi_unroll;
  g_bm.sections_per_lane = out->sections_per_lane;
  if (out->order_lpt) return cm_set_bm_order(full.data(), n_items);
  g_bm.order_n = 0;
  return 0;
}

int cm_set_bm_solver(int jacobi_mode, float tol_rel, int check_every, int max_sweeps, int warm_start) {
  if (jacobi_mode < 0 || jacobi_mode > 1 || check_every < 1 || max_sweeps < 1 || warm_start < 0 ||
      warm_start > 2 || (jacobi_mode == 1 && !(tol_rel > 0.f)) || (warm_start != 0 && jacobi_mode != 1)) {
    set_error("cm_set_bm_solver: jacobi_mode 0|1, tol_rel > 0, check_every/max_sweeps >= 1, warm_start 0..2 "
              "(warm starts need jacobi_mode 1)");
    return -3;
  }
  g_bm.jacobi_mode = jacobi_mode;
  g_bm.tol_rel = tol_rel;
  g_bm.check_every = check_every;
  g_bm.max_sweeps = max_sweeps;
  g_bm.warm_start = warm_start;
  return 0;
}

int cm_set_bm_jacobi_unroll(int unroll) {
  if (unroll != 1 && unroll != 2 && unroll != 4) {
    set_error("cm_set_bm_jacobi_unroll: unroll must be 1, 2 or 4 (got %d)", unroll);
    return -3;
  }
  g_bm.jacobi_unroll = unroll;
  return 0;
}

int cm_set_bm_sections_per_lane(int sections_per_lane) {
  if (sections_per_lane != 1 && sections_per_lane != 2 && sections_per_lane != 4) {
    set_error("cm_set_bm_sections_per_lane: 1, 2 or 4 sections per lane (got %d)", sections_per_lane);
    return -3;
  }
  g_bm.sections_per_lane = sections_per_lane;
  return 0;
}

int cm_set_bm_launch(int waves_per_block, int persistent) {
  if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4) {
    set_error("waves_per_block must be 1, 2 or 4 (got %d)", waves_per_block);
    return -1;
  }
  if (persistent) {
    set_error("cm_set_bm_launch: the persistent work-queue launch is disabled");
    return -3;
  }
  g_bm.waves_per_block = waves_per_block;
  g_bm.persistent = persistent;
  return 0;
}

// include/cochlea_common.h:209-230 -> src/cochlea.cu:763-814.
int BMOHCKernel_Init(const float *gamma, const float *mass, const float *M, const float *U, const float *L,
                     const float *R, const float *S, const float *Q, const float *S_ohc, const float *S_tm,
                     const float *R_tm, float num_frequencies, float dbA, size_t inputBufferNodes,
                     size_t resultBufferNodes, size_t lambdaBufferNodes, int first_time, int Show_Run_Time,
                     int Show_Device_Data) {
  (void)U; (void)L; (void)num_frequencies; (void)dbA; (void)first_time; (void)Show_Run_Time; (void)Show_Device_Data;
  if (!gamma || !mass || !M || !R || !S || !Q || !S_ohc || !S_tm || !R_tm) {
    set_error("BMOHCKernel_Init: null parameter vector");
    return -1;
  }
  if (g_bm.sec.ensure(CM_SEC_COUNT * CM_SECTIONS) || g_bm.gamma.ensure(CM_SECTIONS)) return -1;
  float host[CM_SEC_COUNT * CM_SECTIONS];
  const float *src[CM_SEC_M] = {mass, M, R, S, Q, S_ohc, S_tm, R_tm};
  for (int k = 0; k < CM_SEC_M; ++k) std::memcpy(host + k * CM_SECTIONS, src[k], CM_SECTIONS * sizeof(float));
  // the diagonal itself (only the tolerance-mode residual uses it)
  for (int i = 0; i < CM_SECTIONS; ++i) host[CM_SEC_M * CM_SECTIONS + i] = (float)(1.0 / (double)M[i]);
  if (check(hipMemcpy(g_bm.sec.ptr, host, sizeof(host), hipMemcpyHostToDevice), "upload section params")) return -1;
  if (check(hipMemcpy(g_bm.gamma.ptr, gamma, CM_SECTIONS * sizeof(float), hipMemcpyHostToDevice), "upload gamma"))
    return -1;
  g_bm.n_sets = 1;
  if (g_bm.input.ensure(inputBufferNodes + 2) || g_bm.speeds.ensure(resultBufferNodes)) return -1;
  if (g_bm.failed_nodes.ensure(inputBufferNodes) || g_bm.conv_time.ensure(inputBufferNodes) ||
      g_bm.conv_loops.ensure(inputBufferNodes) || g_bm.conv_jacobi.ensure(inputBufferNodes))
    return -1;
  g_bm.input_nodes = inputBufferNodes;
  g_bm.result_nodes = resultBufferNodes;
  g_bm.lambda_nodes = lambdaBufferNodes;
  return 0;
}

// include/cochlea_common.h:126-176 -> src/cochlea.cu:990-1200, 1436-1493.
int BMOHCNewKernel(const float *input_samples, int override_input_samples, float w_ohc, float time_step,
                   float time_step_out, float delta_x, float alpha_r, int enable_psi, int enable_OW, int base_index,
                   float Ts, float _ohc_alpha_l, float _ohc_alpha_s, float model_Gme, float model_a0, float model_a1,
                   float model_a2, float sigma_ow, float eta_1, float eta_2, int Time_Blocks, int samplesBufferLengthP1,
                   int overlap_nodes_for_block, long overlapTimeMicroSec, int show_transient,
                   float cuda_max_time_step, float cuda_min_time_step, int Decouple_Filter,
                   float Max_M1_SP_Error_Parameter, float Max_Tolerance_Parameter, int Relative_Error_Parameters,
                   int calculate_boundary_conditions, float M1_SP_Fix_Factor, float Tolerance_Fix_Factor,
                   float SPLREfVal, int Show_Calculated_Power, int Show_Device_Data, int Show_Run_Time,
                   int JACOBBY_Loops_Fast, int JACOBBY_Loops_Slow, int Cuda_Outern_Loops,
                   int Run_Fast_BM_Calculation, int BMOHC_Kernel_Configuration) {
  (void)enable_psi; (void)base_index; (void)overlapTimeMicroSec; (void)show_transient;
  (void)Show_Calculated_Power; (void)Show_Device_Data; (void)BMOHC_Kernel_Configuration;
  if (Run_Fast_BM_Calculation != 17) {
    set_error("BMOHCNewKernel: only Run_Fast_BM_Calculation=17 is provided (got %d)", Run_Fast_BM_Calculation);
    return -3;
  }
  if (enable_OW != 1) {
    set_error("BMOHCNewKernel: kernel 17 assumes enable_OW=1");
    return -3;
  }
  if (Time_Blocks <= 0 || Time_Blocks > 65536) {
    set_error("BMOHCNewKernel: Time_Blocks=%d outside [1, 65536]", Time_Blocks);
    return -3;
  }
  if (!g_bm.sec.ptr) {
    set_error("BMOHCNewKernel: BMOHCKernel_Init was not called");
    return -3;
  }
  if ((size_t)samplesBufferLengthP1 + 2 > g_bm.input.cap) {
    set_error("BMOHCNewKernel: input buffer holds %zu nodes, %d needed", g_bm.input.cap, samplesBufferLengthP1 + 2);
    return -3;
  }
  if ((size_t)samplesBufferLengthP1 * CM_SECTIONS > g_bm.speeds.cap) {
    set_error("BMOHCNewKernel: result buffer holds %zu nodes, %zu needed", g_bm.speeds.cap,
              (size_t)samplesBufferLengthP1 * CM_SECTIONS);
    return -3;
  }
  if (!override_input_samples) {
    if (!input_samples) {
      set_error("BMOHCNewKernel: input_samples is null");
      return -1;
    }
    if (check(hipMemcpy(g_bm.input.ptr, input_samples, (size_t)samplesBufferLengthP1 * sizeof(float),
                        hipMemcpyHostToDevice), "upload input"))
      return -1;
    if (check(hipMemset(g_bm.input.ptr + samplesBufferLengthP1, 0, 2 * sizeof(float)), "pad input")) return -1;
  }
  cm_bm_scalars sc;
  std::memset(&sc, 0, sizeof(sc));
  sc.w_ohc = w_ohc;
  sc.time_step = time_step;
  sc.time_step_out = time_step_out;
  sc.delta_x = delta_x;
  sc.alpha_r = alpha_r;
  sc.Ts = Ts;
  sc.alpha_l = _ohc_alpha_l;
  sc.alpha_s = _ohc_alpha_s;
  sc.Gme = model_Gme;
  sc.a0 = model_a0;
  sc.a1 = model_a1;
  sc.a2 = model_a2;
  sc.sigma_ow = sigma_ow;
  sc.eta1 = eta_1;
  sc.eta2 = eta_2;
  sc.dt_min = cuda_min_time_step;
  sc.dt_max = cuda_max_time_step;
  sc.j_fast = JACOBBY_Loops_Fast;
  sc.j_slow = JACOBBY_Loops_Slow;
  sc.max_passes = Cuda_Outern_Loops;
  sc.decision_mode = 0;
  sc.jacobi_mode = g_bm.jacobi_mode;
  sc.check_every = g_bm.check_every;
  sc.max_sweeps = g_bm.max_sweeps;
  sc.tol_rel = g_bm.tol_rel;
  sc.warm_start = g_bm.warm_start;

  std::vector<cm_bm_item> &items = g_bm.host_items;
  block_decomposition(Time_Blocks, samplesBufferLengthP1, overlap_nodes_for_block, Decouple_Filter, items);
  const int I8 = (samplesBufferLengthP1 - overlap_nodes_for_block) / (Time_Blocks + 1);
  if (calculate_boundary_conditions) {
    // src/cochlea.cu:847-963: simple (absolute) or power-relative tolerances
    std::vector<float> host_in;
    if (Relative_Error_Parameters != 0) {
      host_in.resize(samplesBufferLengthP1);
      if (input_samples && !override_input_samples) {
        std::memcpy(host_in.data(), input_samples, samplesBufferLengthP1 * sizeof(float));
      } else if (check(hipMemcpy(host_in.data(), g_bm.input.ptr, samplesBufferLengthP1 * sizeof(float),
                                 hipMemcpyDeviceToHost), "download input")) {
        return -1;
      }
    }
    int max_block_length = 0;
    for (auto &it : items) max_block_length = std::max(max_block_length, it.mesi - it.start);
    int threads = std::min(max_block_length, 1024);
    if (threads < 1024) threads = (int)powf(2.0f, floorf(log2f((float)threads)));
    const float rspl = 1.0f / SPLREfVal;
    for (int b = 0; b < Time_Blocks; ++b) {
      if (Relative_Error_Parameters == 0) {
        items[b].m1tol = Max_M1_SP_Error_Parameter;
        items[b].tol = Max_Tolerance_Parameter;
      } else {
        float mx = 0.f;
        const int s0 = b * I8;
        for (int k = 0; k < threads && s0 + k < samplesBufferLengthP1; ++k) mx = fmaxf(mx, fabsf(host_in[s0 + k]));
        const float power = rspl * mx;
        const float db = power > 0.0f ? 20 * log10f(0.1f * power) : 0.0f;
        items[b].m1tol = Max_M1_SP_Error_Parameter * powf(10.0f, M1_SP_Fix_Factor * db);
        items[b].tol = Max_Tolerance_Parameter * powf(10.0f, Tolerance_Fix_Factor * db);
      }
    }
  } else {
    if (g_bm.tol_blocks < Time_Blocks) {
      set_error("BMOHCNewKernel: setupToleranceProfile covered %d blocks, %d launched", g_bm.tol_blocks, Time_Blocks);
      return -3;
    }
    for (int b = 0; b < Time_Blocks; ++b) {
      items[b].m1tol = g_bm.host_m1tol[b];
      items[b].tol = g_bm.host_tol[b];
    }
  }
  if (g_bm.items.ensure(Time_Blocks) || g_bm.stats.ensure(Time_Blocks) || g_bm.failed_blocks.ensure(Time_Blocks) ||
      g_bm.conv_time_blocks.ensure(Time_Blocks) || g_bm.conv_loops_blocks.ensure(Time_Blocks) ||
      g_bm.conv_jacobi_blocks.ensure(Time_Blocks))
    return -1;
  if (check(hipMemcpy(g_bm.items.ptr, items.data(), Time_Blocks * sizeof(cm_bm_item), hipMemcpyHostToDevice),
            "upload blocks"))
    return -1;
  g_bm.last_blocks = Time_Blocks;
  // rows no block owns (before the first block's output, after the last block's
  // end) are zero rather than left over from a previous run
  {
    const size_t first = (size_t)std::max(0, items[0].write_start);
    const size_t last = (size_t)std::min(samplesBufferLengthP1, items[Time_Blocks - 1].mesi);
    if (first > 0 && check(hipMemsetAsync(g_bm.speeds.ptr, 0, first * CM_SECTIONS * sizeof(float), stream()),
                           "zero leading rows"))
      return -1;
    if ((size_t)samplesBufferLengthP1 > last &&
        check(hipMemsetAsync(g_bm.speeds.ptr + last * CM_SECTIONS, 0,
                             ((size_t)samplesBufferLengthP1 - last) * CM_SECTIONS * sizeof(float), stream()),
              "zero trailing rows"))
      return -1;
  }
  float ms = 0.f;
  int rc = launch_bm_items(sc, g_bm.input.ptr, (int64_t)samplesBufferLengthP1 + 2, g_bm.sec.ptr, g_bm.gamma.ptr,
                           g_bm.items.ptr, Time_Blocks, g_bm.speeds.ptr, (int64_t)samplesBufferLengthP1,
                           g_bm.failed_nodes.ptr, g_bm.conv_time.ptr, g_bm.conv_loops.ptr, g_bm.conv_jacobi.ptr,
                           g_bm.failed_blocks.ptr, g_bm.conv_time_blocks.ptr, g_bm.conv_loops_blocks.ptr,
                           g_bm.conv_jacobi_blocks.ptr, g_bm.stats.ptr, g_bm.waves_per_block, g_bm.persistent,
                           (Show_Run_Time & 1) ? &ms : nullptr, nullptr);
  if (rc) return rc;
  g_bm.last_kernel_ms = ms;
  return check(hipStreamSynchronize(stream()), "BM kernel");
}

// Batched entry (no reference counterpart): an explicit block table over one
// concatenated input buffer, one gamma vector per parameter set, per-block
// tolerances; all pointers are device pointers.  Rows of `speeds` are input
// sample indices, as in the reference.
int cm_bm_run(const cm_bm_scalars *sc, const float *d_input, int64_t input_len, const float *d_sec,
              const float *d_gamma, const cm_bm_item *d_items, int n_items, float *d_speeds, int64_t rows,
              int32_t *d_failed_nodes, float *d_conv_time, float *d_conv_loops, float *d_conv_jacobi,
              int32_t *d_failed_blocks, float *d_conv_time_blocks, float *d_conv_loops_blocks,
              float *d_conv_jacobi_blocks, cm_item_stats *d_stats, int waves_per_block, int persistent,
              int timed) {
  if (!sc || !d_input || !d_sec || !d_gamma || !d_items || n_items < 0) {
    set_error("cm_bm_run: null argument");
    return -1;
  }
  if (persistent) {
    // The work-queue variant went silent on its first MI355X run and its cause
    // has not been shown; it stays switched off.
    set_error("cm_bm_run: the persistent work-queue launch is disabled");
    return -3;
  }
  if (sc->j_fast < 0 || sc->j_slow < 0 || sc->max_passes < 1) {
    set_error("cm_bm_run: invalid Jacobi/pass counts");
    return -3;
  }
  if (sc->jacobi_mode == 1 && (sc->warm_start < 0 || sc->warm_start > 2 || sc->check_every < 1)) {
    set_error("cm_bm_run: warm_start must be 0, 1 or 2 and check_every >= 1");
    return -3;
  }
  float ms = 0.f;
  int rc = launch_bm_items(*sc, d_input, input_len, d_sec, d_gamma, d_items, n_items, d_speeds, rows, d_failed_nodes,
                           d_conv_time, d_conv_loops, d_conv_jacobi, d_failed_blocks, d_conv_time_blocks,
                           d_conv_loops_blocks, d_conv_jacobi_blocks, d_stats, waves_per_block, persistent,
                           timed ? &ms : nullptr, nullptr);
  if (rc) return rc;
  if (timed) g_bm.last_kernel_ms = ms;
  return 0;
}

// Host-side block table of the last BMOHCNewKernel launch (diagnostics).
int cm_get_last_blocks(cm_bm_item *out, int max_blocks) {
  const int n = std::min(max_blocks, (int)g_bm.host_items.size());
  if (out && n > 0) std::memcpy(out, g_bm.host_items.data(), n * sizeof(cm_bm_item));
  return (int)g_bm.host_items.size();
}

int cm_get_last_stats(cm_item_stats *out, int max_blocks) {
  const int n = std::min(max_blocks, g_bm.last_blocks);
  if (n <= 0) return 0;
  if (check(hipMemcpy(out, g_bm.stats.ptr, n * sizeof(cm_item_stats), hipMemcpyDeviceToHost), "download stats"))
    return -1;
  return n;
}

// include/cochlea_common.h:179 -> src/cochlea.cu:819-822: waits for all work
// on the device.
int BMOHCKernel_Wait_Threads(void) { return check(hipDeviceSynchronize(), "BMOHCKernel_Wait_Threads"); }

// include/cochlea_common.h -> src/cochlea.cu:823-827.
int BMOHCKernel_Copy_Results(float *target, size_t resultNodes, size_t offset) {
  if (offset + resultNodes > g_bm.speeds.cap) {
    set_error("BMOHCKernel_Copy_Results: %zu+%zu exceeds %zu nodes", offset, resultNodes, g_bm.speeds.cap);
    return -3;
  }
  return check(hipMemcpy(target, g_bm.speeds.ptr + offset, resultNodes * sizeof(float), hipMemcpyDeviceToHost),
               "copy results");
}

// src/cochlea.cu:828-833: upload BM velocity for a run that starts at the lambda stage.
int ReverseKernel_Copy_Results(const float *src, size_t size) {
  if (g_bm.speeds.ensure(size)) return -1;
  return check(hipMemcpy(g_bm.speeds.ptr, src, size * sizeof(float), hipMemcpyHostToDevice), "reverse copy");
}

int extractConvergenceTimes(float *convergence_times, size_t nodes) {
  if (nodes > g_bm.conv_time.cap) nodes = g_bm.conv_time.cap;
  return check(hipMemcpy(convergence_times, g_bm.conv_time.ptr, nodes * sizeof(float), hipMemcpyDeviceToHost),
               "extractConvergenceTimes");
}

void *getCudaFailedTimeNodes(void) { return g_bm.failed_nodes.ptr; }
void *getCudaFailedBlocks(void) { return g_bm.failed_blocks.ptr; }
void *getCudaConvergeSpeedBlocks(void) { return g_bm.conv_time_blocks.ptr; }
void *getCudaConvergedTimeNodes(void) { return g_bm.conv_loops.ptr; }
void *getCudaConvergedBlocks(void) { return g_bm.conv_loops_blocks.ptr; }
void *getCudaConvergedJacobyLoopsPerIteration(void) { return g_bm.conv_jacobi.ptr; }
void *getCudaConvergedJacobyLoopsPerIterationBlocks(void) { return g_bm.conv_jacobi_blocks.ptr; }
void *getCudaBMVelocity(void) { return g_bm.speeds.ptr; }
void *getCudaInputSamples(void) { return g_bm.input.ptr; }

// GeneralKernel_Copy_Results_Template<T> family (src/cochlea.cu:3678-3693).
int GeneralKernel_Copy_Results_Bytes(void *target, const void *src, size_t bytes, size_t offset_bytes) {
  return check(hipMemcpy(target, (const char *)src + offset_bytes, bytes, hipMemcpyDeviceToHost), "copy results");
}
int GeneralKernel_Copy_Results(float *target, const float *src, size_t size) {
  return GeneralKernel_Copy_Results_Bytes(target, src, size * sizeof(float), 0);
}
int GeneralKernel_Copy_Results_Double(double *target, const double *src, size_t size) {
  return GeneralKernel_Copy_Results_Bytes(target, src, size * sizeof(double), 0);
}

// src/cochlea.cu:4256-4287.
int setupToleranceProfile(const cm_jnd_profile *profiles, int is_first_time_for_parameters_set,
                          float Max_M1_SP_Error_Parameter, float Max_Tolerance_Parameter,
                          int Relative_Error_Parameters, float M1_SP_Fix_Factor, float Tolerance_Fix_Factor,
                          int Blocks_Per_Interval, int from_profile_index, int calculatedProfiles) {
  (void)is_first_time_for_parameters_set;
  if (!profiles || calculatedProfiles < 0 || Blocks_Per_Interval <= 0) {
    set_error("setupToleranceProfile: invalid arguments");
    return -1;
  }
  const int n = calculatedProfiles * Blocks_Per_Interval;
  if (n > 65536) {
    set_error("setupToleranceProfile: %d blocks exceed MAX_NUMBER_OF_BLOCKS", n);
    return -3;
  }
  g_bm.host_tol.assign(n, 0.f);
  g_bm.host_m1tol.assign(n, 0.f);
  for (int i = 0; i < calculatedProfiles; ++i) {
    const int gp = from_profile_index + i;
    float m1f = 1.0f, tf = 1.0f;
    if (Relative_Error_Parameters > 0) {
      m1f = powf(10.0f, M1_SP_Fix_Factor * (float)profiles[gp].dBSPLSignal);
      tf = powf(10.0f, Tolerance_Fix_Factor * (float)profiles[gp].dBSPLSignal);
    }
    for (int j = 0; j < Blocks_Per_Interval; ++j) {
      g_bm.host_m1tol[Blocks_Per_Interval * i + j] = Max_M1_SP_Error_Parameter * m1f;
      g_bm.host_tol[Blocks_Per_Interval * i + j] = Max_Tolerance_Parameter * tf;
    }
  }
  g_bm.tol_blocks = n;
  return 0;
}

// src/cochlea.cu:815-822.
int BMOHCKernel_Free(void) {
  g_bm.input.release();
  g_bm.speeds.release();
  g_bm.sec.release();
  g_bm.gamma.release();
  g_bm.failed_nodes.release();
  g_bm.failed_blocks.release();
  g_bm.queue.release();
  g_bm.conv_time.release();
  g_bm.conv_loops.release();
  g_bm.conv_jacobi.release();
  g_bm.conv_time_blocks.release();
  g_bm.conv_loops_blocks.release();
  g_bm.conv_jacobi_blocks.release();
  g_bm.items.release();
  g_bm.stats.release();
  return 0;
}

}  // extern "C"
