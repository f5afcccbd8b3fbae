// Post-BM stages as one HIP graph: IHC/lambda (AC FIR + dS, DC + lambda) and
// the JND reductions (mean rate, Fisher sums) over device buffers, captured
// once and replayed per batch.  Everything a replay reads besides the BM
// velocity (interval tables, zero-padded tap tables) is uploaded into this
// graph's own buffers before capture, so the graph holds kernel nodes only:
// no memset or host-copy nodes.  Outputs are assigned, not accumulated, so
// replays are idempotent.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "cm_internal.h"

namespace {

struct PostBmGraph {
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  cm_ihc_interval *d_ivs = nullptr;
  double *d_ac_bpad = nullptr, *d_dcpad = nullptr;
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture) (void)hipStreamDestroy(capture);
    if (d_ivs) (void)hipFree(d_ivs);
    if (d_ac_bpad) (void)hipFree(d_ac_bpad);
    if (d_dcpad) (void)hipFree(d_dcpad);
    *this = PostBmGraph{};
  }
};

// taps copied into a zero-padded table of a whole number of 16-tap chunks
int padded_taps(const double *d_src, int n, double **out) {
  const size_t np = (size_t)((n + 15) / 16) * 16;
  std::vector<double> h(np, 0.0);
  if (cm::check(hipMemcpy(h.data(), d_src, (size_t)n * sizeof(double), hipMemcpyDeviceToHost), "taps")) return -1;
  if (cm::check(hipMalloc((void **)out, np * sizeof(double)), "hipMalloc")) return -1;
  return cm::check(hipMemcpy(*out, h.data(), np * sizeof(double), hipMemcpyHostToDevice), "taps");
}

}  // namespace

extern "C" {

// Builds the graph.  IHC/lambda arguments as cm_ihc_lambda_run (FIR AC filter
// only), JND arguments as cm_jnd_run; *handle receives the graph.
int cm_post_bm_graph_create(const float *d_bm, int64_t rows, const cm_ihc_interval *h_ivs, int n_ivs,
                            const double *d_ac_b, int n_ac, const double *d_dc, int n_dc, const double *d_damage,
                            const float *d_A, const float *spont3, float lambda_sat, float eta_ac, float eta_dc,
                            float scale_bm, float *d_scratch, float *d_lambda, int n_intervals, int interval_nodes,
                            int head_nodes, int length_nodes, const float *d_dA, const int32_t *d_ref_index,
                            const int32_t *d_out_index, int out_stride, const double *d_nIHC, const float *W3,
                            double Fs, double delta_alpha_factor, int calc_ai, int calc_rms, float *d_mean_rate,
                            float *d_F_RA, float *d_FisherAISum, void **handle) {
  using namespace cm;
  if (!handle || !d_bm || !h_ivs || n_ivs <= 0 || !d_ac_b || !d_dc || !d_damage || !d_A || !spont3 || !d_scratch ||
      !d_lambda || !d_dA || !d_ref_index || !d_out_index || !d_nIHC || !W3 || !d_mean_rate ||
      (calc_rms && !d_F_RA) || (calc_ai && !d_FisherAISum)) {
    set_error("cm_post_bm_graph_create: null argument");
    return -3;
  }
  for (int i = 0; i < n_ivs; ++i)
    if (h_ivs[i].row0 < 0 || h_ivs[i].row0 + h_ivs[i].len > rows || h_ivs[i].len < 0) {
      set_error("cm_post_bm_graph_create: interval %d outside [0, %lld)", i, (long long)rows);
      return -3;
    }
  if ((int64_t)n_intervals * interval_nodes > rows || head_nodes + length_nodes > interval_nodes) {
    set_error("cm_post_bm_graph_create: JND intervals exceed the lambda planes");
    return -3;
  }
  PostBmGraph *g = new PostBmGraph();
  auto fail = [&](int rc) {
    g->release();
    delete g;
    return rc;
  };
  if (check(hipMalloc((void **)&g->d_ivs, n_ivs * sizeof(cm_ihc_interval)), "hipMalloc") ||
      check(hipMemcpy(g->d_ivs, h_ivs, n_ivs * sizeof(cm_ihc_interval), hipMemcpyHostToDevice), "intervals") ||
      padded_taps(d_ac_b, n_ac, &g->d_ac_bpad) || padded_taps(d_dc, n_dc, &g->d_dcpad) ||
      check(hipStreamCreate(&g->capture), "hipStreamCreate"))
    return fail(-1);
  JndArgs a;
  std::memset(&a, 0, sizeof(a));
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
  if (check(hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
    return fail(-1);
  int rc = cm_launch_ihc_lambda(d_bm, rows, g->d_ivs, h_ivs, n_ivs, nullptr, 0, d_ac_b, g->d_ac_bpad, nullptr, n_ac,
                                d_dc, g->d_dcpad, n_dc, d_damage, d_A, spont3, lambda_sat, eta_ac, eta_dc, scale_bm,
                                d_scratch, d_lambda, nullptr, g->capture);
  if (!rc) rc = cm_launch_jnd(&a, g->capture);
  const hipError_t ec = hipStreamEndCapture(g->capture, &g->graph);
  if (rc || ec != hipSuccess) {
    set_error("cm_post_bm_graph_create: capture failed (%d, %s)", rc, hipGetErrorString(ec));
    return fail(-1);
  }
  if (check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate")) return fail(-1);
  *handle = g;
  return 0;
}

// Replays the graph on the library's stream (cm_set_stream).
int cm_post_bm_graph_launch(void *handle) {
  using namespace cm;
  PostBmGraph *g = static_cast<PostBmGraph *>(handle);
  if (!g || !g->exec) {
    set_error("cm_post_bm_graph_launch: no graph");
    return -3;
  }
  return check(hipGraphLaunch(g->exec, stream()), "hipGraphLaunch");
}

int cm_post_bm_graph_destroy(void *handle) {
  PostBmGraph *g = static_cast<PostBmGraph *>(handle);
  if (g) {
    g->release();
    delete g;
  }
  return 0;
}

}  // extern "C"
