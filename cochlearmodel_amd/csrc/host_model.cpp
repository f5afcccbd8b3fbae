// Host-side physical model: CModel vectors, the tridiagonal of the pressure
// equation and the scalar constants Run_Cuda hands to the BM launch.
//   CModel::CModel            src/model.cpp:18-86 (macros include/const.h:216-255)
//   CSolver::Init_Tri_Matrix  src/solver.cpp:155-176
//   parameter packing         src/solver.cpp:357-376, 629-659
// The reference's constants are float-suffixed macros; they are evaluated in
// float and promoted, as the reference does, before any FP64 arithmetic.
#include <cmath>
#include <cstring>

#include "../../include/cochlea_amd.h"

namespace {
constexpr int NS = CM_SECTIONS;
constexpr float kPI = 3.141592653589793f;  // include/const.h:142
constexpr float kCochleaLen = 3.5f, kArea = 0.5f, kRho = 1.0f, kBeta = 0.003f;
constexpr float kM0 = 1.286e-6f, kM1 = 1.5f, kR0 = 0.25f, kR1 = -0.06f, kS0 = 1.282e4f, kS1 = -1.5f;
constexpr float kSigmaOW = 0.5f, kGammaOW = 20e3f, kGme = 21.4f, kCow = 6e-3f, kRatioK = 1e-3f;
constexpr float kAlphaS = 1e-6f, kAlphaL = 2e-6f, kAlphaR = 0.0f;

double w_ohc() { return (double)(2 * kPI * 1000.0f); }                  // W_OHC
double dx() { return (double)(kCochleaLen / NS); }
double bc0() { return (double)((2.0f * kRho * kCow) / kSigmaOW); }        // BC_0
double bc1() {                                                            // BC_1
  const float wow2 = 2 * kPI * 1500 * 2 * kPI * 1500;
  const float cme_gme = (2 * kPI * 1340) * (2 * kPI * 1340) * 0.059f / (0.49f * 1.4f) * kGme;
  return (double)(kSigmaOW * wow2 + cme_gme);
}
double bc2() { return (double)(kSigmaOW * kGammaOW); }                     // BC_2
double eta1() { return (double)((1.0f) / kAlphaL * 2 * kPI * 1000.0f); }  // ETA_1
}  // namespace

extern "C" {

int cm_model_vectors(double *out) {
  if (!out) return -3;
  const double d = dx();
  for (int i = 0; i < NS; ++i) {
    const double M = (double)kM0 * std::exp((double)kM1 * d * i);
    const double R = (double)kR0 * std::exp((double)kR1 * d * i);
    const double S = (double)kS0 * std::exp((double)kS1 * d * i);
    const double wcf2 = S / M;
    const double Stm = R * wcf2 / w_ohc();
    out[0 * NS + i] = M;
    out[1 * NS + i] = R;
    out[2 * NS + i] = S;
    out[3 * NS + i] = R;                        // R_tm = R_bm
    out[4 * NS + i] = Stm;
    out[5 * NS + i] = (double)kRatioK * Stm;    // S_ohc
    out[6 * NS + i] = 2 * (double)kRho * (double)kBeta / ((double)kArea * M);   // Q
  }
  return 0;
}

int cm_tri_matrix(double *lower, double *mid, double *upper) {
  double v[7 * NS];
  cm_model_vectors(v);
  const double d = dx();
  for (int i = 0; i < NS; ++i) {
    if (mid) mid[i] = -1.0 * (2.0 + d * d * v[6 * NS + i]);
    if (upper) upper[i] = 1.0;
    if (lower) lower[i] = 1.0;
  }
  if (mid) {
    mid[0] = -1.0 * (1.0 + d * bc0());
    mid[NS - 1] = 1.0;
  }
  if (lower) {
    lower[NS - 1] = 0.0;
    lower[NS - 2] = 0.0;
  }
  return 0;
}

// CM_SEC_COUNT x 256 floats: mass, 1/m_i, R, S, Q, S_ohc, S_tm, R_tm (src/solver.cpp:357-376)
// and the diagonal m_i itself.
int cm_pack_section_params(float *out) {
  if (!out) return -3;
  double v[7 * NS], m[NS];
  cm_model_vectors(v);
  cm_tri_matrix(nullptr, m, nullptr);
  for (int i = 0; i < NS; ++i) {
    out[CM_SEC_MASS * NS + i] = (float)v[0 * NS + i];
    out[CM_SEC_RSM * NS + i] = 1.0f / (float)m[i];
    out[CM_SEC_R * NS + i] = (float)v[1 * NS + i];
    out[CM_SEC_S * NS + i] = (float)v[2 * NS + i];
    out[CM_SEC_Q * NS + i] = (float)v[6 * NS + i];
    out[CM_SEC_SOHC * NS + i] = (float)v[5 * NS + i];
    out[CM_SEC_STM * NS + i] = (float)v[4 * NS + i];
    out[CM_SEC_RTM * NS + i] = (float)v[3 * NS + i];
    // the Jacobi residual's diagonal, defined from the 1/m row as the
    // reference's BMOHCKernel_Init interface hands it over (float 1/m)
    out[CM_SEC_M * NS + i] = (float)(1.0 / (double)out[CM_SEC_RSM * NS + i]);
  }
  return 0;
}

// Scalars of the BM launch for sample rate Fs (src/solver.cpp:437-438, 629-659;
// CParams::cudaIterationsPerKernel, include/params.h:378).
int cm_default_bm_scalars(float Fs, cm_bm_scalars *o) {
  if (!o || !(Fs > 0)) return -3;
  std::memset(o, 0, sizeof(*o));
  const float dt_max = 1e-6f, dt_min = 1e-18f, dt0 = 1e-6f;
  o->w_ohc = (float)w_ohc();
  o->time_step = dt0;
  o->time_step_out = dt0 * (float)(int)(1.0f / (Fs * dt_max));
  o->delta_x = (float)dx();
  o->alpha_r = kAlphaR;
  o->Ts = (float)(1.0 / (double)Fs);
  o->alpha_l = kAlphaL;
  o->alpha_s = kAlphaS;
  o->Gme = kGme;
  o->a0 = (float)bc0();
  o->a1 = (float)bc1();
  o->a2 = (float)bc2();
  o->sigma_ow = kSigmaOW;
  o->eta1 = (float)eta1();
  o->eta2 = 0.0f;
  o->dt_min = dt_min;
  o->dt_max = dt_max;
  o->j_fast = 2;
  o->j_slow = 20;
  o->max_passes = 10;
  o->decision_mode = 0;
  o->jacobi_mode = 0;
  o->check_every = 4;
  o->max_sweeps = 1000;
  o->tol_rel = 1e-4f;
  o->warm_start = 0;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- JND ---
// Aggregated JND of complex (Sim_Type 4) runs: ComplexJNDProfile
// (src/ComplexJNDProfile.cpp:19-95) driven by AudioGramCreator::calcComplexJND
// (src/AudioGramCreator.cpp:548-571).  Profile p = noise * n_freq + freq
// collects the raw JND of its n_levels tested levels in level order.
extern "C" int cm_complex_jnd_intervals(int n_freq, int n_levels, int n_noises, int32_t *intervals) {
  // CParams::updateInputProfile (src/params.cpp:693-735): calculated intervals
  // run noise -> level -> frequency, references are not counted
  if (n_freq <= 0 || n_levels <= 0 || n_noises <= 0 || !intervals) return -3;
  for (int nl = 0; nl < n_noises; ++nl)
    for (int pl = 0; pl < n_levels; ++pl)
      for (int fl = 0; fl < n_freq; ++fl)
        intervals[(nl * n_freq + fl) * n_levels + pl] = (nl * n_levels + pl) * n_freq + fl;
  return 0;
}

extern "C" int cm_complex_jnd(const double *raw, const int32_t *failed, int n_raw, const int32_t *intervals,
                              int n_profiles, int n_levels, int methods, float eps, float *jnd_final,
                              int32_t *warnings, int32_t *conv_warnings) {
  if (!raw || !intervals || n_profiles <= 0 || n_levels <= 0 || !jnd_final || !warnings || !conv_warnings ||
      (methods & 3) == 0)
    return -3;
  const float kNotFound = 1111.0f;   // MIN_INF_POWER_LEVEL, include/const.h:342
  int m = 0;
  for (int bit = 0; bit < 2; ++bit) {
    if (!(methods & (1 << bit))) continue;
    for (int p = 0; p < n_profiles; ++p) {
      const int32_t *iv = intervals + (size_t)p * n_levels;
      for (int l = 0; l < n_levels; ++l)
        if (iv[l] < 0 || iv[l] >= n_raw) return -3;
      auto v = [&](int l) { return raw[iv[l]]; };
      auto fl = [&](int l) { return failed ? failed[iv[l]] : 0; };
      float value = kNotFound;
      int warn = 0, conv = 0;
      if (bit == 0) {
        // last strict interior local minimum over the levels
        int last = -1;
        for (int l = 1; l + 1 < n_levels; ++l)
          if (v(l) < v(l - 1) && v(l) < v(l + 1)) last = l;
        if (last >= 0) {
          value = (float)v(last);
          conv = fl(last);
        } else {
          warn = 1;
          // boundary minimum; the reference's convergence flag read here indexes an
          // empty vector, the chosen interval's own flag is used instead
          if (n_levels == 1) {
            value = (float)v(0);
            conv = fl(0);
          } else if (v(0) < v(1)) {
            value = (float)v(0);
            conv = fl(0);
          } else if (v(n_levels - 1) < v(n_levels - 2)) {
            value = (float)v(n_levels - 1);
            conv = fl(n_levels - 1);
          }
        }
      } else {
        // first level whose successor is larger by more than eps
        int at = -1;
        for (int l = 0; l + 1 < n_levels; ++l)
          if (v(l + 1) - v(l) > (double)eps) {
            at = l;
            break;
          }
        if (at >= 0) {
          value = (float)v(at);
        } else {
          value = (float)v(0);
          warn = 1;
        }
        // the reference leaves the profile's convergence flag as the minimum
        // method set it (0 when that method did not run)
        conv = (methods & 1) ? conv_warnings[p] : 0;
      }
      jnd_final[(size_t)m * n_profiles + p] = value;
      warnings[(size_t)m * n_profiles + p] = warn;
      conv_warnings[(size_t)m * n_profiles + p] = conv;
    }
    ++m;
  }
  return 0;
}
