// AC (low-pass) filter design of the IHC stage, host C++.
//
// The reference designs the AC filter at run time from its Function_Filter
// string (SubModel::analyzeACFilter, src/SubModel.cpp:153-360) with the firpm
// library (src/firpm/*, GPL, needs Eigen).  This file has its own
// Parks-McClellan design: the Remez exchange on a dense frequency grid
// (McClellan, Parks & Rabiner, "A computer program for designing optimum FIR
// linear phase digital filters", IEEE Trans. Audio Electroacoust. 21(6), 1973:
// grid of 16 points per basis function, barycentric Lagrange interpolation in
// x = cos(2 pi f), alternation exchange, coefficients by an inverse DFT of the
// amplitude response), and around it the reference's equiripple procedure:
// minimum order (Oppenheim & Schafer eq. 7.104, SubModel::calcMinimumOrder),
// band weights from the ripples, DC-gain normalisation (fixDCGain), the -Apass
// corner correction (FIRFreqError, src/firpm/pm.cpp:742-793) and a redesign,
// the tap search (Fpass/Fstop) or transition-width search (Fc with Astop) that
// brings the stop-band ripple within a factor 3 of the Astop target, then unit
// FIR gain (HFunction::setFIRGain(0), src/HFunction.cpp:43-46).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "cm_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Grid {
  std::vector<double> f, x, des, wt;
};

// Dense grid over the pass band [0, fp] and the stop band [fs, 0.5] (cycles per
// sample): steps of 0.5 / (density * nfcns) from each band's lower edge, the
// last point of a band moved onto its upper edge.  For even lengths the
// amplitude carries a cos(pi f) factor: desired / cos, weight * cos, and no
// point at the Nyquist frequency where that factor vanishes.
Grid make_grid(int ntaps, int nfcns, double fp, double fs, double wpass, double wstop, int density) {
  Grid g;
  const double delf = 0.5 / (density * (double)nfcns);
  const double edges[2][2] = {{0.0, fp}, {fs, 0.5}};
  const double des[2] = {1.0, 0.0}, wt[2] = {wpass, wstop};
  for (int b = 0; b < 2; ++b) {
    const double lo = edges[b][0], hi = edges[b][1];
    const size_t first = g.f.size();
    for (int k = 0;; ++k) {
      const double f = lo + k * delf;
      if (f > hi) break;
      g.f.push_back(f);
    }
    if (g.f.size() == first) g.f.push_back(hi);
    g.f.back() = hi;
    for (size_t i = first; i < g.f.size(); ++i) {
      g.des.push_back(des[b]);
      g.wt.push_back(wt[b]);
    }
  }
  const bool even = ntaps % 2 == 0;
  if (even && g.f.back() > 0.5 - delf) {
    g.f.pop_back();
    g.des.pop_back();
    g.wt.pop_back();
  }
  g.x.resize(g.f.size());
  for (size_t i = 0; i < g.f.size(); ++i) {
    g.x[i] = std::cos(2.0 * kPi * g.f[i]);
    if (even) {
      const double c = std::cos(kPi * g.f[i]);
      g.des[i] /= c;
      g.wt[i] *= c;
    }
  }
  return g;
}

// barycentric weights 1 / prod_{j != i} 2 (x_i - x_j) of the nodes x[0..n)
std::vector<double> bary_weights(const std::vector<double> &x, int n) {
  std::vector<double> w(n);
  for (int i = 0; i < n; ++i) {
    double p = 1.0;
    for (int j = 0; j < n; ++j)
      if (j != i) p *= 2.0 * (x[i] - x[j]);
    w[i] = 1.0 / p;
  }
  return w;
}

struct Interp {
  std::vector<double> x, c, w;   // nodes, values, barycentric weights
  double eval(double xv) const {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
      const double d = xv - x[i];
      if (d == 0.0) return c[i];
      const double t = w[i] / d;
      num += t * c[i];
      den += t;
    }
    return num / den;
  }
};

// sin(x)/x, 1 near 0 (|x| < 1e-5), as the Iowa Hills routines define it
double sinc(double x) { return (x > -1.0e-5 && x < 1.0e-5) ? 1.0 : std::sin(x) / x; }

// the Kaiser window's I0 as the reference evaluates it: the power series
// 1 + sum_{i=1..9} ((x/2)^i / i!)^2, truncated after nine terms
// (src/IowaHillsFilters/FFTCode.cpp:771-783)
double bessel_i0_9(double x) {
  double sum = 0.0, term = 1.0;
  for (int i = 1; i < 10; ++i) {
    term *= (x / 2.0) / i;
    sum += term * term;
  }
  return 1.0 + sum;
}

// rectangular-window low-pass: omega_c sinc(omega_c pi (j - (n-1)/2))
// (RectWinFIR, firLPF, src/IowaHillsFilters/FIRFilterCode.cpp:29-42)
void rect_lowpass(std::vector<double> &h, double omega_c) {
  const int n = (int)h.size();
  for (int j = 0; j < n; ++j) h[j] = omega_c * sinc(omega_c * (j - (n - 1) / 2.0) * kPi);
}

enum class Win { None, Kaiser, KaiserBessel, Sinc, Hanning, Hamming, Blackman, Flattop };

// FIRFilterWindow (src/IowaHillsFilters/FIRFilterCode.cpp:108-150): only the
// Kaiser and Sinc (Lanczos) windows are applied there; the other names leave
// the coefficients unwindowed.  beta is clamped to [0, 10].
void apply_window(std::vector<double> &h, Win w, double beta) {
  if (w != Win::Kaiser && w != Win::Sinc) return;
  beta = std::min(10.0, std::max(0.0, beta));
  const int n = (int)h.size();
  const double dn = n + 1.0;
  std::vector<double> c(n);
  for (int j = 0; j < n; ++j) {
    if (w == Win::Kaiser) {
      const double t = ((double)(2 * j + 2) - dn) / dn;
      c[j] = bessel_i0_9(beta * std::sqrt(1.0 - t * t)) / bessel_i0_9(beta);
    } else {
      c[j] = std::pow(sinc((double)(2 * j + 1 - n) / dn * kPi), beta);
    }
  }
  for (int j = 0; j < n / 2; ++j) c[n - j - 1] = c[j];
  for (int j = 0; j < n; ++j) h[j] *= c[j];
}

}  // namespace

extern "C" {

// Parks-McClellan low-pass of ntaps coefficients: pass band [0, fpass], stop
// band [fstop, 0.5] in cycles per sample, weights wpass / wstop, grid density
// (16 in the classic program).  h: ntaps outputs; delta (may be NULL): the
// final weighted ripple; iterations (may be NULL).
int cm_remez_lowpass(int ntaps, double fpass, double fstop, double wpass, double wstop, int density, double *h,
                     double *delta, int *iterations) {
  if (ntaps < 3 || ntaps > 4096 || !h || !(fpass > 0.0) || !(fstop > fpass) || !(fstop < 0.5) || !(wpass > 0.0) ||
      !(wstop > 0.0) || density < 1) {
    cm::set_error("cm_remez_lowpass: bad arguments (ntaps %d, fpass %g, fstop %g)", ntaps, fpass, fstop);
    return -3;
  }
  const bool even = ntaps % 2 == 0;
  const int nfcns = even ? ntaps / 2 : ntaps / 2 + 1;   // cosine basis functions
  const int r = nfcns + 1;                              // extremal points
  const Grid g = make_grid(ntaps, nfcns, fpass, fstop, wpass, wstop, density);
  const int ng = (int)g.f.size();
  if (ng < 2 * r) {
    cm::set_error("cm_remez_lowpass: grid of %d points too small for %d extremals", ng, r);
    return -3;
  }
  std::vector<int> ext(r);
  for (int j = 0; j < nfcns; ++j) ext[j] = (int)((int64_t)j * (ng - 1) / nfcns);
  ext[nfcns] = ng - 1;
  std::vector<double> err(ng);
  Interp P;
  double dev = 0.0;
  int it = 0;
  for (; it < 250; ++it) {
    // deviation and interpolant through the current extremal set
    std::vector<double> xe(r);
    for (int i = 0; i < r; ++i) xe[i] = g.x[ext[i]];
    const std::vector<double> gam = bary_weights(xe, r);
    double num = 0.0, den = 0.0;
    for (int i = 0; i < r; ++i) {
      num += gam[i] * g.des[ext[i]];
      den += ((i & 1) ? -1.0 : 1.0) * gam[i] / g.wt[ext[i]];
    }
    dev = num / den;
    P.x.assign(xe.begin(), xe.begin() + nfcns);
    P.c.resize(nfcns);
    for (int i = 0; i < nfcns; ++i) P.c[i] = g.des[ext[i]] - ((i & 1) ? -1.0 : 1.0) * dev / g.wt[ext[i]];
    P.w = bary_weights(P.x, nfcns);
    for (int j = 0; j < ng; ++j) err[j] = g.wt[j] * (g.des[j] - P.eval(g.x[j]));
    // local extrema of the weighted error with |E| >= |dev|, alternating in sign
    // (the error at a reference point equals |dev| only up to the rounding of
    // D - (D - dev / W), hence the slack)
    const double floor_abs = std::fabs(dev) * (1.0 - 1e-9) - 1e-14;
    std::vector<int> cand;
    for (int j = 0; j < ng; ++j) {
      const double e = err[j];
      if (std::fabs(e) < floor_abs) continue;
      const bool ge_l = j == 0 || (e > 0 ? e >= err[j - 1] : e <= err[j - 1]);
      const bool ge_r = j == ng - 1 || (e > 0 ? e >= err[j + 1] : e <= err[j + 1]);
      if (ge_l && ge_r) cand.push_back(j);
    }
    std::vector<int> alt;
    for (int j : cand) {
      if (!alt.empty() && (err[alt.back()] > 0) == (err[j] > 0)) {
        if (std::fabs(err[j]) > std::fabs(err[alt.back()])) alt.back() = j;
      } else {
        alt.push_back(j);
      }
    }
    while ((int)alt.size() > r) {
      if (std::fabs(err[alt.front()]) < std::fabs(err[alt.back()])) alt.erase(alt.begin());
      else alt.pop_back();
    }
    if ((int)alt.size() < r) break;   // the current set is already the best alternation on this grid
    double emax = 0.0, emin = 1e300;
    for (int j : alt) {
      emax = std::max(emax, std::fabs(err[j]));
      emin = std::min(emin, std::fabs(err[j]));
    }
    const bool same = alt == ext;
    ext = alt;
    if (same || (emax - emin) <= 1e-13 * emax) {
      ++it;
      break;
    }
  }
  // amplitude at f = m / ntaps, then the inverse DFT of a linear-phase response
  const int half = (ntaps - 1) / 2;
  std::vector<double> A(half + 1);
  for (int m = 0; m <= half; ++m) {
    const double f = (double)m / ntaps;
    double a = P.eval(std::cos(2.0 * kPi * f));
    if (even) a *= std::cos(kPi * f);
    A[m] = a;
  }
  const double mid = 0.5 * (ntaps - 1);
  for (int n = 0; n < ntaps; ++n) {
    double s = A[0];
    for (int m = 1; m <= half; ++m) s += 2.0 * A[m] * std::cos(2.0 * kPi * m * (n - mid) / ntaps);
    h[n] = s / ntaps;
  }
  if (delta) *delta = std::fabs(dev);
  if (iterations) *iterations = it;
  return 0;
}

// |H(e^{j pi omega})| of an FIR, omega in [0, 1] (Nyquist = 1)
static double fir_gain(const std::vector<double> &h, double omega) {
  double re = 0.0, im = 0.0;
  for (size_t n = 0; n < h.size(); ++n) {
    re += h[n] * std::cos(kPi * omega * n);
    im -= h[n] * std::sin(kPi * omega * n);
  }
  return std::sqrt(re * re + im * im);
}

// FIRFreqError (src/firpm/pm.cpp:742-793), low-pass: the last 1/1000-Nyquist
// step whose gain exceeds -corner_db before the response falls below -20 dB;
// returns the corrected cut-off 2 omega_c - that frequency, clamped.
static double corner_correction(const std::vector<double> &h, double omega_c, double corner_db) {
  const double target = std::pow(10.0, -corner_db / 20.0);
  int j3 = 10;
  for (int j = 0; j < 1000; ++j) {
    const double mag = fir_gain(h, j / 1000.0);
    if (mag > target) j3 = j;
    if (mag < 0.1) break;
  }
  return std::min(0.99, std::max(0.001, omega_c * 2.0 - j3 / 1000.0));
}

// Window branch of SubModel::analyzeACFilter (src/SubModel.cpp:366-457).
static int design_window(double Fs, std::map<std::string, std::string> &kv, double *h, int max_taps, int *n_taps) {
  if (!kv.count("windowtype")) {
    cm::set_error("cm_design_ac_filter: FIR filter with Window must have window type");
    return -3;
  }
  const std::string wn = kv["windowtype"];
  const Win w = wn == "kaiser" ? Win::Kaiser : wn == "kaiserbessel" ? Win::KaiserBessel : wn == "sinc" ? Win::Sinc
              : wn == "hanning" ? Win::Hanning : wn == "hamming" ? Win::Hamming : wn == "blackmann" ? Win::Blackman
              : wn == "flattop" ? Win::Flattop : Win::None;
  auto num = [&](const char *k, double def) { return kv.count(k) ? std::atof(kv[k].c_str()) : def; };
  const double nyq = (double)(float)(Fs / 2);
  int taps = kv.count("order") ? std::atoi(kv["order"].c_str()) : -1;
  const double Apass = kv.count("apass") ? (double)(float)num("apass", 3.0) : 3.0;
  const double Astop = kv.count("astop") ? (double)(float)num("astop", 0.0) : 0.0;
  double omega_c, beta = 0.0;
  if (kv.count("fc")) {
    if (!kv.count("order")) {
      cm::set_error("cm_design_ac_filter: FIR filter with Fc must get specific Order");
      return -3;
    }
    omega_c = (double)(float)num("fc", 800.0) / nyq;
  } else {
    if (!kv.count("fpass") || !kv.count("fstop")) {
      cm::set_error("cm_design_ac_filter: FIR filter without Fc must have Fpass,Fstop");
      return -3;
    }
    const float fp = (float)num("fpass", 600.0), fs = (float)num("fstop", 1600.0);
    omega_c = (double)(fp + fs) / (2 * nyq);
    const double tw = (fs - fp) / nyq;
    switch (w) {
      case Win::None: taps = (int)(4 * kPi / tw); break;
      case Win::Hamming:
      case Win::Hanning: taps = (int)(8 * kPi / tw); break;
      case Win::Blackman: taps = (int)(12 * kPi / tw); break;
      case Win::Kaiser: {
        if (Astop == 0 && !kv.count("apass")) {
          cm::set_error("cm_design_ac_filter: Kaiser window must have Astop or Apass");
          return -3;
        }
        const double twp = tw * kPi, a = std::max(Apass, Astop);
        taps = (int)std::ceil(a > 21 ? (a - 7.95) / (2.285 * twp) : 5.79 / twp);
        if (kv.count("beta")) beta = num("beta", 0.0);
        else if (a > 50) beta = 0.1102 * (a - 8.7);
        // the reference raises (A - 21) to the 4th power here (Kaiser's rule has 0.4)
        else if (a > 21) beta = 0.5842 * std::pow(a - 21, 4) + 0.07886 * (a - 21);
        break;
      }
      default: break;   // Sinc, Flattop, KaiserBessel: the order must be given
    }
    if (kv.count("order")) taps = std::atoi(kv["order"].c_str());
  }
  if (taps < 3 || taps > max_taps) {
    cm::set_error("cm_design_ac_filter: window design of %d taps (give Order; room for %d)", taps, max_taps);
    return -3;
  }
  std::vector<double> b(taps);
  auto design = [&](double wc) {
    rect_lowpass(b, wc);
    apply_window(b, w, beta);
    const double dc = fir_gain(b, 0.0);   // fixDCGain
    for (double &x : b) x /= dc;
  };
  design(omega_c);
  design(corner_correction(b, omega_c, 3.0));
  double sum = 0.0;
  for (double x : b) sum += x;
  for (int i = 0; i < taps; ++i) h[i] = b[i] / sum;
  *n_taps = taps;
  return 0;
}

// polynomial product (coefficients in ascending powers of z^-1)
static std::vector<double> poly_mul(const std::vector<double> &p, const std::vector<double> &q) {
  std::vector<double> r(p.size() + q.size() - 1, 0.0);
  for (size_t i = 0; i < p.size(); ++i)
    for (size_t j = 0; j < q.size(); ++j) r[i + j] += p[i] * q[j];
  return r;
}

// Butterworth branch of SubModel::analyzeACFilter (src/SubModel.cpp:458-503):
// order ceil(log(sqrt(Gs^2 - 1) / eps) / log(ws / wp)) with eps^2 = Gp^2 - 1,
// cut-off wp eps^(-1/N) (pass-band mode) or ws (eps / sqrt(Gs^2 - 1))^(1/N)
// (stop-band mode); the analog prototype's poles exp(j pi (2k + N + 1) / 2N)
// taken most negative real part first, one 1st- or 2nd-order factor each,
// mapped by the bilinear transform with s = T (1 - z^-1) / (1 + z^-1),
// T = 2 tan(pi omega_c / 2), each factor with unit DC gain
// (CalcIIRFilterCoeff, src/IowaHillsFilters/IIRFilterCode.cpp:32-140), and the
// factors multiplied as 5-coefficient sections (HFunction products,
// src/HFunction.cpp:91-126), so N poles give 4 ceil(N/2) + 1 coefficients.
static int design_butterworth(double Fs, std::map<std::string, std::string> &kv, double *b_out, double *a_out,
                              int max_len, int *n_out) {
  auto num = [&](const char *k, double def) { return kv.count(k) ? std::atof(kv[k].c_str()) : def; };
  const double nyq = (double)(float)(Fs / 2);
  const double Apass = (float)num("apass", 3.0), Astop = (float)num("astop", 60.0);
  const double wp = (double)(float)num("fpass", 600.0) / nyq, ws = (double)(float)num("fstop", 1600.0) / nyq;
  const double h0 = std::pow(10.0, Apass / 20.0), gs = std::pow(10.0, Astop / 20.0);
  const double eps = std::sqrt(h0 * h0 - 1.0);
  if (!(eps > 0.0) || !(gs > 1.0) || !(ws > wp) || !(wp > 0.0)) {
    cm::set_error("cm_design_ac_filter: Butterworth needs Apass > 0, Astop > 0 and Fstop > Fpass > 0");
    return -3;
  }
  const int order = (int)std::ceil(std::log(std::sqrt(gs * gs - 1.0) / eps) / std::log(ws / wp));
  const bool stopband = kv.count("mode") && kv["mode"].find("stopband") != std::string::npos;
  const double wc = stopband ? ws * std::pow(eps / std::sqrt(gs * gs - 1.0), 1.0 / order)
                             : wp * std::pow(eps, -1.0 / order);
  const int sections = (order + 1) / 2;
  if (order < 1 || order > 40 || 4 * sections + 1 > max_len) {
    cm::set_error("cm_design_ac_filter: Butterworth order %d out of range", order);
    return -3;
  }
  // prototype poles, most negative real part first
  std::vector<std::pair<double, double>> poles;
  for (int j = 0; j < order / 2; ++j) {
    const double th = kPi * (double)(2 * j + order + 1) / (double)(2 * order);
    poles.push_back({std::cos(th), std::sin(th)});
  }
  if (order % 2) poles.push_back({-1.0, 0.0});
  std::stable_sort(poles.begin(), poles.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
  const double T = 2.0 * std::tan(wc * kPi / 2.0);
  std::vector<double> B{1.0}, A{1.0};
  for (const auto &p : poles) {
    std::vector<double> nb(5, 0.0), na(5, 0.0);
    na[0] = 1.0;
    if (p.second == 0.0) {   // real pole: s + |re|
      const double Bc = 1.0, Cc = -p.first, arg = 2.0 * Bc + Cc * T;
      na[1] = (-2.0 * Bc + Cc * T) / arg;
      nb[0] = Cc * T / arg;
      nb[1] = Cc * T / arg;
    } else {                 // pair: s^2 - 2 re s + |p|^2
      const double Ac = 1.0, Bc = -2.0 * p.first, Cc = p.first * p.first + p.second * p.second;
      const double arg = 4.0 * Ac + 2.0 * Bc * T + Cc * T * T;
      na[1] = (2.0 * Cc * T * T - 8.0 * Ac) / arg;
      na[2] = (4.0 * Ac - 2.0 * Bc * T + Cc * T * T) / arg;
      nb[0] = Cc * T * T / arg;
      nb[1] = 2.0 * Cc * T * T / arg;
      nb[2] = Cc * T * T / arg;
    }
    B = B.size() == 1 ? nb : poly_mul(B, nb);
    A = A.size() == 1 ? na : poly_mul(A, na);
  }
  for (size_t i = 0; i < B.size(); ++i) {
    b_out[i] = B[i];
    a_out[i] = A[i];
  }
  *n_out = (int)B.size();
  return 0;
}

// AC low-pass from a Function_Filter string ("type=equiripple,fpass=600,
// fstop=1600,apass=3,astop=60", keys case-insensitive, as parsed by
// src/params.cpp:1091-1170).  Types: equiripple (fpass/fstop, order or minimum
// order, apass/astop or wpass/wstop, mode=constrained), window (windowtype,
// fc+order or fpass/fstop, astop/apass, beta), butterworth (fpass, fstop,
// apass, astop, mode=stopband).  h: numerator, a: denominator (Butterworth
// only; may be NULL otherwise), up to max_taps each; n_taps: the length;
// is_iir: 1 for the Butterworth IIR.
int cm_design_ac_filter(double Fs, const char *function_filter, double *h, double *a, int max_taps, int *n_taps,
                        int *is_iir) {
  if (!function_filter || !h || !n_taps || !is_iir || !(Fs > 0.0)) {
    cm::set_error("cm_design_ac_filter: bad arguments");
    return -3;
  }
  *is_iir = 0;
  // key=value pairs ([a-zA-Z0-9_]+=[a-zA-Z0-9_.]+), keys and values lower-cased
  std::map<std::string, std::string> kv;
  const std::string s(function_filter);
  auto word = [](char c) { return std::isalnum((unsigned char)c) || c == '_'; };
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] != '=') continue;
    size_t ks = i;
    while (ks > 0 && word(s[ks - 1])) --ks;
    size_t ve = i + 1;
    while (ve < s.size() && (word(s[ve]) || s[ve] == '.')) ++ve;
    if (ks == i || ve == i + 1) continue;
    std::string k = s.substr(ks, i - ks), v = s.substr(i + 1, ve - i - 1);
    for (char &c : k) c = (char)std::tolower((unsigned char)c);
    for (char &c : v) c = (char)std::tolower((unsigned char)c);
    kv.insert({k, v});
  }
  auto num = [&](const char *k, double def) { return kv.count(k) ? std::atof(kv[k].c_str()) : def; };
  const std::string type = kv.count("type") ? kv["type"] : "equiripple";
  if (type == "window") return design_window(Fs, kv, h, max_taps, n_taps);
  if (type == "butterworth") {
    if (!a) {
      cm::set_error("cm_design_ac_filter: a Butterworth design needs the denominator array");
      return -3;
    }
    *is_iir = 1;
    return design_butterworth(Fs, kv, h, a, max_taps, n_taps);
  }
  if (type != "equiripple") {
    cm::set_error("cm_design_ac_filter: filter type '%s' is not supported", type.c_str());
    return -3;
  }
  // equiripple branch of SubModel::analyzeACFilter (src/SubModel.cpp:167-360)
  const bool has_fc = kv.count("fc") > 0;
  if (has_fc && (kv.count("fpass") || kv.count("fstop"))) {   // src/params.cpp:1136-1140
    cm::set_error("cm_design_ac_filter: Fc cannot be given together with Fpass/Fstop");
    return -3;
  }
  if (!has_fc && (!kv.count("fpass") || !kv.count("fstop"))) {
    cm::set_error("cm_design_ac_filter: equiripple needs either Fpass,Fstop or Fc");
    return -3;
  }
  const double nyq = (double)(float)(Fs / 2);
  const double Fpass = (float)num("fpass", 600.0), Fstop = (float)num("fstop", 1600.0);
  double omega_c, pw = 0.0, max_fstop = 0.0, min_fstop = 0.0;
  if (has_fc) {
    omega_c = (double)(float)num("fc", 0.0) / nyq;
  } else {
    omega_c = Fpass / nyq;
    pw = (Fstop - Fpass) / nyq;
    max_fstop = Fpass + nyq * 0.3;
    min_fstop = Fpass + nyq * 0.01;
  }
  int order = kv.count("order") && kv["order"] != "minimum" ? std::atoi(kv["order"].c_str()) : -1;
  const bool min_order = order == -1 || (kv.count("minorder") && std::atoi(kv["minorder"].c_str()) == 1);
  const bool deltas_found = kv.count("astop") || kv.count("apass") ||
                            (kv.count("mode") && kv["mode"].find("constrained") != std::string::npos);
  // Apass is 1 dB for the ripple when only Astop is given; the first corner
  // correction still aims at -3 dB, the ones inside the searches at -Apass
  // (src/SubModel.cpp:210-213, 273, 317, 339)
  const double Apass = (float)num("apass", 1.0);
  const double corner_db = kv.count("apass") ? Apass : 3.0;
  double dp, ds;
  if (kv.count("wstop")) {
    dp = kv.count("wpass") ? (float)num("wpass", 10.0) : 10.0;
    ds = (float)num("wstop", 100.0);
  } else if (kv.count("astop")) {
    dp = std::pow(10.0, -Apass / 20.0);
    ds = std::pow(10.0, -(double)(float)num("astop", 60.0) / 20.0);
  } else if (min_order) {
    cm::set_error("cm_design_ac_filter: minimum order needs Wstop or Astop");
    return -3;
  } else {
    dp = 1.0;
    ds = 10.0;
  }
  // Fc with Astop: the transition width is searched, from 0.15 between
  // MIN_PARK_WIDTH 0.01 and MAX_PARK_WIDTH 0.3 (include/const.h:344-345)
  const bool search_pw = has_fc && kv.count("astop");
  if (search_pw) pw = 0.15;
  if (min_order) {
    if (!(pw > 0.0)) {
      cm::set_error("cm_design_ac_filter: a minimum order needs a transition band (Fc needs Order or Astop)");
      return -3;
    }
    order = (int)std::ceil((-10.0 * std::log10(dp * ds) - 13.0) / (2.324 * pw * kPi));
  }
  double w_stop = 10.0;
  double w_pass = std::max(0.01, 10.0 * ds / dp);
  if (w_pass > 10.0 * ds / dp) w_stop *= w_pass / (10.0 * ds / dp);
  if (kv.count("wstop")) w_stop = (float)num("wstop", 100.0);
  if (kv.count("wpass")) w_pass = (float)num("wpass", 1.0);
  if (!has_fc && (Fstop > max_fstop || Fstop < min_fstop)) {
    cm::set_error("cm_design_ac_filter: Fstop should be between %g and %g", min_fstop, max_fstop);
    return -3;
  }
  // One design of `order` (firpm's degree n gives n + 1 coefficients) at cut-off
  // wc and transition width; rip: the achieved stop-band ripple of the
  // DC-normalised result (the exchange's equiripple deviation over the stop
  // weight).  A zero width (Fc without Astop) starts the stop band one grid
  // step above the pass band edge.
  std::vector<double> b;
  double rip = 0.0;
  auto design = [&](int n, double wc, double width) {
    const int taps = n + 1;
    if (n < 2 || taps > max_taps) {
      cm::set_error("cm_design_ac_filter: %d taps (room for %d)", taps, max_taps);
      return -3;
    }
    b.assign(taps, 0.0);
    const int nfcns = taps % 2 == 0 ? taps / 2 : taps / 2 + 1;
    const double fp = 0.5 * wc;
    const double fs = width > 0.0 ? 0.5 * (wc + width) : fp + 0.5 / (16.0 * nfcns);
    double dev = 0.0;
    if (cm_remez_lowpass(taps, fp, fs, w_pass, w_stop, 16, b.data(), &dev, nullptr)) return -3;
    const double dc = fir_gain(b, 0.0);   // fixDCGain(pmo, 1)
    for (double &x : b) x /= dc;
    rip = dev / w_stop / dc;
    return 0;
  };
  if (design(order, omega_c, pw)) return -3;
  const double omega_c0 = omega_c;
  if (deltas_found) {
    omega_c = corner_correction(b, omega_c, corner_db);
    if (design(order, omega_c, pw)) return -3;
  }
  // each search step: the design at the locked cut-off, its corner correction
  // and the redesign (src/SubModel.cpp:314-320, 336-342)
  auto redesign = [&]() {
    if (design(order, omega_c0, pw)) return -3;
    omega_c = corner_correction(b, omega_c0, Apass);
    return design(order, omega_c, pw);
  };
  // The searches stop once the stop-band ripple is within a factor 3 of the
  // Astop target.  The reference tests firpm's convergence measure Q there
  // (pmo.Q, src/firpm/pm.cpp:602-603), which is not a ripple and depends on
  // the exchange's iteration history; the achieved ripple is the quantity the
  // comparisons and the 10 log10 tap rule are written for.
  if (search_pw) {                               // src/SubModel.cpp:280-324
    double pw_high = 0.3, pw_low = 0.01;
    for (int tests = 0; (rip > 3 * ds || 3 * rip < ds) && tests < 10; ++tests) {
      if (min_order && (rip > 10 * ds || 10 * rip < ds)) {
        order += (int)std::ceil(10 * std::log10(rip / ds));
      } else if (rip > 3 * ds) {                 // too little attenuation: widen the transition
        pw_low = pw;
        pw = (pw + pw_high) / 2;
      } else {                                   // more than needed: narrow it
        pw_high = pw;
        pw = (pw + pw_low) / 2;
      }
      if (redesign()) return -3;
    }
  } else if (!has_fc && !kv.count("wstop")) {    // src/SubModel.cpp:325-346
    double fix = 10, prev = rip;
    for (int tests = 0; (rip > 3 * ds || 3 * rip < ds) && tests < 5; ++tests) {
      if ((prev > ds && rip < ds) || (prev < ds && rip > ds)) fix *= 0.5;
      prev = rip;
      order += (int)std::ceil(fix * std::log10(rip / ds));
      if (redesign()) return -3;
    }
  }
  const int taps = (int)b.size();
  double sum = 0.0;   // setFIRGain(0): unit gain at DC
  for (double x : b) sum += x;
  for (int i = 0; i < taps; ++i) h[i] = b[i] / sum;
  *n_taps = taps;
  return 0;
}

}  // extern "C"
