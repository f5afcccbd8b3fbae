// cochlea_run -- command-line host of the VOICE path with the reference's
// solver/params API, calling libcochlea_amd.so through its C-ABI.
//
//   cochlea_run <file.par>           run and write Output_Result
//   cochlea_run --parse <file.par>   print the parsed key = value map
//
// Parameter file (ConfigFileHandler, src/ConfigFileHandler.cpp:32-56): one
// `key = value` or `key = "value"` per line; keys are [_A-Za-z0-9]+; the first
// occurrence of a key wins (std::map::insert); CRLF line ends are accepted.
// Values are read like parseToScalar / parseToVector (include/cvector.cpp:403-424):
// a stream extraction, 0 when it fails; vectors split on white space.  Lines
// are matched one at a time (the reference's regex relies on a per-line `$`).
//
// Run (CSolver::Run_Cuda, src/solver.cpp:380-917, one VOICE chunk):
// BMOHCKernel_Init -> BMOHCNewKernel -> copies -> write failed nodes/blocks,
// convergence data per Convergence_Tests bits, BM velocity (Allowed_Outputs
// bit 0) -> IHCNewKernel + RunIHCKernel -> lambda high/medium/low (bits 1-3)
// -> timers.  Output file (OutputBuffer<float>, include/OutputBuffer.h:47-123):
// the float32 concatenation of every written vector, in that order.  The
// emitted rows of a first chunk are [overlap, totalTimeNodes) (the transient
// overlap is discarded, src/solver.cpp:704-727).
//
// Supported keys: Fs, Sim_Type (0 VOICE, 1 sine, 3/4 JND runs below),
// Input_Signal, Input_File (raw float64, CBin), Sin_Freq and Sin_dB (the tone
// of Sim_Type 1), Duration, Time_Blocks,
// Time_Block_Length, Overlap_Time, OHC_Vector, IHC_Vector, Allowed_Outputs,
// Discard_BM_Velocity_Output, Discard_Lambdas_Output, Output_Result,
// JACOBBY_Loops_Fast, JACOBBY_Loops_Slow, Cuda_Outern_Loops, MAX_M1_SP_ERROR,
// MAX_TOLERANCE, RELATIVE_ERRORS, M1_SP_Fix_Factor, Tolerance_Fix_Factor, SPLRef,
// Decouple_Filter, Convergence_Tests, Lambda_SAT, Disable_Lambda,
// AC_Filter_Vector (Filter_Mode 2 vector: [order, isFIR, (a...), b...]),
// Function_Filter (equiripple, window or Butterworth low-pass designed by the
// library).
// Run_Stage_Calculation (1: from BM velocity, [totalTimeNodes][256] floats;
// 2: from lambda, three such planes) with Run_Stage_File_name (raw float32) or
// Run_Stage_Vector: the stages before it are skipped, as are the convergence
// and failure vectors, which only a run from the input produces.
// JND runs (Sim_Type 3 or 4 with Calculate_JND = 1): testedFrequencies,
// testedPowerLevels, testedNoises, Generate_One_Ref_Per_Noise_Level,
// Complex_Profile_* vectors, JND_Interval_Duration/Head/Tail,
// JND_Delta_Alpha_Time_Factor, Calculate_JND_Types, Calculate_From_Mean_Rate,
// Complex_JND_Calculation_Types, Approximated_JND_EPS_Diff, Max_dB_SPL_JND, W,
// Input_Noise / Input_Noise_File, JND_File_Name, JND_Raw_File_Name.  Noise_Seed
// (no reference counterpart) seeds the white noise when none is given.
// TEST_File_Target with Type_TEST_Output (default MeanRate): one stage written
// as raw float32 -- the IHC stages AC, DC, dS, Shigh, IHC, PRE_IHC, the lambda
// planes JND_Lambda[k], and with Calculate_JND the JND stages MeanRate,
// AvgMeanRate, dMeanRate, dLambda[k], preFisherAI, preFisherAITimeReduced and
// CRLB_RA (src/AudioGramCreator.cpp:884-971).  Where the reference writes a
// stale host vector for AvgMeanRate, preFisherAI and preFisherAITimeReduced
// (no branch of :938-964 copies them), the stage the kernel computed is
// written.
// Keys of this library's solver options (no reference counterpart):
// Jacobi_Tolerance (> 0: Jacobi to that relative residual instead of the fixed
// sweeps), Jacobi_Check_Every (4), Jacobi_Max_Sweeps (1000), Warm_Start (0
// previous pressure, 1 extrapolated, 2 learned), Warm_Start_Model (default
// data/warmstart_mlp.bin beside this executable) and Warm_Start_Every (the
// network on every k-th step, default 1).
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cochlea_amd.h"

namespace {

constexpr int kSections = 256;
constexpr int kLambdas = 3;

// ------------------------------------------------------------ par file ----
class ParFile {
 public:
  explicit ParFile(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open parameter file " + path);
    std::string line;
    while (std::getline(f, line)) parse_line(line);
  }
  bool has(const std::string &k) const { return kv_.count(k) != 0; }
  std::string str(const std::string &k, const std::string &def = "") const {
    auto it = kv_.find(k);
    return it == kv_.end() ? def : it->second;
  }
  template <class T> T scalar(const std::string &k, T def) const {
    if (!has(k)) return def;
    std::stringstream ss(str(k));
    T v;
    return (ss >> v) ? v : T(0);
  }
  template <class T> std::vector<T> vec(const std::string &k) const {
    std::vector<T> out;
    std::stringstream ss(str(k));
    std::string tok;
    while (ss >> tok) {
      std::stringstream ts(tok);
      T v;
      out.push_back((ts >> v) ? v : T(0));
    }
    return out;
  }
  const std::map<std::string, std::string> &map() const { return kv_; }

 private:
  static bool key_char(char c) { return std::isalnum((unsigned char)c) || c == '_'; }
  void parse_line(std::string line) {
    while (!line.empty() && (line.back() == '\r' || line.back() == '\n')) line.pop_back();
    // find "key\s*=\s*value" anywhere on the line (regex search semantics)
    for (size_t eq = line.find('='); eq != std::string::npos; eq = line.find('=', eq + 1)) {
      size_t ke = eq;
      while (ke > 0 && std::isspace((unsigned char)line[ke - 1])) --ke;
      size_t ks = ke;
      while (ks > 0 && key_char(line[ks - 1])) --ks;
      if (ks == ke) continue;
      size_t vs = eq + 1;
      while (vs < line.size() && std::isspace((unsigned char)line[vs])) ++vs;
      if (vs >= line.size()) continue;
      std::string value;
      if (line[vs] == '"') {
        const size_t close = line.rfind('"');
        if (close == vs || close + 1 != line.size()) continue;   // quoted value must end the line
        value = line.substr(vs + 1, close - vs - 1);
      } else {
        value = line.substr(vs);
        if (value.find('"') != std::string::npos) continue;
      }
      kv_.insert({line.substr(ks, ke - ks), value});
      return;
    }
  }
  std::map<std::string, std::string> kv_;
};

// --------------------------------------------------------------- helpers ---
void ok(int rc, const char *what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + cm_last_error());
}

// expandVectorToSize (include/cvector.cpp:609-624): each entry repeated in a block
std::vector<float> expand(const std::vector<double> &v, int size) {
  std::vector<float> out;
  if (v.empty()) return std::vector<float>(size, 0.f);
  const int n = (int)v.size(), times = size / n, rem = size % n;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < times + (i < rem ? 1 : 0); ++j) out.push_back((float)v[i]);
  return out;
}

std::vector<double> read_float64_file(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open input file " + path);
  f.seekg(0, std::ios::end);
  const size_t n = (size_t)f.tellg() / sizeof(double);
  f.seekg(0);
  std::vector<double> v(n);
  f.read(reinterpret_cast<char *>(v.data()), n * sizeof(double));
  return v;
}

// CParams::get_stage_data (src/params.cpp:1710-1745): n floats of the stage
// data from element `offset` on, from Run_Stage_Vector or the raw float32 file
// Run_Stage_File_name (TBin::read_padd, include/TBin.tpp:138-196: what the file
// lacks stays 0).
std::vector<float> stage_data(const ParFile &par, size_t offset, size_t n) {
  std::vector<float> out(n, 0.f);
  if (par.has("Run_Stage_Vector")) {
    const std::vector<float> v = par.vec<float>("Run_Stage_Vector");
    for (size_t i = 0; i < n && offset + i < v.size(); ++i) out[i] = v[offset + i];
    return out;
  }
  const std::string path = par.str("Run_Stage_File_name");
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open stage file " + path);
  f.seekg(0, std::ios::end);
  const size_t len = (size_t)f.tellg() / sizeof(float);
  if (offset < len) {
    f.seekg((std::streamoff)(offset * sizeof(float)));
    f.read(reinterpret_cast<char *>(out.data()), (std::streamsize)(std::min(n, len - offset) * sizeof(float)));
  }
  return out;
}

// the committed default AC filter: the numbers of "filter_mode_2": [ ... ]
std::vector<double> read_filter_fixture(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open AC filter fixture " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  const size_t k = s.find("\"filter_mode_2\"");
  const size_t a = s.find('[', k), b = s.find(']', a);
  if (k == std::string::npos || a == std::string::npos || b == std::string::npos)
    throw std::runtime_error("no filter_mode_2 vector in " + path);
  std::string body = s.substr(a + 1, b - a - 1);
  for (char &c : body)
    if (c == ',') c = ' ';
  std::stringstream bs(body);
  std::vector<double> v;
  double x;
  while (bs >> x) v.push_back(x);
  return v;
}

// The AC filter (SubModel, src/SubModel.cpp:86-131): AC_Filter_Vector
// (Filter_Mode 2: [order, isFIR, (a...), b...]), else a Function_Filter
// equiripple design (cm_design_ac_filter), else the committed equiripple
// design of the default band (data/ac_filter_equiripple_<Fs>.json).
// Returns true for an FIR (a stays empty).
bool ac_filter(const ParFile &par, const std::string &exe_dir, float Fs, std::vector<double> &b,
               std::vector<double> &a) {
  b.clear();
  a.clear();
  if (!par.has("AC_Filter_Vector") && par.has("Function_Filter") && par.scalar<int>("Filter_Mode", 1) == 1) {
    std::vector<double> h(4096), den(4096);
    int n = 0, iir = 0;
    ok(cm_design_ac_filter(Fs, par.str("Function_Filter").c_str(), h.data(), den.data(), (int)h.size(), &n, &iir),
       "cm_design_ac_filter");
    b.assign(h.begin(), h.begin() + n);
    if (iir) a.assign(den.begin(), den.begin() + n);
    return !iir;
  }
  const std::vector<double> fm2 = par.has("AC_Filter_Vector")
                                      ? par.vec<double>("AC_Filter_Vector")
                                      : read_filter_fixture(exe_dir + "/data/ac_filter_equiripple_" +
                                                            std::to_string((int)Fs) + ".json");
  const int order = (int)fm2.at(0);
  const bool fir = (int)fm2.at(1) == 1;
  if (fir) {
    b.assign(fm2.begin() + 2, fm2.begin() + 2 + order);
  } else {
    a.assign(fm2.begin() + 2, fm2.begin() + 2 + order);
    b.assign(fm2.begin() + 2 + order, fm2.begin() + 2 + 2 * order);
  }
  return fir;
}

class BinWriter {
 public:
  explicit BinWriter(const std::string &path) : f_(path, std::ios::binary) {
    if (!f_) throw std::runtime_error("cannot open output file " + path);
  }
  void write(const float *v, size_t n) { f_.write(reinterpret_cast<const char *>(v), n * sizeof(float)); }
  template <class T> void write_cast(const T *v, size_t n) {
    std::vector<float> t(v, v + n);
    write(t.data(), n);
  }

 private:
  std::ofstream f_;
};

std::string dir_of(const std::string &p) {
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? "." : p.substr(0, s);
}

// Stage dumps (TEST_File_Target, Type_TEST_Output; src/params.cpp:109, 165-187,
// 1695-1707): the backup stage number of a Type_TEST_Output name.  1..6 are
// stages of the JND kernel, 9..14 of the IHC chain; 0 names a stage read
// without a backup (JND_Lambda[k], MeanRate).  Unknown names are refused.
struct TestTarget {
  std::string path, type;
  int stage = -1;      // -1: no TEST_File_Target
  int plane = -1;      // dLambdaK / JND_LambdaK: lambda plane k, else all three
};

TestTarget test_target(const ParFile &par) {
  TestTarget t;
  if (!par.has("TEST_File_Target")) return t;
  t.path = par.str("TEST_File_Target");
  t.type = par.str("Type_TEST_Output");
  if (t.type.empty()) t.type = "MeanRate";
  static const std::map<std::string, int> kStages = {
      {"AvgMeanRate", 1}, {"dMeanRate", 2}, {"dLambda", 3}, {"preFisherAI", 4}, {"preFisherAITimeReduced", 5},
      {"CRLB_RA", 6}, {"AC", 9}, {"DC", 10}, {"dS", 11}, {"Shigh", 12}, {"IHC", 13}, {"PRE_IHC", 14},
      {"JND_Lambda", 0}, {"MeanRate", 0}};
  std::string base = t.type;
  // the reference matches the per-plane forms as dLambda<digit> / JND_Lambda<digit>
  // (src/AudioGramCreator.cpp:925, 942); "dLambda(k)" is accepted as well
  for (const char *pre : {"dLambda", "JND_Lambda"}) {
    const size_t n = std::strlen(pre);
    if (t.type.size() > n && t.type.compare(0, n, pre) == 0) {
      std::string rest = t.type.substr(n);
      if (rest.size() == 3 && rest[0] == '(' && rest[2] == ')') rest = rest.substr(1, 1);
      if (rest.size() == 1 && std::isdigit((unsigned char)rest[0])) {
        t.plane = rest[0] - '0';
        if (t.plane >= kLambdas) throw std::runtime_error("index of lambda must be smaller than 3, you chose " + rest);
        base = pre;
      }
    }
  }
  const auto it = kStages.find(base);
  if (it == kStages.end()) throw std::runtime_error("unknown Type_TEST_Output " + t.type);
  t.stage = it->second;
  return t;
}

// saveArrayToDisk without the time column (src/AudioGramCreator.cpp:25-60):
// raw float32
void write_test_target(const TestTarget &t, const float *v, size_t n) {
  BinWriter w(t.path);
  w.write(v, n);
  std::printf("cochlea_run: %s (%zu values) written to %s\n", t.type.c_str(), n, t.path.c_str());
}

// ------------------------------------------------------------------ run ---
int run(const ParFile &par, const std::string &exe_dir) {
  const float Fs = par.scalar<float>("Fs", 44100.f);
  const float block_len = par.scalar<float>("Time_Block_Length", 0.02f);
  const double overlap_time = par.scalar<double>("Overlap_Time", 0.005);
  const int block = (int)std::ceil(block_len * Fs);
  const long overlap_us = (long)std::llround(overlap_time * 1000.0) * 1000;
  const int overlap = (int)(overlap_us * (double)Fs / 1000000.0);

  // Run_Stage_Calculation (include/params.h:249, 390-395): 0 from the input,
  // 1 from BM velocity, 2 from lambda, read from Run_Stage_File_name or
  // Run_Stage_Vector
  const int run_stage = par.scalar<int>("Run_Stage_Calculation", 0);
  if (run_stage < 0 || run_stage > 2) throw std::runtime_error("Run_Stage_Calculation must be 0, 1 or 2");
  const TestTarget tt = test_target(par);
  if ((tt.stage >= 1 && tt.stage <= 6) || tt.type == "MeanRate")
    throw std::runtime_error("Type_TEST_Output " + tt.type + " is a JND stage: it needs Calculate_JND");
  if (run_stage > 0 && !par.has("Run_Stage_Vector") && par.str("Run_Stage_File_name").empty())
    throw std::runtime_error("Chosen Run stage calculation (=" + std::to_string(run_stage) + ") need data or file");

  // input signal (Input_Signal, Input_File, or Sin_Frequency/Sin_dB)
  std::vector<double> signal;
  if (run_stage > 0) {
    if (!par.has("Time_Blocks") && !par.has("Duration"))
      throw std::runtime_error("a run from a stage needs Time_Blocks or Duration");
  } else if (par.has("Input_Signal")) {
    signal = par.vec<double>("Input_Signal");
  } else if (par.has("Input_File") && !par.str("Input_File").empty()) {
    signal = read_float64_file(par.str("Input_File"));
  } else if (par.scalar<int>("Sim_Type", 0) == 1 || (par.has("Sin_Freq") && par.has("Sin_dB"))) {
    // SIM_TYPE_SIN: CState::gen_input_data (src/state.cpp:411-435) from time 0,
    // amplitude CONV_dB_TO_AMP(Sin_dB) in float (include/const.h:53), the
    // sample time accumulated in steps of 1/Fs; generated below for the
    // whole buffer
  } else {
    throw std::runtime_error("no input: set Input_Signal, Input_File, or Sin_Freq/Sin_dB (Sim_Type 1)");
  }
  const bool sine = run_stage == 0 && signal.empty() &&
                    (par.scalar<int>("Sim_Type", 0) == 1 || (par.has("Sin_Freq") && par.has("Sin_dB")));
  int TB = par.scalar<int>("Time_Blocks", 0);
  if (TB <= 0) {
    const double duration = par.scalar<double>("Duration", (double)signal.size() / Fs);
    TB = std::max(1, (int)std::lround(duration / block_len));
  }
  const int bufP1 = overlap + block * (TB + 1);
  const int total = overlap + block * TB;
  const size_t in_nodes = (size_t)block * (TB + 1) + 2 * overlap;
  const size_t res_nodes = in_nodes * kSections;

  unsigned allowed = 0;
  if (!par.scalar<int>("Discard_BM_Velocity_Output", 0)) allowed |= 1;
  if (!par.scalar<int>("Discard_Lambdas_Output", 0)) allowed |= 14;
  allowed = par.scalar<unsigned>("Allowed_Outputs", allowed);
  const int conv_tests = par.scalar<int>("Convergence_Tests", 0);
  const std::string out_path = par.str("Output_Result");
  if (out_path.empty()) throw std::runtime_error("Output_Result is not set");

  // model (CModel, src/model.cpp:18-86; Init_Tri_Matrix, src/solver.cpp:155-176)
  std::vector<float> sec(9 * kSections);
  ok(cm_pack_section_params(sec.data()), "cm_pack_section_params");
  std::vector<double> lo(kSections), mid(kSections), up(kSections);
  ok(cm_tri_matrix(lo.data(), mid.data(), up.data()), "cm_tri_matrix");
  std::vector<float> upf(up.begin(), up.end()), lof(lo.begin(), lo.end());
  std::vector<double> ohc = par.has("OHC_Vector") ? par.vec<double>("OHC_Vector") : std::vector<double>{0.5};
  const std::vector<float> gamma = expand(ohc, kSections);
  const float *S = sec.data();
  ok(BMOHCKernel_Init(gamma.data(), S + 0 * kSections, S + 1 * kSections, upf.data(), lof.data(), S + 2 * kSections,
                      S + 3 * kSections, S + 4 * kSections, S + 5 * kSections, S + 6 * kSections, S + 7 * kSections,
                      1.f, 0.f, in_nodes, res_nodes, res_nodes * 3, 1, 0, 0),
     "BMOHCKernel_Init");

  std::vector<float> input(bufP1, 0.f);
  if (sine) {
    const float freq = par.scalar<float>("Sin_Freq", 1000.f), db = par.scalar<float>("Sin_dB", 60.f);
    const float amp = 10.0f * 20.0e-6f * std::pow(10.0f, db / 20.0f);
    const double Ts = 1.0 / (double)Fs;
    double t = 0.0;
    for (int i = 0; i < bufP1; ++i) {
      input[i] = (float)((double)amp * std::sin(2.0 * M_PI * (double)freq * t));
      t += Ts;
    }
  }
  for (int i = 0; i < bufP1 && i < (int)signal.size(); ++i) input[i] = (float)signal[i];
  cm_bm_scalars sc;
  ok(cm_default_bm_scalars(Fs, &sc), "cm_default_bm_scalars");
  const int j_fast = par.scalar<int>("JACOBBY_Loops_Fast", 2), j_slow = par.scalar<int>("JACOBBY_Loops_Slow", 20);
  const int passes = par.scalar<int>("Cuda_Outern_Loops", 10);
  const int DF = par.scalar<int>("Decouple_Filter", 0);
  const float tau = par.scalar<float>("Jacobi_Tolerance", 0.f);
  const int warm = par.scalar<int>("Warm_Start", 0);
  if (warm == 2) {
    ok(cm_ws_load_file(par.str("Warm_Start_Model", exe_dir + "/data/warmstart_mlp.bin").c_str()), "cm_ws_load_file");
    ok(cm_ws_set_every(par.scalar<int>("Warm_Start_Every", 1)), "cm_ws_set_every");
  }
  ok(cm_set_bm_solver(tau > 0.f ? 1 : 0, tau > 0.f ? tau : 1e-4f, par.scalar<int>("Jacobi_Check_Every", 4),
                      par.scalar<int>("Jacobi_Max_Sweeps", 1000), warm),
     "cm_set_bm_solver");
  std::vector<float> speeds((size_t)bufP1 * kSections, 0.f);
  float bm_ms = 0.f;
  BinWriter out(out_path);
  const size_t rows = (size_t)(total - overlap);
  if (run_stage == 0) {
    ok(BMOHCNewKernel(input.data(), 0, sc.w_ohc, sc.time_step, sc.time_step_out, sc.delta_x, sc.alpha_r, 1, 1, 0,
                      sc.Ts, sc.alpha_l, sc.alpha_s, sc.Gme, sc.a0, sc.a1, sc.a2, sc.sigma_ow, sc.eta1, sc.eta2, TB,
                      bufP1, overlap, std::lround(overlap / (double)Fs * 1e6), 1, sc.dt_max, sc.dt_min, DF,
                      par.scalar<float>("MAX_M1_SP_ERROR", 1e-15f), par.scalar<float>("MAX_TOLERANCE", 1e-10f),
                      par.scalar<int>("RELATIVE_ERRORS", 0), 1, par.scalar<float>("M1_SP_Fix_Factor", 0.f),
                      par.scalar<float>("Tolerance_Fix_Factor", 0.f), par.scalar<float>("SPLRef", 2e-5f), 0, 0, 1,
                      j_fast, j_slow, passes, 17, 2),
       "BMOHCNewKernel");
    bm_ms = cm_last_kernel_ms();
    ok(BMOHCKernel_Copy_Results(speeds.data(), speeds.size(), 0), "BMOHCKernel_Copy_Results");
    auto dev_copy = [&](void *src, size_t n_bytes, void *dst) {
      ok(GeneralKernel_Copy_Results_Bytes(dst, src, n_bytes, 0), "GeneralKernel_Copy_Results_Bytes");
    };
    std::vector<int32_t> failed_nodes(bufP1), failed_blocks(TB);
    dev_copy(getCudaFailedTimeNodes(), failed_nodes.size() * 4, failed_nodes.data());
    dev_copy(getCudaFailedBlocks(), failed_blocks.size() * 4, failed_blocks.data());
    std::vector<float> conv_times(bufP1), conv_times_b(TB), conv_loops(bufP1), conv_loops_b(TB), conv_j(bufP1),
        conv_j_b(TB);
    ok(extractConvergenceTimes(conv_times.data(), conv_times.size()), "extractConvergenceTimes");
    dev_copy(getCudaConvergeSpeedBlocks(), TB * 4, conv_times_b.data());
    dev_copy(getCudaConvergedTimeNodes(), bufP1 * 4, conv_loops.data());
    dev_copy(getCudaConvergedBlocks(), TB * 4, conv_loops_b.data());
    dev_copy(getCudaConvergedJacobyLoopsPerIteration(), bufP1 * 4, conv_j.data());
    dev_copy(getCudaConvergedJacobyLoopsPerIterationBlocks(), TB * 4, conv_j_b.data());
    // convergence and failure data exist only for a run from the input (src/solver.cpp:706-721)
    if (conv_tests & 1) out.write(conv_times.data() + overlap, rows);
    if (conv_tests & 2) out.write(conv_times_b.data(), TB);
    out.write_cast(failed_nodes.data() + overlap, rows);
    out.write_cast(failed_blocks.data(), TB);
    if (conv_tests & 4) {
      out.write(conv_loops.data() + overlap, rows);
      out.write(conv_loops_b.data(), TB);
    }
    if (conv_tests & 8) {
      out.write(conv_j.data() + overlap, rows);
      out.write(conv_j_b.data(), TB);
    }
  } else if (run_stage == 1) {
    // resume from BM velocity: [totalTimeNodes][256] uploaded into the BM
    // buffer (src/solver.cpp:683-692)
    const std::vector<float> v = stage_data(par, 0, (size_t)total * kSections);
    std::copy(v.begin(), v.end(), speeds.begin());
    ok(ReverseKernel_Copy_Results(speeds.data(), (size_t)total * kSections), "ReverseKernel_Copy_Results");
  }
  if (allowed & 1) out.write(speeds.data() + (size_t)overlap * kSections, rows * kSections);

  float ihc_ms = 0.f;
  if (!par.scalar<int>("Disable_Lambda", 0) && (allowed & 14)) {
    // AudioGramCreator::setupIHCOHC (src/AudioGramCreator.cpp:603-823)
    std::vector<double> b, a;
    const bool fir = ac_filter(par, exe_dir, Fs, b, a);
    const int ndc = (int)std::lround(Fs * 2e-3f);
    const double dcv = (double)(1.0f / (Fs * 2e-3f));
    std::vector<double> dc(ndc, dcv);
    std::vector<double> ihcv = par.has("IHC_Vector") ? par.vec<double>("IHC_Vector") : std::vector<double>{8.0};
    const std::vector<float> ihc = expand(ihcv, kSections);
    std::vector<double> dmg(kSections);
    for (int i = 0; i < kSections; ++i) dmg[i] = std::pow(10.0, (double)ihc[i]);
    const float nc[9] = {60.f, 3.f, 0.1f, 0.f, 0.f, 0.f, 0.61f, 0.23f, 0.16f};   // spont, -, W
    const float sat = par.scalar<float>("Lambda_SAT", 500.f);
    ok(IHCNewKernel(dmg.data(), nc, dc.data(), ndc, b.data(), fir ? nullptr : a.data(), fir ? 0 : 1, (int)b.size(),
                    kSections, TB, par.scalar<double>("SPLRef", 2e-5), nullptr, 0, total, total, total, block, 0, 0,
                    sat, 1.f, 100.f, 1, 1, 0, 0, 0, 0, 1.f, 1, 1, 0, DF, 0),
       "IHCNewKernel");
    ok(enableloadAihc(), "enableloadAihc");
    std::vector<float> aihc(kLambdas * kSections);
    const float A3[3] = {70.f, 50.f, 30.f};
    for (int j = 0; j < kLambdas; ++j)
      for (int i = 0; i < kSections; ++i) aihc[j * kSections + i] = A3[j];
    ok(loadAihc(aihc.data()), "loadAihc");
    std::vector<float> backup;
    const auto t0 = std::chrono::steady_clock::now();
    if (run_stage < 2) {
      if (tt.stage >= 9) backup.resize((size_t)total * kSections);
      ok(RunIHCKernel(backup.empty() ? nullptr : backup.data(), 0, 1, tt.stage >= 9 ? tt.stage : 0, 0),
         "RunIHCKernel");
      if (tt.stage >= 9) write_test_target(tt, backup.data(), backup.size());
    } else {
      // resume from lambda (src/AudioGramCreator.cpp:895-902): plane j of the
      // stage data is rows [j*total, (j+1)*total); upload, then the
      // spontaneous-rate fix-up, keeping max(spont, lambda) for the output
      const size_t plane = (size_t)total * kSections;
      const std::vector<float> lam_in = stage_data(par, 0, kLambdas * plane);
      ok(ReverseKernel_Copy_Results_Float(lam_in.data(), static_cast<float *>(getCudaLambda()), 0,
                                          (size_t)kLambdas * total, kSections),
         "ReverseKernel_Copy_Results_Float");
      ok(updateCUDALambdaArray(static_cast<float *>(getCudaLambda()), static_cast<float *>(getCudaBuffer()),
                               (size_t)total, kSections, 0, 0, 1),
         "updateCUDALambdaArray");
    }
    ok(cm_device_synchronize(), "cm_device_synchronize");
    ihc_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<float> lam((size_t)kLambdas * total * kSections);
    ok(BMOHCKernel_Copy_Lambda(lam.data(), lam.size(), 0), "BMOHCKernel_Copy_Lambda");
    for (int j = 0; j < kLambdas; ++j)
      if (allowed & (2u << j))
        out.write(lam.data() + ((size_t)j * total + overlap) * kSections, rows * kSections);
    if (tt.stage == 0) {
      // JND_Lambda[k] (src/AudioGramCreator.cpp:921-934): the lambda planes
      const size_t plane = (size_t)total * kSections;
      std::vector<float> v(tt.plane < 0 ? kLambdas * plane : plane);
      ok(GeneralKernel_Copy_Results_Bytes(v.data(), getCudaLambda(), v.size() * sizeof(float),
                                          (tt.plane < 0 ? 0 : tt.plane) * plane * sizeof(float)),
         "JND_Lambda copy");
      write_test_target(tt, v.data(), v.size());
    }
  }
  // timers_marked: this host's stage times in ms (BM kernel, IHC/lambda)
  const float timers[2] = {bm_ms, ihc_ms};
  out.write(timers, 2);
  std::printf("cochlea_run: %d time blocks, %zu rows written to %s (BM %.2f ms, IHC/lambda %.2f ms)\n", TB, rows,
              out_path.c_str(), bm_ms, ihc_ms);
  ok(BMOHCKernel_Free(), "BMOHCKernel_Free");
  ok(IHCKernel_Free(), "IHCKernel_Free");
  return 0;
}

// --------------------------------------------------------- JND runs ---
// Profile runs with Calculate_JND (Sim_Type 3) and complex-JND runs (Sim_Type
// 4): CParams::updateInputProfile (src/params.cpp:661-782) builds one interval
// per (noise, level, frequency) plus the reference intervals; then input
// synthesis, BM, IHC/lambda and the JND reductions run on the device, and the
// final dB, failed-signal counts and (Sim_Type 4) the complex-JND aggregation
// on the host (src/solver.cpp:735-835, src/AudioGramCreator.cpp:454-571).
struct JndLayout {
  std::vector<cm_jnd_profile> prof;
  std::vector<int> refs, serial, calc, to_ref;
};

JndLayout jnd_layout(const std::vector<float> &freqs, const std::vector<float> &levels,
                     const std::vector<float> &noises, bool one_ref_per_noise, double spl_ref,
                     const std::vector<double> &divisor, const std::vector<double> &fix_factor,
                     const std::vector<double> &fix_add, bool calculate_jnd = true) {
  constexpr float kNoNoise = 1111.f;   // MIN_INF_POWER_LEVEL
  auto spl2pa = [&](double db) { return db == kNoNoise ? 0.0 : 10.0 * std::pow(10.0, db / 20.0) * spl_ref; };
  const int tf = (int)freqs.size(), tp = (int)levels.size(), tn = (int)noises.size();
  if (tf == 0 || tp == 0 || tn == 0) throw std::runtime_error("testedFrequencies/PowerLevels/Noises must not be empty");
  JndLayout L;
  const int add_f = calculate_jnd && !one_ref_per_noise ? 1 : 0;
  const int add_n = calculate_jnd && one_ref_per_noise ? 1 : 0;
  L.prof.assign((size_t)((tf + add_f) * tp + add_n) * tn, cm_jnd_profile{});
  int g = 0, calc = 0;
  auto reference = [&](int nl) {
    cm_jnd_profile &r = L.prof[g];
    r = cm_jnd_profile{};
    r.dBSPLNoise = noises[nl] != kNoNoise ? noises[nl] : 0.f;
    r.Wn = spl2pa(noises[nl]);
    r.isReference = 1;
    r.referenceIndex = g;
    r.calculateIndexPosition = -1;
    L.refs.push_back(g++);
  };
  for (int nl = 0; nl < tn; ++nl) {
    for (int pl = 0; pl < tp; ++pl) {
      for (int fl = 0; fl < tf; ++fl) {
        const size_t cid = (size_t)tf * nl + fl;
        float base = levels[pl];
        if (cid < divisor.size() && divisor[cid] != 0 && noises[nl] != kNoNoise && noises[nl] > 0)
          base = levels[0] + (levels[pl] - levels[0]) * (float)divisor[cid] / noises[nl];
        if (noises[nl] != kNoNoise)
          base += (float)(cid < fix_factor.size() ? fix_factor[cid] : 0.0) * noises[nl] +
                  (float)(cid < fix_add.size() ? fix_add[cid] : 0.0);
        cm_jnd_profile &p = L.prof[g];
        p = cm_jnd_profile{};
        const double pw = spl2pa(base);
        p.dBSPLSignal = base;
        p.dBSPLNoise = noises[nl] != kNoNoise ? noises[nl] : 0.f;
        p.referenceIndex = !calculate_jnd ? -1 : one_ref_per_noise ? tp * tf * (nl + 1) + nl : g + tf - fl;
        p.calculateIndexPosition = calc++;
        p.dA = pw;
        p.recipdA = pw > 0 ? 1.0 / pw : 0.0;
        p.Wn = spl2pa(noises[nl]);
        p.frequency = freqs[fl];
        ++g;
      }
      if (add_f) reference(nl);
    }
    if (add_n) reference(nl);
  }
  // JND_Serial_Intervals_Positions / JND_Interval_To_Reference (src/params.cpp:1596-1616)
  const int n = (int)L.prof.size(), nr = (int)L.refs.size();
  L.serial.assign(n, 0);
  if (nr == 0) {   // no references: every interval is calculated in place
    for (int i = 0; i < n; ++i) {
      L.serial[i] = i;
      L.calc.push_back(i);
    }
    return L;
  }
  for (int i = 0; i < n; ++i) {
    const int rr = (int)((int64_t)i * nr / n);
    if (std::find(L.refs.begin(), L.refs.end(), i) == L.refs.end()) {
      L.serial[i] = (int)L.calc.size();
      L.calc.push_back(i);
      L.to_ref.push_back(L.refs[rr]);
    } else {
      L.serial[i] = rr;
    }
  }
  return L;
}

// transposeVector (include/cvector.cpp:520-530)
std::vector<float> transpose(const std::vector<float> &v, size_t column_size) {
  std::vector<float> r(v.size());
  const size_t row = v.size() / column_size;
  for (size_t i = 0; i < v.size(); ++i) r[(i % row) * column_size + i / row] = v[i];
  return r;
}

// white noise of the profile runs: Input_Noise, Input_Noise_File (raw float64)
// or, in place of the reference's random_device draw, N(0,1) from a seeded
// 64-bit LCG with Box-Muller (Noise_Seed, default 0)
std::vector<float> jnd_noise(const ParFile &par, int n) {
  std::vector<double> src;
  if (par.has("Input_Noise")) src = par.vec<double>("Input_Noise");
  else if (!par.str("Input_Noise_File").empty()) src = read_float64_file(par.str("Input_Noise_File"));
  std::vector<float> wn(n, 0.f);
  if (!src.empty()) {
    for (int i = 0; i < n && i < (int)src.size(); ++i) wn[i] = (float)src[i];
    return wn;
  }
  uint64_t st = 0x9E3779B97F4A7C15ull ^ (uint64_t)par.scalar<long long>("Noise_Seed", 0);
  auto uni = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  for (int i = 0; i < n; i += 2) {
    const double r = std::sqrt(-2.0 * std::log(uni())), t = 2.0 * M_PI * uni();
    wn[i] = (float)(r * std::cos(t));
    if (i + 1 < n) wn[i + 1] = (float)(r * std::sin(t));
  }
  return wn;
}

int run_jnd(const ParFile &par, const std::string &exe_dir) {
  const int sim_type = par.scalar<int>("Sim_Type", 0);
  const float Fs = par.scalar<float>("Fs", 44100.f);
  const double spl_ref = par.scalar<double>("SPLRef", 2e-5);
  const std::vector<float> freqs = par.has("testedFrequencies")
                                       ? par.vec<float>("testedFrequencies")
                                       : std::vector<float>{250.f, 500.f, 1000.f, 2000.f, 3000.f, 4000.f, 6000.f};
  const std::vector<float> levels = par.has("testedPowerLevels") ? par.vec<float>("testedPowerLevels")
                                                                 : std::vector<float>{10.f};
  const std::vector<float> noises = par.has("testedNoises") ? par.vec<float>("testedNoises")
                                                            : std::vector<float>{1111.f};
  const bool one_ref = par.scalar<int>("Generate_One_Ref_Per_Noise_Level", 1) == 1;
  const bool calculate_jnd = par.scalar<int>("Calculate_JND", 0) == 1;
  const JndLayout lay = jnd_layout(freqs, levels, noises, one_ref, spl_ref,
                                   par.vec<double>("Complex_Profile_Power_Level_Divisor"),
                                   par.vec<double>("Complex_Profile_Noise_Level_Fix_Factor"),
                                   par.vec<double>("Complex_Profile_Noise_Level_Fix_Addition"), calculate_jnd);
  const std::string jnd_file = par.str("JND_File_Name");
  if (calculate_jnd && jnd_file.empty()) throw std::runtime_error("JND File is missing, cannot output results...");
  const TestTarget tt = test_target(par);
  if (!calculate_jnd && ((tt.stage >= 1 && tt.stage <= 6) || tt.type == "MeanRate"))
    throw std::runtime_error("Type_TEST_Output " + tt.type + " is a JND stage: it needs Calculate_JND");
  const std::string out_path = par.str("Output_Result");
  unsigned allowed = 0;
  if (!par.scalar<int>("Discard_BM_Velocity_Output", 0)) allowed |= 1;
  if (!par.scalar<int>("Discard_Lambdas_Output", 0)) allowed |= 14;
  allowed = par.scalar<unsigned>("Allowed_Outputs", allowed);
  if (!calculate_jnd && out_path.empty()) throw std::runtime_error("Output_Result is not set");
  const float dur = par.scalar<float>("JND_Interval_Duration", 0.02f);
  const float head_s = par.scalar<float>("JND_Interval_Head", 0.014f);
  const float tail_s = par.scalar<float>("JND_Interval_Tail", 0.002f);
  const double dalpha = par.scalar<double>("JND_Delta_Alpha_Time_Factor", 0.2);
  const int interval = (int)std::lround(Fs * dur);          // JND_Interval_Nodes_Full (roundf)
  const int head = (int)std::lround(Fs * head_s);
  const int length = (int)std::lround(Fs * (dur - head_s - tail_s));
  const float block_len = par.scalar<float>("Time_Block_Length", 0.02f);
  const int block = (int)std::ceil(block_len * Fs);
  const int DF = par.scalar<int>("Decouple_Filter", interval / block);
  if (DF <= 0 || DF * block != interval)
    throw std::runtime_error("a profile run needs JND_Interval_Duration = Decouple_Filter x Time_Block_Length");
  const int n_prof = (int)lay.prof.size(), n_calc = (int)lay.calc.size();
  const int TB = n_prof * DF;
  const long overlap_us = (long)std::llround(par.scalar<double>("Overlap_Time", 0.005) * 1000.0) * 1000;
  const int overlap = (int)(overlap_us * (double)Fs / 1000000.0);
  const int bufP1 = overlap + block * (TB + 1);
  const int rows = n_prof * interval;
  const unsigned types = par.scalar<unsigned>("Calculate_JND_Types", 0xFFFFu);
  const int calc_rms = (types & 1) != 0, calc_ai = (types & 2) != 0;
  std::printf("cochlea_run: Sim_Type %d, %d intervals (%d calculated), %d time blocks\n", sim_type, n_prof, n_calc, TB);

  // model and buffers
  std::vector<float> sec(9 * kSections);
  ok(cm_pack_section_params(sec.data()), "cm_pack_section_params");
  std::vector<double> lo(kSections), mid(kSections), up(kSections);
  ok(cm_tri_matrix(lo.data(), mid.data(), up.data()), "cm_tri_matrix");
  std::vector<float> upf(up.begin(), up.end()), lof(lo.begin(), lo.end());
  std::vector<double> ohc = par.has("OHC_Vector") ? par.vec<double>("OHC_Vector") : std::vector<double>{0.5};
  const std::vector<float> gamma = expand(ohc, kSections);
  const float *S = sec.data();
  const size_t in_nodes = (size_t)bufP1 + 2 * overlap;
  ok(BMOHCKernel_Init(gamma.data(), S + 0 * kSections, S + 1 * kSections, upf.data(), lof.data(), S + 2 * kSections,
                      S + 3 * kSections, S + 4 * kSections, S + 5 * kSections, S + 6 * kSections, S + 7 * kSections,
                      1.f, 0.f, in_nodes, in_nodes * kSections, 0, 1, 0, 0),
     "BMOHCKernel_Init");

  // input synthesis into the BM input buffer (generateInputFromProfile)
  const std::vector<float> wn = jnd_noise(par, interval);
  const int e1 = (int)((dur - tail_s) * Fs);
  ok(InputProfilesArrayInitialization(n_prof, interval, 0, 0, 0), "InputProfilesArrayInitialization");
  const float one = 1.f;
  ok(generateInputFromProfile(lay.prof.data(), wn.data(), interval, nullptr, 0, 0, 0, 0,
                              par.scalar<double>("Normalize_Noise_Energy_To_Given_Interval", 0.16), 0.0, 0, e1, 0, e1,
                              n_prof, 0, 0, n_prof, overlap, interval, head, length, Fs, 0, nullptr, 0, 1,
                              par.scalar<float>("MAX_M1_SP_ERROR", 1e-15f), par.scalar<float>("MAX_TOLERANCE", 1e-10f),
                              0, 0.f, 0.f, DF, 0, &one, 1, &one, 1),
     "generateInputFromProfile");
  cm_bm_scalars sc;
  ok(cm_default_bm_scalars(Fs, &sc), "cm_default_bm_scalars");
  ok(cm_set_bm_solver(0, 1e-4f, 4, 1000, 0), "cm_set_bm_solver");
  ok(BMOHCNewKernel(nullptr, 1, sc.w_ohc, sc.time_step, sc.time_step_out, sc.delta_x, sc.alpha_r, 1, 1, 0, sc.Ts,
                    sc.alpha_l, sc.alpha_s, sc.Gme, sc.a0, sc.a1, sc.a2, sc.sigma_ow, sc.eta1, sc.eta2, TB, bufP1,
                    overlap, overlap_us, 1, sc.dt_max, sc.dt_min, DF, par.scalar<float>("MAX_M1_SP_ERROR", 1e-15f),
                    par.scalar<float>("MAX_TOLERANCE", 1e-10f), 0, 0, 0.f, 0.f, (float)spl_ref, 0, 0, 0,
                    par.scalar<int>("JACOBBY_Loops_Fast", 2), par.scalar<int>("JACOBBY_Loops_Slow", 20),
                    par.scalar<int>("Cuda_Outern_Loops", 10), 17, 2),
     "BMOHCNewKernel");
  const float bm_ms = cm_last_kernel_ms();
  std::vector<int32_t> failed_blocks(TB);
  ok(GeneralKernel_Copy_Results_Bytes(failed_blocks.data(), getCudaFailedBlocks(), TB * 4, 0), "failed blocks");
  // Output_Result of a profile run (src/solver.cpp:704-734): decoupled
  // intervals show the transient and drop the tail (Remove_Transient_Tail), so
  // rows [0, rows) are written
  std::unique_ptr<BinWriter> out;
  const int conv_tests = par.scalar<int>("Convergence_Tests", 0);
  if (!out_path.empty()) {
    out.reset(new BinWriter(out_path));
    auto dev_copy = [&](void *src, size_t n_bytes, void *dst) {
      ok(GeneralKernel_Copy_Results_Bytes(dst, src, n_bytes, 0), "GeneralKernel_Copy_Results_Bytes");
    };
    std::vector<int32_t> failed_nodes(bufP1);
    dev_copy(getCudaFailedTimeNodes(), failed_nodes.size() * 4, failed_nodes.data());
    std::vector<float> conv_times(bufP1), conv_times_b(TB), conv_loops(bufP1), conv_loops_b(TB), conv_j(bufP1),
        conv_j_b(TB);
    ok(extractConvergenceTimes(conv_times.data(), conv_times.size()), "extractConvergenceTimes");
    dev_copy(getCudaConvergeSpeedBlocks(), TB * 4, conv_times_b.data());
    dev_copy(getCudaConvergedTimeNodes(), bufP1 * 4, conv_loops.data());
    dev_copy(getCudaConvergedBlocks(), TB * 4, conv_loops_b.data());
    dev_copy(getCudaConvergedJacobyLoopsPerIteration(), bufP1 * 4, conv_j.data());
    dev_copy(getCudaConvergedJacobyLoopsPerIterationBlocks(), TB * 4, conv_j_b.data());
    if (conv_tests & 1) out->write(conv_times.data(), rows);
    if (conv_tests & 2) out->write(conv_times_b.data(), TB);
    out->write_cast(failed_nodes.data(), rows);
    out->write_cast(failed_blocks.data(), TB);
    if (conv_tests & 4) {
      out->write(conv_loops.data(), rows);
      out->write(conv_loops_b.data(), TB);
    }
    if (conv_tests & 8) {
      out->write(conv_j.data(), rows);
      out->write(conv_j_b.data(), TB);
    }
    if (allowed & 1) {
      std::vector<float> speeds((size_t)rows * kSections);
      ok(BMOHCKernel_Copy_Results(speeds.data(), speeds.size(), 0), "BMOHCKernel_Copy_Results");
      out->write(speeds.data(), speeds.size());
    }
  }
  const bool need_lambda = calculate_jnd || (out && (allowed & 14));

  if (!need_lambda) {
    const float timers[2] = {bm_ms, 0.f};
    if (out) out->write(timers, 2);
    std::printf("cochlea_run: profile run, %d intervals written to %s (BM %.2f ms)\n", n_prof, out_path.c_str(),
                bm_ms);
    ok(BMOHCKernel_Free(), "BMOHCKernel_Free");
    return 0;
  }
  // IHC / lambda over the decoupled intervals
  std::vector<double> b, a;
  const bool fir = ac_filter(par, exe_dir, Fs, b, a);
  const int ndc = (int)std::lround(Fs * 2e-3f);
  std::vector<double> dc(ndc, (double)(1.0f / (Fs * 2e-3f)));
  std::vector<double> ihcv = par.has("IHC_Vector") ? par.vec<double>("IHC_Vector") : std::vector<double>{8.0};
  const std::vector<float> ihc = expand(ihcv, kSections);
  std::vector<double> dmg(kSections);
  for (int i = 0; i < kSections; ++i) dmg[i] = std::pow(10.0, (double)ihc[i]);
  std::vector<float> W = par.has("W") ? par.vec<float>("W") : std::vector<float>{0.61f, 0.23f, 0.16f};
  W.resize(3, 0.f);
  const float nc[9] = {60.f, 3.f, 0.1f, 0.f, 0.f, 0.f, W[0], W[1], W[2]};
  ok(IHCNewKernel(dmg.data(), nc, dc.data(), ndc, b.data(), fir ? nullptr : a.data(), fir ? 0 : 1, (int)b.size(),
                  kSections, TB, spl_ref, nullptr, 0, rows, rows, rows, block, 0, 0,
                  par.scalar<float>("Lambda_SAT", 500.f), 1.f, 100.f, 1, 1, 0, 0, 0, 0, 1.f, 1, 1, 0, DF, 0),
     "IHCNewKernel");
  ok(enableloadAihc(), "enableloadAihc");
  std::vector<float> aihc(kLambdas * kSections);
  const float A3[3] = {70.f, 50.f, 30.f};
  for (int j = 0; j < kLambdas; ++j)
    for (int i = 0; i < kSections; ++i) aihc[j * kSections + i] = A3[j];
  ok(loadAihc(aihc.data()), "loadAihc");
  std::vector<float> backup;
  if (tt.stage >= 9) backup.resize((size_t)rows * kSections);
  const auto t_ihc = std::chrono::steady_clock::now();
  ok(RunIHCKernel(backup.empty() ? nullptr : backup.data(), 0, 0, tt.stage >= 9 ? tt.stage : 0, 0), "RunIHCKernel");
  ok(cm_device_synchronize(), "cm_device_synchronize");
  const float ihc_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_ihc).count();
  if (out) {
    if (allowed & 14) {
      std::vector<float> lam((size_t)kLambdas * rows * kSections);
      ok(BMOHCKernel_Copy_Lambda(lam.data(), lam.size(), 0), "BMOHCKernel_Copy_Lambda");
      for (int j = 0; j < kLambdas; ++j)
        if (allowed & (2u << j)) out->write(lam.data() + (size_t)j * rows * kSections, (size_t)rows * kSections);
    }
    const float timers[2] = {bm_ms, ihc_ms};
    out->write(timers, 2);
  }
  auto dump_stage = [&]() {
    if (tt.stage < 0) return;
    // the stage dumps of src/AudioGramCreator.cpp:913-966: mean-rate layout
    // [3][intervals][256] (getMeanNodes), lambda layout [3][rows][256]
    // (matrixSize per plane); the IHC stages as RunIHCKernel backed them up
    const size_t plane = (size_t)rows * kSections, mean_nodes = (size_t)kLambdas * n_prof * kSections;
    std::vector<float> v;
    auto copy = [&](void *src, size_t n, size_t off) {
      v.resize(n);
      ok(GeneralKernel_Copy_Results_Bytes(v.data(), src, n * sizeof(float), off * sizeof(float)), "stage copy");
    };
    if (tt.stage >= 9) v = backup;
    else if (tt.type == "MeanRate") copy(getCudaMeanRate(), mean_nodes, 0);
    else if (tt.stage == 0) copy(getCudaLambda(), tt.plane < 0 ? kLambdas * plane : plane, tt.plane < 0 ? 0 : tt.plane * plane);
    else if (tt.stage == 1) copy(getCudaBuffer(), (size_t)kLambdas * n_prof, 0);
    else if (tt.stage == 3 || tt.stage == 4)
      copy(getCudaBuffer(), tt.plane < 0 ? kLambdas * plane : plane, tt.plane < 0 ? 0 : tt.plane * plane);
    else copy(getCudaBuffer(), mean_nodes, 0);
    write_test_target(tt, v.data(), v.size());
  };
  if (!calculate_jnd) {
    dump_stage();
    std::printf("cochlea_run: profile run, %d intervals written to %s (BM %.2f ms)\n", n_prof, out_path.c_str(),
                bm_ms);
    ok(BMOHCKernel_Free(), "BMOHCKernel_Free");
    ok(IHCKernel_Free(), "IHCKernel_Free");
    return 0;
  }

  // JND reductions, final dB, failed signals
  std::vector<double> nIHC(kSections, 30000.0 / kSections);
  std::vector<float> F_RA(3 * (size_t)n_prof, 0.f), FAI(3 * (size_t)n_prof, 0.f);
  ok(CudaCalculateJND(calc_ai, calc_rms, 0, 3 * n_prof, spl_ref, Fs, 1.0, nIHC.data(), lay.calc.data(), n_calc,
                      lay.refs.data(), (int)lay.refs.size(), 0, n_prof, n_prof, head, overlap, interval, length,
                      lay.serial.data(), lay.to_ref.data(), F_RA.data(), FAI.data(), rows * kSections,
                      rows * kSections, dalpha == 0 ? (double)length : dalpha * Fs, 1, 0, 0, 0,
                      tt.stage >= 1 && tt.stage <= 6 ? tt.stage : 0),
     "CudaCalculateJND");
  dump_stage();
  std::vector<double> ai(n_prof, 0.0), rate(n_prof, 0.0);
  ok(cm_jnd_final(F_RA.data(), FAI.data(), n_prof, n_prof, 0, lay.refs.data(), (int)lay.refs.size(),
                  lay.serial.data(), spl_ref, par.scalar<double>("Max_dB_SPL_JND", 150.0), calc_ai, calc_rms,
                  ai.data(), rate.data(), nullptr, nullptr),
     "cm_jnd_final");
  std::vector<int32_t> failed_signals(n_prof, 0);
  ok(cm_jnd_failed_signals(failed_blocks.data(), TB, n_prof, 0, lay.refs.data(), (int)lay.refs.size(),
                           lay.serial.data(), DF, failed_signals.data()),
     "cm_jnd_failed_signals");

  // raw JND output (src/solver.cpp:756-806)
  const std::string raw_name = par.str("JND_Raw_File_Name");
  const bool write_raw = sim_type != 4 || !raw_name.empty();
  if (write_raw) {
    BinWriter raw(sim_type == 4 ? raw_name : jnd_file);
    const size_t mcols = (size_t)(n_calc / (levels.size() * freqs.size())) * levels.size();
    raw.write_cast(failed_signals.data(), n_calc);
    for (int pass = 0; pass < 2; ++pass) {
      if ((pass == 0 && !calc_ai) || (pass == 1 && !calc_rms)) continue;
      const std::vector<double> &src = pass == 0 ? ai : rate;
      std::vector<float> v(src.begin(), src.begin() + n_calc);
      raw.write(v.data(), v.size());                 // jnd_ai_power / jnd_rms_power
      const std::vector<float> t = transpose(v, mcols);
      raw.write(t.data(), t.size());                 // jnd_ai / jnd_rms
    }
  }
  // complex JND (Sim_Type 4, src/solver.cpp:809-835)
  if (sim_type == 4) {
    const unsigned methods = par.scalar<unsigned>("Complex_JND_Calculation_Types", 1u) & 3u;
    const int n_methods = (methods & 1) + ((methods >> 1) & 1);
    const int n_cplx = (int)(freqs.size() * noises.size());
    std::vector<int32_t> iv((size_t)n_cplx * levels.size());
    ok(cm_complex_jnd_intervals((int)freqs.size(), (int)levels.size(), (int)noises.size(), iv.data()),
       "cm_complex_jnd_intervals");
    const std::vector<double> &src = par.scalar<int>("Calculate_From_Mean_Rate", 1) == 1 ? rate : ai;
    std::vector<float> fin((size_t)n_methods * n_cplx);
    std::vector<int32_t> warn(fin.size()), conv(fin.size());
    ok(cm_complex_jnd(src.data(), failed_signals.data(), n_calc, iv.data(), n_cplx, (int)levels.size(), (int)methods,
                      par.scalar<float>("Approximated_JND_EPS_Diff", 1.f), fin.data(), warn.data(), conv.data()),
       "cm_complex_jnd");
    BinWriter fo(jnd_file);
    fo.write(fin.data(), fin.size());                // jnd_final
    fo.write_cast(warn.data(), warn.size());         // jnd_final_warnings
    fo.write_cast(conv.data(), conv.size());         // jnd_final_signal_convergence_warning
  }
  std::printf("cochlea_run: JND written to %s (BM %.2f ms)\n", jnd_file.c_str(), bm_ms);
  ok(BMOHCKernel_Free(), "BMOHCKernel_Free");
  ok(IHCKernel_Free(), "IHCKernel_Free");
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    if (argc == 3 && std::string(argv[1]) == "--parse") {
      ParFile par(argv[2]);
      for (const auto &kv : par.map()) std::printf("%s = %s\n", kv.first.c_str(), kv.second.c_str());
      return 0;
    }
    if (argc == 3 && std::string(argv[1]) == "--profiles") {
      // the JND interval layout of a profile run, one line per interval
      ParFile par(argv[2]);
      const std::vector<float> freqs = par.vec<float>("testedFrequencies"), levels = par.vec<float>("testedPowerLevels");
      const std::vector<float> noises = par.has("testedNoises") ? par.vec<float>("testedNoises")
                                                                : std::vector<float>{1111.f};
      const JndLayout lay = jnd_layout(freqs, levels, noises, par.scalar<int>("Generate_One_Ref_Per_Noise_Level", 1) == 1,
                                       par.scalar<double>("SPLRef", 2e-5), {}, {}, {},
                                       par.scalar<int>("Calculate_JND", 1) == 1);
      for (size_t g = 0; g < lay.prof.size(); ++g) {
        const cm_jnd_profile &q = lay.prof[g];
        std::printf("%zu %.17g %.17g %.17g %d %d %d %d\n", g, q.dA, q.Wn, q.frequency, (int)q.isReference,
                    q.calculateIndexPosition, q.referenceIndex, lay.serial[g]);
      }
      return 0;
    }
    if (argc != 2) {
      std::fprintf(stderr, "usage: %s <file.par> | --parse <file.par> | --profiles <file.par>\n", argv[0]);
      return 2;
    }
    ParFile par(argv[1]);
    const int sim_type = par.scalar<int>("Sim_Type", 0);
    if (sim_type == 3 || (sim_type == 4 && par.scalar<int>("Calculate_JND", 0) == 1))
      return run_jnd(par, dir_of(argv[0]));
    return run(par, dir_of(argv[0]));
  } catch (const std::exception &e) {
    std::fprintf(stderr, "cochlea_run: %s\n", e.what());
    return 1;
  }
}
