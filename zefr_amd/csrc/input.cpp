// input.cpp -- input deck reader and similarity / non-dimensional parameters.
//
// Deck format as in the reference: one "key value" pair per line, first token is the
// key, anything after the value is ignored, unknown keys are ignored
// (include/input.hpp:254-312 read_param).  Defaults and the post-processing of the
// freestream state follow src/input.cpp:31-539.  Only the residual-path subset of the
// reference's options is accepted; options that select code outside that path
// (implicit schemes, adaptive dt, filters, motion, overset, AdvDiff) are rejected with
// an error instead of being silently ignored.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"

namespace zfr {

namespace {

struct Deck {
  std::vector<std::pair<std::string, std::vector<std::string>>> lines;

  explicit Deck(const std::string &file) {
    std::ifstream f(file);
    if (!f.is_open()) raise("cannot open input file " + file);
    std::string s;
    while (std::getline(f, s)) {
      std::stringstream ss(s);
      std::string key, tok;
      if (!(ss >> key)) continue;
      std::vector<std::string> vals;
      while (ss >> tok) vals.push_back(tok);
      lines.emplace_back(key, vals);
    }
  }
  const std::vector<std::string> *find(const std::string &key) const {
    for (auto &l : lines)
      if (l.first == key) return &l.second;
    return nullptr;
  }
  bool has(const std::string &key) const { return find(key) != nullptr; }
  std::string str(const std::string &key, const char *def = nullptr) const {
    auto *v = find(key);
    if (!v || v->empty()) {
      if (!def) raise("Input parameter " + key + " not found!");
      return def;
    }
    return (*v)[0];
  }
  double num(const std::string &key, const double *def = nullptr) const {
    auto *v = find(key);
    if (!v || v->empty()) {
      if (!def) raise("Input parameter " + key + " not found!");
      return *def;
    }
    try {
      return std::stod((*v)[0]);
    } catch (...) {
      raise("Input parameter " + key + " is not a number");
    }
  }
  double num(const std::string &key, double def) const { return num(key, &def); }
  unsigned uint(const std::string &key) const { return (unsigned)num(key); }
  unsigned uint(const std::string &key, unsigned def) const { return (unsigned)num(key, (double)def); }
};

void copy_str(char *dst, size_t n, const std::string &s) {
  if (s.size() >= n) raise("input string too long: " + s);
  std::memset(dst, 0, n);
  std::memcpy(dst, s.data(), s.size());
}

std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), ::tolower);
  return s;
}

}  // namespace

// src/input.cpp:466-539
void apply_nondim(zefr_input_t &in) {
  if (in.disable_nondim) {
    in.T_fs = in.P_fs / (in.rho_fs * in.R);
    in.mu = in.rho_fs * in.v_mag_fs * in.L_fs / in.Re_fs;
    in.T_tot_fs = in.T_fs * (1.0 + 0.5 * (in.gamma - 1.0) * in.mach_fs * in.mach_fs);
    in.P_tot_fs = in.P_fs * std::pow(1.0 + 0.5 * (in.gamma - 1.0) * in.mach_fs * in.mach_fs,
                                     in.gamma / (in.gamma - 1.0));
    for (unsigned d = 0; d < in.nDims; d++) in.V_fs[d] = in.v_mag_fs * in.norm_fs[d];
    in.fix_vis = 1;
    return;
  }
  const double V_fs_mag = in.mach_fs * std::sqrt(in.gamma * in.R * in.T_fs);
  for (unsigned d = 0; d < in.nDims; d++) in.V_fs[d] = V_fs_mag * in.norm_fs[d];
  in.mu = (in.rho_fs * V_fs_mag * in.L_fs) / in.Re_fs;
  if (!in.fix_vis)
    in.mu = in.mu * std::pow(in.T_fs / in.T_gas, 1.5) * ((in.T_gas + in.S) / (in.T_fs + in.S));
  in.rho_fs = in.mu * in.Re_fs / (V_fs_mag * in.L_fs);
  in.P_fs = in.rho_fs * in.R * in.T_fs;
  in.rt = in.T_gas * in.R / (V_fs_mag * V_fs_mag);
  in.c_sth = in.S / in.T_gas;
  const double T_ref = in.T_fs, rho_ref = in.rho_fs, P_ref = in.rho_fs * V_fs_mag * V_fs_mag;
  const double mu_ref = in.rho_fs * V_fs_mag, R_ref = in.R * in.T_fs / (V_fs_mag * V_fs_mag);
  in.mu = in.mu / mu_ref;
  in.rho_fs = in.rho_fs / rho_ref;
  for (unsigned d = 0; d < in.nDims; d++) in.V_fs[d] = in.V_fs[d] / V_fs_mag;
  in.P_fs = in.P_fs / P_ref;
  in.T_tot_fs = (in.T_fs / T_ref) * (1.0 + 0.5 * (in.gamma - 1.0) * in.mach_fs * in.mach_fs);
  in.P_tot_fs = in.P_fs * std::pow(1.0 + 0.5 * (in.gamma - 1.0) * in.mach_fs * in.mach_fs,
                                   in.gamma / (in.gamma - 1.0));
  in.v_mag_fs = 1.0;
  in.T_fs = 1.0;
  in.R = R_ref;
}

// src/input.cpp:431-460 (EulerNS branch)
void finish_freestream(zefr_input_t &in) {
  in.v_mag_fs = in.mach_fs * std::sqrt(in.gamma * in.P_fs / in.rho_fs);
  double magn = 0.;
  for (int i = 0; i < 3; i++) magn += in.norm_fs[i] * in.norm_fs[i];
  magn = std::sqrt(magn);
  for (int i = 0; i < 3; i++) in.norm_fs[i] /= magn;
  if (in.viscous) apply_nondim(in);
  else
    for (int i = 0; i < 3; i++) in.V_fs[i] = in.norm_fs[i] * in.v_mag_fs;
}

void set_stages(zefr_input_t &in) {
  const std::string s = in.dt_scheme;
  if (s == "Euler") in.nStages = 1;
  else if (s == "RK44") in.nStages = 4;
  else raise("dt_scheme '" + s + "' is not supported (Euler, RK44)");
}

void read_input_file(const std::string &file, zefr_input_t &in) {
  std::memset(&in, 0, sizeof(in));
  Deck f(file);
  in.nDims = f.uint("nDims");
  copy_str(in.meshfile, sizeof(in.meshfile), f.str("meshfile"));
  for (auto &l : f.lines) {
    if (l.first != "mesh_bound") continue;
    if (l.second.size() < 2) raise("mesh_bound needs a mesh name and a boundary condition");
    if (in.nMeshBounds >= ZEFR_MAX_BOUNDS) raise("too many mesh_bound entries");
    copy_str(in.meshBoundName[in.nMeshBounds], ZEFR_NAME_LEN, lower(l.second[0]));
    copy_str(in.meshBoundBC[in.nMeshBounds], ZEFR_NAME_LEN, lower(l.second[1]));
    in.nMeshBounds++;
  }
  in.order = f.uint("order");
  const std::string eq = f.str("equation", "EulerNS");
  if (eq == "EulerNS") in.equation = 1;
  else if (eq == "AdvDiff") raise("equation AdvDiff is outside the supported residual path (EulerNS)");
  else raise("Equation not recognized!");
  in.viscous = f.uint("viscous") != 0;
  in.disable_nondim = f.uint("disable_nondim", 0) != 0;
  if (f.uint("source", 0)) raise("source terms are not supported");
  if (f.uint("squeeze", 0)) raise("squeeze stabilisation is not supported");
  if (f.uint("motion", 0)) raise("moving grids are not supported");
  if (f.uint("filt_on", 0)) raise("filtering is not supported");
  if (f.uint("adapt_dt", 0)) raise("adaptive time stepping is not supported");
  // solution files (src/input.cpp:130-147)
  in.restart = f.uint("restart", 0) != 0;
  copy_str(in.restart_case, sizeof(in.restart_case), f.str("restart_case", ""));
  in.restart_type = f.uint("restart_type", 0);
  in.restart_iter = f.uint("restart_iter", 0);
  in.write_pyfr = f.uint("write_pyfr", 0) != 0;
  in.write_paraview = f.uint("write_paraview", 1) != 0;
  // p-multigrid (src/input.cpp:137-142); the level checks of PMGrid::setup
  // (src/multigrid.cpp:24-29) are made here, where the lists are read
  in.p_multi = f.uint("p_multi", 0) != 0;
  copy_str(in.mg_cycle, sizeof(in.mg_cycle), f.str("mg_cycle", "V"));
  in.FMG_vcycles = f.uint("FMG_vcycles", 1);
  in.rel_fac = f.num("rel_fac", 1.0);
  {
    const auto *lv = f.find("mg_levels");
    const auto *st = f.find("mg_steps");
    const size_t nl = lv ? lv->size() : 0, ns = st ? st->size() : 0;
    if (in.p_multi) {
      if (nl == 0 || ns == 0) raise("Need to provide mg_levels and/or mg_steps to run multigrid!");
      if (nl != ns) raise("Inconsistent number of mg_levels and mg_steps provided!");
      if (nl > ZEFR_MAX_LEVELS) raise("at most 8 multigrid levels are supported");
    }
    in.nLevels = (unsigned)std::min<size_t>(std::min(nl, ns), ZEFR_MAX_LEVELS);
    for (unsigned i = 0; i < in.nLevels; i++) {
      in.mg_levels[i] = (unsigned)std::stoul((*lv)[i]);
      in.mg_steps[i] = (unsigned)std::stoul((*st)[i]);
    }
  }
  in.n_steps = f.uint("n_steps");
  copy_str(in.dt_scheme, sizeof(in.dt_scheme), f.str("dt_scheme"));
  set_stages(in);
  in.dt = f.num("dt");
  in.dt_type = f.uint("dt_type", 0);
  if (in.dt_type != 0) raise("only dt_type 0 (supplied dt) is supported");
  copy_str(in.output_prefix, sizeof(in.output_prefix), f.str("output_prefix", "zefr"));
  in.write_freq = f.uint("write_freq", 0);
  in.report_freq = f.uint("report_freq", 0);
  in.res_type = f.uint("res_type", 2);
  in.res_field = f.uint("res_field", 0);
  in.test_case = f.uint("test_case", 0);
  const std::string fc = f.str("fconv_type", "Rusanov");
  // the reference knows Rusanov only (src/input.cpp:177-182); Roe is this build's addition
  if (fc == "Rusanov") in.fconv_type = 0;
  else if (fc == "Roe") in.fconv_type = 2;
  else raise("Convective interface flux not recognized!");
  const std::string fv = f.str("fvisc_type", "LDG");
  if (fv != "LDG") raise("Viscous interface flux not recognized!");
  in.fvisc_type = 1;
  in.rus_k = f.num("rus_k", 0.0);
  in.ldg_b = f.num("ldg_b", 0.5);
  in.ldg_tau = f.num("ldg_tau", 1.0);
  copy_str(in.spt_type, sizeof(in.spt_type), f.str("spt_type", "Legendre"));
  if (std::string(in.spt_type) != "Legendre") raise("spt_type must be Legendre");
  in.ic_type = f.uint("ic_type", 0);
  in.T_gas = f.num("T_gas", 291.15);
  in.gamma = f.num("gamma", 1.4);
  in.R = f.num("R", 286.9);
  in.prandtl = f.num("prandtl", 0.72);
  in.S = f.num("S", 120.0);
  in.rho_fs = f.num("rho_fs", 1.4);
  in.P_fs = f.num("P_fs", 1.0);
  in.fix_vis = f.uint("fix_vis", 0) != 0;
  in.mach_fs = f.num("mach_fs", 0.2);
  in.Re_fs = f.num("Re_fs", 200.0);
  in.L_fs = f.num("L_fs", 1.0);
  in.T_fs = f.num("T_fs", 300.0);
  in.norm_fs[0] = f.num("nx_fs", 1.0);
  in.norm_fs[1] = f.num("ny_fs", 0.0);
  in.norm_fs[2] = f.num("nz_fs", 0.0);
  finish_freestream(in);
}

// src/input.cpp:542-562
void initialize_inputs(zefr_input_t &in) {
  switch (in.nStages) {
    case 1: copy_str(in.dt_scheme, sizeof(in.dt_scheme), "Euler"); break;
    case 4: copy_str(in.dt_scheme, sizeof(in.dt_scheme), "RK44"); break;
    default: raise("nStages must be 1 or 4");
  }
  in.dt_type = 0;
  in.initIter = in.iter;
  if (in.viscous && in.equation == 1) apply_nondim(in);
}

}  // namespace zfr
