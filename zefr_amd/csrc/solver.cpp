// solver.cpp -- Elements / Faces / FRSolver host classes.
//
// Setup follows the reference's sequence (src/solver.cpp:55-180): read the mesh, build
// connectivity, compute geometry and metrics, set the initial state, move everything to
// the device.  Geometry is computed element by element straight into the element-major
// device layout (the reference keeps point-major [pt][dim][ele] arrays of several GB at
// the benchmark size).  The residual and the RK stages are kernel launches on the
// solver's own stream; nothing on the hot path touches the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"

namespace zfr {

// ---------------------------------------------------------------- DevBuf
DevBuf::~DevBuf() { release(); }

void DevBuf::alloc(size_t nbytes) {
  release();
  if (nbytes == 0) return;
  hip_or_throw(hipMalloc(&p, nbytes), "hipMalloc");
  bytes = nbytes;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

template <typename T>
static void upload(DevBuf &b, const std::vector<T> &v) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) hip_or_throw(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
}

// ---------------------------------------------------------------- Elements
void Elements::setup_geometry(const Mesh &m) {
  const int D = nDims, nN = nNodes, E = nEles, nS = nSpts, nM = D * D + 1;
  // shape function tables at spts and fpts (src/elements.cpp:214-323)
  auto tables = [&](const std::vector<double> &locs, int npts, std::vector<double> &shp, std::vector<double> &dshp) {
    shp.assign((size_t)npts * nN, 0.0);
    dshp.assign((size_t)npts * D * nN, 0.0);
    std::vector<double> val(nN), dval((size_t)nN * D);
    for (int p = 0; p < npts; p++) {
      re.shape(&locs[(size_t)p * D], val.data(), dval.data());
      for (int n = 0; n < nN; n++) {
        shp[(size_t)p * nN + n] = val[n];
        for (int r = 0; r < D; r++) dshp[((size_t)p * D + r) * nN + n] = dval[(size_t)n * D + r];
      }
    }
  };
  tables(re.loc_spts, nSpts, shp_s, dshp_s);
  tables(re.loc_fpts, nFpts, shp_f, dshp_f);
  X.assign((size_t)E * nN * D, 0.0);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; e++)
    for (int n = 0; n < nN; n++)
      for (int d = 0; d < D; d++)
        X[((size_t)e * nN + n) * D + d] = m.coord[(size_t)m.ele2nodes[(size_t)e * nN + n] * D + d];
  // metric terms at spts: adjugate of the mapping Jacobian and its determinant
  // (src/elements.cpp:845-875, 964-996)
  metric.assign((size_t)E * nM * nS, 0.0);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int e = 0; e < E; e++)
    for (int s = 0; s < nS; s++) {
      double J[3][3], A[3][3];
      jacobian(e, false, s, J);
      const double det = adjugate(D, J, A);
      if (!(det > 0)) bad++;
      for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) metric[((size_t)e * nM + a * D + b) * nS + s] = A[a][b];
      metric[((size_t)e * nM + D * D) * nS + s] = det;
    }
  if (bad) raise("Negative Jacobian at solution points.");
}

// x = sum_n shape_n(xi) x_n  (src/elements.cpp:325-388)
void Elements::point_coord(int e, bool fpt, int p, double *x) const {
  const int D = nDims, nN = nNodes;
  const double *shp = fpt ? &shp_f[(size_t)p * nN] : &shp_s[(size_t)p * nN];
  const double *Xe = &X[(size_t)e * nN * D];
  for (int d = 0; d < D; d++) {
    double s = 0.0;
    for (int n = 0; n < nN; n++) s += shp[n] * Xe[(size_t)n * D + d];
    x[d] = s;
  }
}

// J[r][d] = dx_d / dxi_r = sum_n dshape_n/dxi_r x_n,d  (src/elements.cpp:845-868)
void Elements::jacobian(int e, bool fpt, int p, double J[3][3]) const {
  const int D = nDims, nN = nNodes;
  const double *dshp = fpt ? &dshp_f[(size_t)p * D * nN] : &dshp_s[(size_t)p * D * nN];
  const double *Xe = &X[(size_t)e * nN * D];
  for (int r = 0; r < D; r++)
    for (int d = 0; d < D; d++) {
      double s = 0.0;
      for (int n = 0; n < nN; n++) s += dshp[(size_t)r * nN + n] * Xe[(size_t)n * D + d];
      J[r][d] = s;
    }
}

// Adjugate (det * J^-1, stored as the reference's inv_jaco(a, pt, b)) and determinant
// (src/elements.cpp:964-996).
double Elements::adjugate(int D, const double J[3][3], double A[3][3]) {
  if (D == 2) {
    A[0][0] = J[1][1];  A[0][1] = -J[1][0];
    A[1][0] = -J[0][1]; A[1][1] = J[0][0];
    return J[0][0] * J[1][1] - J[0][1] * J[1][0];
  }
  const double xr = J[0][0], xs = J[1][0], xt = J[2][0];
  const double yr = J[0][1], ys = J[1][1], yt = J[2][1];
  const double zr = J[0][2], zs = J[1][2], zt = J[2][2];
  A[0][0] = ys * zt - yt * zs; A[0][1] = xt * zs - xs * zt; A[0][2] = xs * yt - xt * ys;
  A[1][0] = yt * zr - yr * zt; A[1][1] = xr * zt - xt * zr; A[1][2] = xt * yr - xr * yt;
  A[2][0] = yr * zs - ys * zr; A[2][1] = xs * zr - xr * zs; A[2][2] = xr * ys - xs * yr;
  return xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
}

// Reference-layout geometry for inspection (small meshes): coord_spts / coord_fpts
// [n][D][E], jaco_det_spts [nS][E], inv_jaco_spts / inv_jaco_fpts [D][n][D][E].
std::vector<double> Elements::export_array(const std::string &name) const {
  const int D = nDims, E = nEles;
  std::vector<double> out;
  if (name == "coord_spts" || name == "coord_fpts") {
    const bool f = name == "coord_fpts";
    const int np = f ? nFpts : nSpts;
    out.assign((size_t)np * D * E, 0.0);
    for (int e = 0; e < E; e++)
      for (int p = 0; p < np; p++) {
        double x[3];
        point_coord(e, f, p, x);
        for (int d = 0; d < D; d++) out[((size_t)p * D + d) * E + e] = x[d];
      }
  } else if (name == "jaco_det_spts") {
    out.assign((size_t)nSpts * E, 0.0);
    for (int e = 0; e < E; e++)
      for (int s = 0; s < nSpts; s++) out[(size_t)s * E + e] = metric[((size_t)e * (D * D + 1) + D * D) * nSpts + s];
  } else if (name == "inv_jaco_spts" || name == "inv_jaco_fpts") {
    const bool f = name == "inv_jaco_fpts";
    const int np = f ? nFpts : nSpts;
    out.assign((size_t)D * np * D * E, 0.0);
    for (int e = 0; e < E; e++)
      for (int p = 0; p < np; p++) {
        double J[3][3], A[3][3];
        jacobian(e, f, p, J);
        adjugate(D, J, A);
        for (int a = 0; a < D; a++)
          for (int b = 0; b < D; b++) out[(((size_t)a * np + p) * D + b) * E + e] = A[a][b];
      }
  } else {
    raise("unknown element array " + name);
  }
  return out;
}

// Initial solution (src/elements.cpp:998-1074, src/funcs.cpp:31-155), element-major.
std::vector<double> Elements::initial_state(const zefr_input_t &in) const {
  const int D = nDims, V = nVars, E = nEles;
  const bool tc_ok = in.ic_type == 0 || (in.ic_type == 1 && (in.test_case == 1 || (in.test_case == 4 && D == 3)));
  if (!tc_ok) raise("unsupported initial condition (ic_type 0, or ic_type 1 with test_case 1 / 4 (3D))");
  std::vector<double> U((size_t)E * V * nSpts, 0.0);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; e++)
    for (int s = 0; s < nSpts; s++) {
      double u[5] = {0, 0, 0, 0, 0};
      if (in.ic_type == 0) {
        u[0] = in.rho_fs;
        double Vsq = 0.0;
        for (int d = 0; d < D; d++) { u[d + 1] = in.rho_fs * in.V_fs[d]; Vsq += in.V_fs[d] * in.V_fs[d]; }
        u[D + 1] = in.P_fs / (in.gamma - 1.0) + 0.5 * in.rho_fs * Vsq;
      } else {
        double xyz[3] = {0, 0, 0};
        point_coord(e, false, s, xyz);
        const double x = xyz[0], y = xyz[1], z = xyz[2];
        if (in.test_case == 1) {  // isentropic vortex, src/funcs.cpp:63-75
          const double G = 5.0, R = 1.;
          const double f = (1.0 - x * x - y * y) / R;
          const double rho = std::pow(1.0 - (G * G * (in.gamma - 1.)) / (8.0 * in.gamma * M_PI * M_PI) * std::exp(f),
                                      1.0 / (in.gamma - 1.0));
          const double Vx = 1.0 - G * y / (2.0 * M_PI) * std::exp(0.5 * f);
          const double Vy = 1.0 + G * x / (2.0 * M_PI) * std::exp(0.5 * f);
          const double P = std::pow(rho, in.gamma);
          u[0] = rho; u[1] = rho * Vx; u[2] = rho * Vy;
          u[D + 1] = P / (in.gamma - 1.0) + 0.5 * rho * (Vx * Vx + Vy * Vy);
        } else {  // Taylor-Green vortex, src/funcs.cpp:125-155
          const double L = in.L_fs, U0 = in.v_mag_fs, Rho0 = in.rho_fs, P0 = in.P_fs;
          const double uu = U0 * std::sin(x / L) * std::cos(y / L) * std::cos(z / L);
          const double vv = -U0 * std::cos(x / L) * std::sin(y / L) * std::cos(z / L);
          const double p = P0 + Rho0 * U0 * U0 / 16. * (std::cos(2 * x / L) + std::cos(2 * y / L)) * (std::cos(2 * z / L) + 2);
          const double rho = p / (in.R * in.T_fs);
          u[0] = rho; u[1] = rho * uu; u[2] = rho * vv; u[3] = 0.0;
          u[4] = p / (in.gamma - 1.0) + 0.5 * rho * (uu * uu + vv * vv);
        }
      }
      for (int v = 0; v < V; v++) U[((size_t)e * V + v) * nSpts + s] = u[v];
    }
  return U;
}

zfr_ele_desc_t Elements::desc(const zefr_input_t &in, int phase) const {
  zfr_ele_desc_t d;
  std::memset(&d, 0, sizeof(d));
  d.nDims = nDims; d.order = order; d.nVars = nVars; d.nEles = nEles; d.nSpts = nSpts; d.nFpts = nFpts;
  d.viscous = in.viscous;
  d.eles_per_block = tile_for(phase);
  d.gamma = in.gamma; d.prandtl = in.prandtl; d.mu = in.mu; d.rt = in.rt; d.c_sth = in.c_sth;
  d.fix_vis = in.fix_vis;
  d.rus_k = in.rus_k; d.ldg_b = in.ldg_b; d.ldg_tau = in.ldg_tau;
  d.ops1d = ops_d.as<double>();
  d.metric = metric_d.as<double>();
  d.fgeo = fgeo_d.as<double>();
  d.nbr = nbr_d.as<unsigned>();
  d.Utr = Utr_d.as<double>();
  d.Gtr = Gtr_d.as<double>();
  d.Ftr = Ftr_d.as<double>();
  d.fconv_type = (int)in.fconv_type;
  return d;
}

// ---------------------------------------------------------------- Faces
// Normal and area at every element flux point (src/elements.cpp:876-941):
// n = adj(J)^T tnorm, dA(slot) = |n|; the unit normal of a global flux point is the one
// seen from its slot-0 element.  Pass 1 stores each side's own normal, pass 2 replaces
// slot-1 normals by the partner's.
void Faces::setup(const Elements &el, const Connectivity &c) {
  const int D = el.nDims, E = el.nEles, nF = el.nFpts, R = D + 1;
  const int B = nbr_fpt_bits(nF);
  if ((long)E >= (1L << (31 - B)))
    raise("element count exceeds the " + std::to_string(31 - B) + "-bit partner index");
  nGfpts = c.nGfpts;
  fgeo.assign((size_t)E * R * nF, 0.0);
  nbr.assign((size_t)E * nF, 0u);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; e++)
    for (int f = 0; f < nF; f++) {
      double J[3][3], A[3][3];
      el.jacobian(e, true, f, J);
      Elements::adjugate(D, J, A);
      double n[3] = {0, 0, 0};
      for (int d1 = 0; d1 < D; d1++)
        for (int d2 = 0; d2 < D; d2++) n[d1] += A[d2][d1] * el.re.tnorm[(size_t)f * D + d2];
      double a = 0.0;
      for (int d = 0; d < D; d++) a += n[d] * n[d];
      a = std::sqrt(a);
      const int s = c.slot[(size_t)f * E + e];
      for (int d = 0; d < D; d++) fgeo[((size_t)e * R + d) * nF + f] = n[d] / a;
      fgeo[((size_t)e * R + D) * nF + f] = s ? -a : a;
      const int g = c.fpt2gfpt[(size_t)f * E + e];
      const int pe = c.gf_ele[(size_t)(1 - s) * c.nGfpts + g], pf = c.gf_fpt[(size_t)(1 - s) * c.nGfpts + g];
      nbr[(size_t)e * nF + f] = ((unsigned)pe << (B + 1)) | ((unsigned)s << B) | (unsigned)pf;
    }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; e++)
    for (int f = 0; f < nF; f++) {
      const unsigned p = nbr[(size_t)e * nF + f];
      if (((p >> B) & 1u) == 0) continue;
      const size_t pe = p >> (B + 1), pf = p & ((1u << B) - 1u);
      for (int d = 0; d < D; d++) fgeo[((size_t)e * R + d) * nF + f] = fgeo[(pe * R + d) * nF + pf];
    }
}

std::vector<double> Faces::export_array(const std::string &name, const Elements &el, const Connectivity &c) const {
  const int D = el.nDims, E = el.nEles, nF = el.nFpts, R = D + 1;
  std::vector<double> out;
  if (name == "norm") out.assign((size_t)D * nGfpts, 0.0);
  else if (name == "dA") out.assign((size_t)2 * nGfpts, 0.0);
  else raise("unknown face array " + name);
  for (int e = 0; e < E; e++)
    for (int f = 0; f < nF; f++) {
      const int g = c.fpt2gfpt[(size_t)f * E + e], s = c.slot[(size_t)f * E + e];
      if (name == "norm") {
        if (s == 0)
          for (int d = 0; d < D; d++) out[(size_t)d * nGfpts + g] = fgeo[((size_t)e * R + d) * nF + f];
      } else {
        out[(size_t)s * nGfpts + g] = std::abs(fgeo[((size_t)e * R + D) * nF + f]);
      }
    }
  return out;
}

// ---------------------------------------------------------------- FRSolver
FRSolver::FRSolver(zefr_input_t *in, int ord) : input(in), order(ord < 0 ? (int)in->order : ord) {}

FRSolver::~FRSolver() {
  drop_graph();
  if (stream && own_stream) (void)hipStreamDestroy(stream);
}

namespace {
// Setup stage timing, printed when ZEFR_VERBOSE is set (the reference prints its setup
// stages unconditionally, src/solver.cpp:62-180).
struct StageClock {
  bool on = std::getenv("ZEFR_VERBOSE") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void lap(const char *what) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[zefr] %-28s %8.3f s\n", what, std::chrono::duration<double>(now - t).count());
    t = now;
  }
};
}  // namespace

void FRSolver::setup_host() {
  if (input->equation != 1) raise("only the EulerNS equation set is supported");
  if (input->nStages != 1 && input->nStages != 4) raise("nStages must be 1 (Euler) or 4 (RK44)");
  StageClock clk;
  mesh = read_gmsh(input->meshfile, *input);
  clk.lap("read mesh + periodic pairs");
  eles = std::make_shared<Elements>();
  Elements &el = *eles;
  el.re.build((int)input->nDims, order);
  el.etype = mesh.etype; el.nDims = mesh.nDims; el.order = order; el.nEles = mesh.nEles;
  el.nSpts = el.re.nSpts; el.nFpts = el.re.nFpts; el.nVars = mesh.nDims + 2; el.nNodes = mesh.nNodesPerEle;
  el.setup_geometry(mesh);
  clk.lap("shape tables + metrics");
  conn = build_connectivity(mesh, el.re, [&el](int e, int f, double *x) { el.point_coord(e, true, f, x); });
  clk.lap("connectivity");
  faces = std::make_shared<Faces>();
  faces->setup(el, conn);
  clk.lap("face normals + partner map");
  el.U_host = el.initial_state(*input);
  clk.lap("initial state");
}

void FRSolver::setup_device() {
  Elements &el = *eles;
  const int D = el.nDims, E = el.nEles, nF = el.nFpts, V = el.nVars;
  StageClock clk;
  if (!stream) hip_or_throw(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  upload(el.ops_d, el.re.ops1d);
  upload(el.metric_d, el.metric);
  upload(el.fgeo_d, faces->fgeo);
  upload(el.nbr_d, faces->nbr);
  const size_t st = el.state_size();
  upload(el.U_spts_d, el.U_host);
  el.U_ini_d.alloc(st * sizeof(double));
  el.divF_d.alloc((size_t)input->nStages * st * sizeof(double));
  hip_or_throw(hipMemset(el.divF_d.p, 0, el.divF_d.bytes), "hipMemset");
  el.Utr_d.alloc((size_t)E * V * nF * sizeof(double));
  if (!onesided() && input->nStages > 1) el.Utr2_d.alloc((size_t)E * V * nF * sizeof(double));
  if (onesided()) el.Ftr_d.alloc((size_t)E * V * nF * sizeof(double));
  else if (input->viscous) el.Gtr_d.alloc((size_t)E * D * V * nF * sizeof(double));
  norm_work_d.alloc(1024 * 8 * sizeof(double));
  norm_out_d.alloc(8 * sizeof(double));
  device_ready = true;
  clk.lap("device upload");
}

void FRSolver::synchronize() { hip_or_throw(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

static void need_device(const FRSolver &s) {
  if (!s.device_ready) raise("solver is not set up on the device (call setup_solver)");
}

bool FRSolver::onesided() const {
  return input->viscous && (input->ldg_b == 0.5 || input->ldg_b == -0.5);
}

// compute_residual (src/solver.cpp:1494-1645) and its split (:1647-1840):
//   start  = U traces (extrapolate_U);
//   mid    = one-sided LDG: gradient, volume divergence and owned common fluxes;
//            general LDG beta: gradient and its traces;
//   finish = one-sided LDG: face correction of divF;
//            otherwise: common flux + full divergence.
const double *FRSolver::cur_state() const { return state_in ? state_in : eles->U_spts_d.as<double>(); }

void FRSolver::compute_residual_start(unsigned, int color) {
  need_device(*this);
  if (color > -1) raise("multi-colour (MCGS) residuals are not supported");
  const auto d = eles->desc(*input, 0);
  // the one-sided and the Euler kernels extrapolate the partner traces they need themselves
  if (input->viscous && !onesided()) status_or_throw(zfr_extrapolate(&d, cur_state(), stream));
}

void FRSolver::compute_residual_mid(unsigned stage, int) {
  if (!input->viscous) return;
  if (stage >= input->nStages) raise("stage out of range");
  auto d = eles->desc(*input, 1);
  if (trace_in) d.Utr = trace_in;
  if (onesided()) {
    double *divF = eles->divF_d.as<double>() + (size_t)stage * eles->state_size();
    // with trace_in (RK stages 1..), the previous stage's update has written the traces
    if (trace_in) status_or_throw(zfr_gradient_flux_traced(&d, cur_state(), divF, stream));
    else status_or_throw(zfr_gradient_flux(&d, cur_state(), divF, stream));
  } else {
    status_or_throw(zfr_gradient(&d, cur_state(), nullptr, stream));
  }
}

void FRSolver::compute_residual_finish(unsigned stage, int) {
  if (stage >= input->nStages) raise("stage out of range");
  auto d = eles->desc(*input, 2);
  if (trace_in) d.Utr = trace_in;
  double *divF = eles->divF_d.as<double>() + (size_t)stage * eles->state_size();
  if (onesided()) status_or_throw(zfr_face_correction(&d, divF, stream));
  else status_or_throw(zfr_divergence(&d, cur_state(), divF, stream));
}

void FRSolver::compute_residual(unsigned stage, int color) {
  if (stage >= input->nStages) raise("stage out of range");
  compute_residual_start(stage, color);
  compute_residual_mid(stage, color);
  compute_residual_finish(stage, color);
}

void FRSolver::residual_at(const double *U, double *divF, hipStream_t st) {
  need_device(*this);
  const auto d0 = eles->desc(*input, 0), d1 = eles->desc(*input, 1), d2 = eles->desc(*input, 2);
  if (onesided()) {
    status_or_throw(zfr_gradient_flux(&d1, U, divF, st));
    status_or_throw(zfr_face_correction(&d2, divF, st));
  } else {
    if (input->viscous) {
      status_or_throw(zfr_extrapolate(&d0, U, st));
      status_or_throw(zfr_gradient(&d1, U, nullptr, st));
    }
    status_or_throw(zfr_divergence(&d2, U, divF, st));  // Euler: extrapolates the partner traces itself
  }
}

void FRSolver::residual_vjp(const double *U, const double *w, double *v, hipStream_t st) {
  need_device(*this);
  Elements &el = *eles;
  const size_t tr = (size_t)el.nEles * el.nVars * el.nFpts * sizeof(double);
  if (!Wtr_d.p) Wtr_d.alloc(tr);
  if (!Xtr_d.p) Xtr_d.alloc(tr);
  const auto d = el.desc(*input);
  status_or_throw(zfr_residual_vjp(&d, U, w, v, Wtr_d.as<double>(), Xtr_d.as<double>(), st));
}

void FRSolver::compute_gradient_export() {
  need_device(*this);
  if (!input->viscous) raise("the gradient is only formed for viscous runs");
  Elements &el = *eles;
  if (!el.dU_spts_d.p) el.dU_spts_d.alloc((size_t)el.nDims * el.state_size() * sizeof(double));
  if (!el.Gtr_d.p) el.Gtr_d.alloc((size_t)el.nEles * el.nDims * el.nVars * el.nFpts * sizeof(double));
  const auto d = el.desc(*input);
  status_or_throw(zfr_extrapolate(&d, el.U_spts_d.as<double>(), stream));
  status_or_throw(zfr_gradient(&d, el.U_spts_d.as<double>(), el.dU_spts_d.as<double>(), stream));
}

static const double kRkAlpha[3] = {0.5, 0.5, 1.0};                         // src/solver.cpp:726-727
static const double kRkBeta[4] = {1. / 6., 1. / 3., 1. / 3., 1. / 6.};     // :732-734

void FRSolver::rk_stage_update(int stage) {
  const auto d = eles->desc(*input);
  const size_t st = eles->state_size();
  status_or_throw(zfr_rk_stage_source(&d, eles->U_spts_d.as<double>(), eles->U_ini_d.as<double>(),
                                      eles->divF_d.as<double>() + (size_t)stage * st, rk_source, kRkAlpha[stage],
                                      input->dt, stream));
}

void FRSolver::rk_final_update() {
  const auto d = eles->desc(*input);
  const int ns = (int)input->nStages;
  const double one = 1.0;
  const double *beta = (ns > 1) ? kRkBeta : &one;
  const double *Uini = (ns > 1) ? eles->U_ini_d.as<double>() : eles->U_spts_d.as<double>();
  status_or_throw(zfr_rk_final_source(&d, eles->U_spts_d.as<double>(), Uini, eles->divF_d.as<double>(), rk_source,
                                      beta, ns, input->dt, stream));
}

// compute_residual_finish with the RK update of this stage done by the same kernel; with
// trace_out, that kernel also writes the new state's trace there (stage updates only).
// split: the step keeps its starting state in U_spts_d until the final combination and
// its stage states in U_ini_d (see step_RK_eager).
void FRSolver::finish_with_rk(int stage, double *trace_out, bool split) {
  auto d = eles->desc(*input, 2);
  if (trace_in) d.Utr = trace_in;
  const int ns = (int)input->nStages;
  const size_t st = eles->state_size();
  double *const A = eles->U_spts_d.as<double>(), *const I = eles->U_ini_d.as<double>();
  zfr_rk_args_t r{};
  r.U = (split && stage < ns - 1) ? I : A;
  r.U_ini = (split || ns == 1) ? A : I;
  r.divF0 = eles->divF_d.as<double>();
  r.stage = stage;
  r.nStages = ns;
  r.dt = input->dt;
  r.source = rk_source;
  if (split && stage == 0) r.U_ini_out = rk_save;
  if (stage == 0 && rk_form_source) r.source_out = rk_form_source;  // source -= divF(0) first
  if (stage < ns - 1) {
    r.mode = 1;
    r.alpha = kRkAlpha[stage];
    r.trace_out = trace_out;
  } else {
    r.mode = 2;
    for (int s = 0; s < ns; s++) r.beta[s] = (ns > 1) ? kRkBeta[s] : 1.0;
    if (split) r.U_ini_out = I;  // the starting state, where the unfused step leaves it
    r.corr_out = rk_corr;
    r.corr_base = rk_corr_base;
    if (rk_corr && rk_prolong.out) {
      r.prolong_A = rk_prolong.A;
      r.prolong_n = rk_prolong.n;
      r.prolong_alpha = rk_prolong.alpha;
      r.prolong_out = rk_prolong.out;
    }
  }
  double *divF = eles->divF_d.as<double>() + (size_t)stage * st;
  if (onesided()) status_or_throw(zfr_face_correction_rk(&d, divF, &r, stream));
  else status_or_throw(zfr_divergence_rk(&d, cur_state(), divF, &r, stream));
}

// One RK step.  Unfused: U_ini = U, then per stage a residual and an update kernel.
// Fused with several stages (split): the starting state stays in U_spts_d, stage 0's
// residual is formed from it and its update writes the stage state into U_ini_d; stages
// 1.. update U_ini_d in place; the final combination writes U_spts_d and stores the
// starting state into U_ini_d (so U_ini_d ends as the unfused step leaves it, and no copy
// precedes the step).  Each stage update also writes the next stage's trace, so only
// stage 0 forms it from the state (the extrapolation kernel, or on the one-sided path the
// gradient_flux kernel itself).  On the one-sided path the writing kernel
// (face_correction) reads no trace and writes Utr_d in place, and stages 1.. read it
// (gradient_flux_traced).  On the Euler and general-LDG paths the writing kernel
// (divergence) reads the partners' traces, so the stages alternate Utr_d and Utr2_d.  An
// Euler divergence whose epilogue writes the state it reads (stages 1.., and the final
// combination, which writes U_ini_d) reads the partners' traces instead of their state;
// stage 0 of a split step writes only U_ini_d and extrapolates them itself.
void FRSolver::step_RK_eager() {
  const int ns = (int)input->nStages;
  const bool split = fuse_rk && ns > 1;
  if (ns > 1 && !split)
    hip_or_throw(hipMemcpyAsync(eles->U_ini_d.p, eles->U_spts_d.p, eles->U_spts_d.bytes, hipMemcpyDeviceToDevice,
                                stream), "hipMemcpyAsync");
  if (rk_save && !split)
    hip_or_throw(hipMemcpyAsync(rk_save, eles->U_spts_d.p, eles->U_spts_d.bytes, hipMemcpyDeviceToDevice, stream),
                 "hipMemcpyAsync");
  double *const tr[2] = {eles->Utr_d.as<double>(),
                         onesided() ? eles->Utr_d.as<double>() : eles->Utr2_d.as<double>()};
  const bool chain = fuse_rk && ns > 1 && tr[1] != nullptr;
  for (int stage = 0; stage < ns; stage++) {
    if (fuse_rk) {
      trace_in = (chain && stage > 0) ? tr[stage % 2] : nullptr;
      state_in = (split && stage > 0) ? eles->U_ini_d.as<double>() : nullptr;
      try {
        // the fused divergence reads a trace (viscous, or Euler writing the state it
        // reads): form it at stage 0, or at every stage without the chain
        const bool reads_tr = input->viscous || !split || stage > 0;
        if ((!chain || stage == 0) && !onesided() && reads_tr) {
          const auto d0 = eles->desc(*input, 0);
          status_or_throw(zfr_extrapolate(&d0, cur_state(), stream));
        }
        compute_residual_mid(stage);
        finish_with_rk(stage, (chain && stage < ns - 1) ? tr[(stage + 1) % 2] : nullptr, split);
      } catch (...) {
        trace_in = nullptr;
        state_in = nullptr;
        throw;
      }
      trace_in = nullptr;
      state_in = nullptr;
    } else {
      compute_residual(stage);
      if (stage < ns - 1) rk_stage_update(stage);
      else rk_final_update();
    }
  }
  if (rk_corr && !fuse_rk)
    status_or_throw(zfr_axpby((long)eles->state_size(), 1.0, eles->U_spts_d.as<double>(), -1.0, rk_corr_base,
                              rk_corr, stream));
}

void FRSolver::drop_graph() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  graph_exec = nullptr;
  graph = nullptr;
}

void FRSolver::set_step_mode(bool fuse, bool use) {
  fuse_rk = fuse;
  use_graph = use;
  drop_graph();
}

// step_RK (src/solver.cpp:2865-3097) with a supplied dt.  With use_graph the step's
// launches (the U_ini copy, three kernels per stage) are captured once and replayed; the
// kernel arguments bake in dt and the tile, so a change of either recaptures.
void FRSolver::step_RK() {
  need_device(*this);
  if (!use_graph) {
    step_RK_eager();
    return;
  }
  const std::array<const void *, 6> extra = {rk_save, rk_corr, rk_corr_base, rk_prolong.A, rk_prolong.out,
                                             rk_form_source};
  if (!graph_exec || graph_dt != input->dt || graph_tile != eles->tile_key() || graph_fused != fuse_rk ||
      graph_source != rk_source || graph_extra != extra || graph_pro_alpha != rk_prolong.alpha ||
      graph_pro_n != rk_prolong.n) {
    drop_graph();
    hip_or_throw(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      step_RK_eager();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hip_or_throw(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
    hip_or_throw(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    graph_dt = input->dt;
    graph_tile = eles->tile_key();
    graph_fused = fuse_rk;
    graph_source = rk_source;
    graph_extra = extra;
    graph_pro_alpha = rk_prolong.alpha;
    graph_pro_n = rk_prolong.n;
  }
  hip_or_throw(hipGraphLaunch(graph_exec, stream), "hipGraphLaunch");
}

// FRSolver::restart_solution (src/solver.cpp:212-222) for PyFR files.
void FRSolver::restart_solution() {
  if (input->restart_type == 2) restart_pyfr(input->restart_case, input->restart_iter);
  else if (input->restart_type == 1) raise("ParaView restart files are not supported; use restart_type 2 (PyFR)");
  else raise("restart_type must be 2 (PyFR .pyfrs files)");
}

void FRSolver::update() { update(nullptr, nullptr, nullptr, nullptr); }

void FRSolver::update(const double *source) { update(source, nullptr, nullptr, nullptr); }

void FRSolver::update(const double *source, double *save, double *corr, const double *corr_base) {
  update(source, save, corr, corr_base, Prolong{});
}

void FRSolver::update(const double *source, double *save, double *corr, const double *corr_base,
                      const Prolong &pro, bool form_source) {
  if (corr && !corr_base) raise("update: corr needs corr_base");
  if (pro.out && (!fuse_rk || !corr)) raise("update: the prolongation is fused into a fused step with corr");
  if (form_source && (!fuse_rk || input->nStages < 2 || !source))
    raise("update: the source term is formed by stage 0 of a fused step of several stages, from a source buffer");
  prev_time = flow_time;
  rk_source = source;
  rk_save = save;
  rk_corr = corr;
  rk_corr_base = corr_base;
  rk_prolong = pro;
  // the buffer is the caller's own source array (PMGrid's sourcesBT), written in place
  rk_form_source = form_source ? const_cast<double *>(source) : nullptr;
  auto clear = [&] {
    rk_source = nullptr;
    rk_save = nullptr;
    rk_corr = nullptr;
    rk_corr_base = nullptr;
    rk_prolong = Prolong{};
    rk_form_source = nullptr;
  };
  try {
    step_RK();
  } catch (...) {
    clear();
    throw;
  }
  clear();
  flow_time = prev_time + input->dt;
  current_iter++;
}

void FRSolver::residual_restrict(const double *source, const double *R, int n_out, double *Uc, double *Rc) {
  if (!fuse_rk) raise("residual_restrict: the restriction is fused into the residual's last kernel (fuse_rk)");
  compute_residual_start(0);
  compute_residual_mid(0);
  auto d = eles->desc(*input, 2);
  zfr_rk_args_t r{};
  r.mode = source ? 4 : 0;
  r.source = source;
  r.restrict_A = R;
  r.restrict_n = n_out;
  r.restrict_U = Uc;
  r.restrict_divF = Rc;
  r.restrict_src = cur_state();
  double *divF = eles->divF_d.as<double>();
  if (onesided()) status_or_throw(zfr_face_correction_rk(&d, divF, &r, stream));
  else status_or_throw(zfr_divergence_rk(&d, cur_state(), divF, &r, stream));
}

void FRSolver::residual_source(int mode, double *source_out, const double *source) {
  if (mode != 3 && mode != 4) raise("residual_source: mode must be 3 or 4");
  const long n = (long)eles->state_size();
  double *divF = eles->divF_d.as<double>();
  if (!fuse_rk) {
    compute_residual(0);
    if (mode == 3) status_or_throw(zfr_axpby(n, 1.0, source_out, -1.0, divF, source_out, stream));
    else status_or_throw(zfr_axpby(n, 1.0, divF, 1.0, source, divF, stream));
    return;
  }
  compute_residual_start(0);
  compute_residual_mid(0);
  auto d = eles->desc(*input, 2);
  zfr_rk_args_t r{};
  r.mode = mode;
  r.source = source;
  r.source_out = source_out;
  if (onesided()) status_or_throw(zfr_face_correction_rk(&d, divF, &r, stream));
  else status_or_throw(zfr_divergence_rk(&d, cur_state(), divF, &r, stream));
}

// step_RK_stage (src/solver.cpp:3458-3690)
void FRSolver::step_RK_stage(int stage) {
  step_RK_stage_start(stage);
  step_RK_stage_mid(stage);
  step_RK_stage_finish(stage);
}

void FRSolver::step_RK_stage_start(int stage) {
  need_device(*this);
  const int ns = (int)input->nStages;
  if (stage < 0 || stage >= ns) raise("stage out of range");
  if (ns > 1 && stage == 0) {
    prev_time = flow_time;
    hip_or_throw(hipMemcpyAsync(eles->U_ini_d.p, eles->U_spts_d.p, eles->U_spts_d.bytes, hipMemcpyDeviceToDevice,
                                stream), "hipMemcpyAsync");
  }
  compute_residual_start(stage);
}

void FRSolver::step_RK_stage_mid(int stage) { compute_residual_mid(stage); }

void FRSolver::step_RK_stage_finish(int stage) {
  const int ns = (int)input->nStages;
  compute_residual_finish(stage);
  if (stage < ns - 1) {
    rk_stage_update(stage);
  } else {
    rk_final_update();
    flow_time = prev_time + input->dt;
    current_iter++;
  }
}

// report_residuals (src/solver.cpp:6990-7137): norm of divF(0)/detJ over all spts,
// reduced on the device, divided by nDoF = nSpts * nEles.
void FRSolver::report_residuals(double *res) {
  need_device(*this);
  const auto d = eles->desc(*input);
  const int V = eles->nVars;
  status_or_throw(zfr_residual_norm(&d, eles->divF_d.as<double>(), (int)input->res_type,
                                    norm_work_d.as<double>(), norm_out_d.as<double>(), stream));
  std::vector<double> r(V);
  hip_or_throw(hipMemcpyAsync(r.data(), norm_out_d.p, V * sizeof(double), hipMemcpyDeviceToHost, stream), "D2H");
  synchronize();
  const double nDoF = (double)eles->nSpts * eles->nEles;
  for (int v = 0; v < V; v++) {
    if (std::isnan(r[v])) raise("NaN in residual");
    res[v] = ((input->res_type == 2) ? std::sqrt(r[v]) : r[v]) / nDoF;
  }
  res_max = res[std::min<int>(input->res_field, V - 1)];
}

// Time the residual with each feasible tile and keep the fastest.
// Workgroup tiles by measurement: every candidate (1, 2, 4, ... elements per workgroup up
// to 256 threads, and the largest tile of 1..4 full waves) is timed per launch with
// time_kernels, twice in alternation; each of the residual's three launches keeps its own
// fastest tile.
// Returns the tile of the middle (dominant) launch.
int FRSolver::autotune(int reps) {
  need_device(*this);
  Elements &el = *eles;
  std::vector<int> cands;
  for (int eb = 1; eb * el.nSpts <= 256; eb *= 2) cands.push_back(eb);
  // and the largest tile that fits 1, 2, 3, 4 waves (e.g. 2, 5, 7, 10 elements of 25
  // points), which leave the fewest lanes idle
  for (int w = 1; w <= 4; w++) {
    const int eb = 64 * w / el.nSpts;
    if (eb >= 1 && std::find(cands.begin(), cands.end(), eb) == cands.end()) cands.push_back(eb);
  }
  el.tile_phase[0] = el.tile_phase[1] = el.tile_phase[2] = 0;
  double best[3] = {1e300, 1e300, 1e300};
  int best_eb[3] = {el.tile, el.tile, el.tile};
  for (int pass = 0; pass < 2; pass++)
    for (int eb : cands) {
      el.tile = eb;
      double ms[3];
      time_kernels(1, ms);  // warm-up at this shape
      time_kernels(std::max(reps, 1), ms);
      for (int k = 0; k < 3; k++)
        if (ms[k] > 0.0 && ms[k] < best[k]) { best[k] = ms[k]; best_eb[k] = eb; }
    }
  // an inviscid residual has no middle launch: phase 1 timed only the gap between two
  // events, so its tile is the divergence launch's
  if (!input->viscous) best_eb[1] = best_eb[2];
  // one-sided and Euler residuals have no trace launch, so phase 0 timed nothing; its tile
  // (used by a fused RK step's stage-0 extrapolation) falls back to the solver-wide tile
  if (!(input->viscous && !onesided())) best_eb[0] = 0;
  el.tile = best_eb[1];
  for (int k = 0; k < 3; k++) el.tile_phase[k] = best_eb[k];
  return best_eb[1];
}

// Residuals with an event pair around every launch (benchmark instrumentation).
void FRSolver::time_kernels(int reps, double *ms_out) {
  need_device(*this);
  Elements &el = *eles;
  const auto d0 = el.desc(*input, 0), d1 = el.desc(*input, 1), d2 = el.desc(*input, 2);
  reps = std::max(reps, 1);
  // one event before each launch and one after the last: residuals run back to back
  std::vector<hipEvent_t> ev(3 * reps + 1);
  for (auto &e : ev) hip_or_throw(hipEventCreate(&e), "hipEventCreate");
  double *U = el.U_spts_d.as<double>(), *divF = el.divF_d.as<double>();
  for (int r = 0; r < reps; r++) {
    hip_or_throw(hipEventRecord(ev[3 * r], stream), "hipEventRecord");
    if (input->viscous && !onesided()) status_or_throw(zfr_extrapolate(&d0, U, stream));  // else no trace launch
    hip_or_throw(hipEventRecord(ev[3 * r + 1], stream), "hipEventRecord");
    if (onesided()) status_or_throw(zfr_gradient_flux(&d1, U, divF, stream));
    else if (input->viscous) status_or_throw(zfr_gradient(&d1, U, nullptr, stream));
    hip_or_throw(hipEventRecord(ev[3 * r + 2], stream), "hipEventRecord");
    if (onesided()) status_or_throw(zfr_face_correction(&d2, divF, stream));
    else status_or_throw(zfr_divergence(&d2, U, divF, stream));
  }
  hip_or_throw(hipEventRecord(ev[3 * reps], stream), "hipEventRecord");
  hip_or_throw(hipEventSynchronize(ev[3 * reps]), "hipEventSynchronize");
  double acc[3] = {0, 0, 0};
  for (int i = 0; i < 3 * reps; i++) {
    float ms = 0.f;
    hip_or_throw(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "hipEventElapsedTime");
    acc[i % 3] += ms;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  for (int k = 0; k < 3; k++) ms_out[k] = acc[k] / reps;
}


}  // namespace zfr
