// zefr_oracle.cpp -- CPU restatement of ZEFR's FRSolver::compute_residual path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it restates, in plain
// C++17 (+ optional OpenMP), the reference's CPU algorithm for the residual hot path
// and the setup that feeds it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker / the timed CPU baseline.
// The product (zefr_amd/) never links, loads or calls anything here.
//
// Reference files followed (paths relative to the reference checkout):
//   Gmsh 2.2 reader            src/geometry.cpp:307-1157 (read_boundary_ids, read_node_coords,
//                              read_element_connectivity, read_boundary_faces)
//   periodic face pairing      src/geometry.cpp:1159-1255 (couple_periodic_bnds)
//   global flux points         src/geometry.cpp:1257-1744 (setup_global_fpts)
//   flux-point orientation     src/solver.cpp:312-437 (orient_fpts), include/funcs.hpp:254-276
//   points / polynomials       src/points.cpp:9-209, include/polynomials.hpp:15-76
//   quad / hex local data      src/quads.cpp:45-361,520-560; src/hexas.cpp:51-490,760-840
//   shape, coords, metrics     src/elements.cpp:214-388, 845-996
//   FR operators               src/elements.cpp:555-654
//   residual orchestration     src/solver.cpp:1494-1645
//   operator products          src/elements.cpp:1337-1501 (gimmik_mm_cpu: see op_mm below)
//   pointwise flux             src/elements.cpp:1588-1699, include/flux.hpp:1-202
//   common U / common F        src/faces.cpp:1244-1614
//   RK44 + stage updates       src/solver.cpp:713-735, 2865-3097
//   residual norm              src/solver.cpp:6990-7137
//
// Scope: linear quads (Gmsh type 3) and linear hexes (type 5), one element type per
// mesh, periodic or fluid physical groups only (any other boundary raises), Euler /
// Navier-Stokes with Rusanov + LDG, RK44 or forward Euler with a fixed dt.
//
// What pins each part (tests/test_oracle.py): operators against the reference's baked
// GiMMiK hex p=3/p=4 constants and polynomial exactness; Gauss-Legendre rules (1..10
// points) against numpy, where the reference's 10-point table keeps one node pair 5e-4
// off the true nodes (+-0.433895394129247; restated as the reference has it); metrics against affine closed forms and cell volumes; connectivity by brute
// force (each global flux point has one slot-0 and one slot-1 point at the same
// physical location); the slot of every element flux point against the first-incidence
// rule derived from the mesh file by an independent reader (periodic faces included; the
// swapped rule is rejected); fluxes against textbook Euler/NS from analytic gradients;
// residuals by freestream preservation and discrete conservation; RK44 and the norms
// by their textbook definitions.  Rounding: the same source built with
// -DORACLE_REAL="long double" is the exact answer the FP64 results are measured against.
// Parity unpinned: the exact global flux-point index permutation (setup_global_fpts +
// orient_fpts).  No reference-held output records it; it is restated from the cited
// code and checked only for being a valid pairing.  Residuals do not depend on it.
//
// gimmik_mm_cpu: its generated source (src/gimmik_cpu.c) is absent from the reference.
// GiMMiK bakes the nonzero entries of the constant operator into an unrolled dot product,
// so op_mm restates it as C = alpha*A*B + beta*C over the nonzeros of A, summed in
// increasing k.  Exact zeros of A are skipped, which is what the generated code does.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

// Arithmetic type of the oracle: double for the FP64 restatement; the extended-precision
// build (libzefr_oracle_ld.so, -DORACLE_REAL="long double") evaluates the same algorithm
// on the same double inputs with a 64-bit significand, as the exact answer that FP64
// evaluations are measured against.  The C interface below stays in double either way.
#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
using real = ORACLE_REAL;

namespace {

enum { LINE_T = 0, QUAD_T = 2, HEX_T = 4 };
enum { BC_PERIODIC = 1 };

struct Phys {
  int viscous;
  real gamma, prandtl, mu, rt, c_sth;
  int fix_vis;
  real rus_k, ldg_b, ldg_tau, dt;
  int nStages;
  int res_type;
  int fconv;  // 0 Rusanov (the reference's), 2 Roe (textbook, below)
};

[[noreturn]] void fail(const std::string &msg) { throw std::runtime_error(msg); }

// ---------------------------------------------------------------- points / polynomials
// src/points.cpp:9-98 (closed forms up to 5 points, 15-digit tables beyond).
std::vector<double> gl_points_d(unsigned n) {
  switch (n) {
    case 1: return {0.0};
    case 2: return {-std::sqrt(1. / 3.), std::sqrt(1. / 3.)};
    case 3: return {-std::sqrt(3. / 5.), 0.0, std::sqrt(3. / 5.)};
    case 4: {
      double a = std::sqrt((3. + 2. * std::sqrt(6. / 5.)) / 7.);
      double b = std::sqrt((3. - 2. * std::sqrt(6. / 5.)) / 7.);
      return {-a, -b, b, a};
    }
    case 5: {
      double a = (1. / 3.) * std::sqrt(5. + 2. * std::sqrt(10. / 7.));
      double b = (1. / 3.) * std::sqrt(5. - 2. * std::sqrt(10. / 7.));
      return {-a, -b, 0.0, b, a};
    }
    case 6:
      return {-0.932469514203152, -0.661209386466264, -0.238619186083197,
              0.238619186083197,  0.661209386466264,  0.932469514203152};
    case 7:
      return {-0.949107912342758, -0.741531185599394, -0.405845151377397, 0.0,
              0.405845151377397,  0.741531185599394,  0.949107912342758};
    case 8:
      return {-0.960289856497536, -0.796666477413627, -0.525532409916329, -0.183434642495650,
              0.183434642495650,  0.525532409916329,  0.796666477413627,  0.960289856497536};
    case 9:
      return {-0.968160239507626, -0.836031107326636, -0.613371432700590, -0.324253423403809, 0.0,
              0.324253423403809,  0.613371432700590,  0.836031107326636,  0.968160239507626};
    case 10:
      return {-0.973906528517172, -0.865063366688985, -0.679409568299024, -0.433895394129247,
              -0.148874338981631, 0.148874338981631,  0.433895394129247,  0.679409568299024,
              0.865063366688985,  0.973906528517172};
    default: fail("oracle: Gauss-Legendre points only for 1..10 points");
  }
}

// src/points.cpp:123-209
std::vector<double> gl_weights_d(unsigned n) {
  switch (n) {
    case 1: return {2.0};
    case 2: return {1.0, 1.0};
    case 3: return {5. / 9., 8. / 9., 5. / 9.};
    case 4: {
      double a = 1. / 36. * (18. - std::sqrt(30.)), b = 1. / 36. * (18. + std::sqrt(30.));
      return {a, b, b, a};
    }
    case 5: {
      double a = 1. / 900. * (322 - 13. * std::sqrt(70.)), b = 1. / 900. * (322 + 13. * std::sqrt(70.));
      return {a, b, 128. / 225., b, a};
    }
    case 6:
      return {0.171324492379171, 0.360761573048139, 0.467913934572691,
              0.467913934572691, 0.360761573048139, 0.171324492379171};
    case 7:
      return {0.129484966168870, 0.279705391489277, 0.381830050505119, 0.417959183673469,
              0.381830050505119, 0.279705391489277, 0.129484966168870};
    case 8:
      return {0.101228536290377, 0.222381034453374, 0.313706645877887, 0.362683783378362,
              0.362683783378362, 0.313706645877887, 0.222381034453374, 0.101228536290377};
    case 9:
      return {0.081274388361575, 0.180648160694857, 0.260610696402936, 0.312347077040003, 0.330239355001260,
              0.312347077040003, 0.260610696402936, 0.180648160694857, 0.081274388361575};
    case 10:
      return {0.066671344308688, 0.149451349150581, 0.219086362515982, 0.269266719309996, 0.295524224714753,
              0.295524224714753, 0.269266719309996, 0.219086362515982, 0.149451349150581, 0.066671344308688};
    default: fail("oracle: Gauss-Legendre weights only for 1..10 points");
  }
}

// The same rules as the arithmetic type of this build: the double values widened, so
// that an extended-precision build evaluates the same discretisation as the FP64 one.
std::vector<real> gl_points(unsigned n) {
  const std::vector<double> g = gl_points_d(n);
  return std::vector<real>(g.begin(), g.end());
}

std::vector<real> gl_weights(unsigned n) {
  const std::vector<double> g = gl_weights_d(n);
  return std::vector<real>(g.begin(), g.end());
}

// include/polynomials.hpp:15-28 -- product form, factors taken in increasing grid order.
real lagrange(const std::vector<real> &g, unsigned mode, real x) {
  real v = 1.0;
  for (unsigned i = 0; i < g.size(); i++)
    if (i != mode) v *= (x - g[i]) / (g[mode] - g[i]);
  return v;
}

// include/polynomials.hpp:47-76 -- sum over j != mode of prod_{i != mode, j} (x - g_i), / denominator.
real lagrange_d1(const std::vector<real> &g, unsigned mode, real x) {
  real den = 1.0;
  for (unsigned i = 0; i < g.size(); i++)
    if (i != mode) den *= (g[mode] - g[i]);
  real val = 0.0;
  for (unsigned j = 0; j < g.size(); j++) {
    if (j == mode) continue;
    real term = 1.0;
    for (unsigned i = 0; i < g.size(); i++)
      if (i != mode && i != j) term *= (x - g[i]);
    val += term;
  }
  return val / den;
}

// ---------------------------------------------------------------- fuzzy lexicographic sort
// include/funcs.hpp:254-276: sort indices by coordinate 'dim'; runs whose values stay
// within tol of the run's first entry are re-sorted by the next coordinate.
// mat is [nDims][ncols] row-major (mat(dim, col)).
void fuzzysort_ind(const std::vector<real> &mat, unsigned ncols, unsigned *inds, unsigned n,
                   unsigned nDims, unsigned dim = 0, real tol = 1e-6) {
  auto val = [&](unsigned c) { return mat[(size_t)dim * ncols + c]; };
  std::sort(inds, inds + n, [&](unsigned a, unsigned b) { return val(a) < val(b); });
  unsigned i = 0, j;
  unsigned ix = inds[0];
  for (j = 1; j < n; j++) {
    unsigned jx = inds[j];
    if (val(jx) - val(ix) >= tol) {
      if (j - i > 1 && dim + 1 < nDims) fuzzysort_ind(mat, ncols, inds + i, j - i, nDims, dim + 1, tol);
      i = j;
      ix = jx;
    }
  }
  if (i != j && dim + 1 < nDims) fuzzysort_ind(mat, ncols, inds + i, j - i, nDims, dim + 1, tol);
}

// ---------------------------------------------------------------- operator product
// C[M][N] = alpha * A[M][K] * B[K][N] + beta * C, row-major; nonzeros of A only (see header).
void op_mm(int M, int N, int K, real alpha, const real *A, const real *B, real beta, real *C) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; i++) {
    real *c = C + (size_t)i * N;
    if (beta == 0.0)
      for (int j = 0; j < N; j++) c[j] = 0.0;
    else if (beta != 1.0)
      for (int j = 0; j < N; j++) c[j] *= beta;
    for (int k = 0; k < K; k++) {
      real a = A[(size_t)i * K + k];
      if (a == 0.0) continue;
      a *= alpha;
      const real *b = B + (size_t)k * N;
      for (int j = 0; j < N; j++) c[j] += a * b[j];
    }
  }
}

// ---------------------------------------------------------------- pointwise fluxes
// include/flux.hpp:11-63 with zero grid velocity.
template <int V, int D>
void fconv(const real *U, real F[V][D], real &P, real gamma) {
  if (D == 2) {
    real invrho = 1. / U[0];
    real momF = (U[1] * U[1] + U[2] * U[2]) * invrho;
    P = (gamma - 1.0) * (U[3] - 0.5 * momF);
    real H = (U[3] + P) * invrho;
    F[0][0] = U[1];
    F[1][0] = U[1] * U[1] * invrho + P;
    F[2][0] = U[1] * U[2] * invrho;
    F[3][0] = U[1] * H;
    F[0][1] = U[2];
    F[1][1] = U[2] * U[1] * invrho;
    F[2][1] = U[2] * U[2] * invrho + P;
    F[3][1] = U[2] * H;
  } else {
    real invrho = 1. / U[0];
    real momF = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * invrho;
    P = (gamma - 1.0) * (U[4] - 0.5 * momF);
    real H = (U[4] + P) * invrho;
    for (int d = 0; d < 3; d++) {
      F[0][d] = U[d + 1];
      for (int m = 0; m < 3; m++) F[m + 1][d] = U[d + 1] * U[m + 1] * invrho;
      F[d + 1][d] += P;
      F[4][d] = U[d + 1] * H;
    }
  }
}

// Roe's flux (Roe 1981) with Harten's entropy fix on the acoustic waves, delta = 0.1 a~.
// The reference has no Roe flux: this is a textbook restatement, not a restatement of
// reference code; tests/test_roe.py pins it against numpy's eigen-decomposition of the
// flux Jacobian at the Roe-averaged state.  Fn = F.n for the unit normal n.
template <int V, int D>
void roe_flux(const real *UL, const real *UR, const real *n, real gamma, real *Fn) {
  real FL[V][D], FR[V][D], PL, PR;
  fconv<V, D>(UL, FL, PL, gamma);
  fconv<V, D>(UR, FR, PR, gamma);
  real uL[D], uR[D];
  for (int d = 0; d < D; d++) { uL[d] = UL[d + 1] / UL[0]; uR[d] = UR[d + 1] / UR[0]; }
  const real HL = (UL[V - 1] + PL) / UL[0], HR = (UR[V - 1] + PR) / UR[0];
  const real rl = std::sqrt(UL[0]), rr = std::sqrt(UR[0]);
  real ut[D], unt = 0, q2 = 0, unL = 0, unR = 0;
  for (int d = 0; d < D; d++) {
    ut[d] = (rl * uL[d] + rr * uR[d]) / (rl + rr);
    unt += ut[d] * n[d]; q2 += ut[d] * ut[d];
    unL += uL[d] * n[d]; unR += uR[d] * n[d];
  }
  const real Ht = (rl * HL + rr * HR) / (rl + rr);
  const real rhot = rl * rr;
  const real at = std::sqrt((gamma - 1.0) * (Ht - 0.5 * q2));
  const real delta = 0.1 * at;
  auto fix = [&](real lam) {
    real l = std::abs(lam);
    return (l < delta) ? (lam * lam + delta * delta) / (2.0 * delta) : l;
  };
  const real lam[3] = {fix(unt - at), std::abs(unt), fix(unt + at)};
  const real dP = PR - PL, dUn = unR - unL, dRho = UR[0] - UL[0];
  // wave strengths and right eigenvectors: acoustic (-), entropy, acoustic (+), shear
  const real alpha[3] = {(dP - rhot * at * dUn) / (2.0 * at * at), dRho - dP / (at * at),
                           (dP + rhot * at * dUn) / (2.0 * at * at)};
  real r[3][V];
  r[0][0] = 1.0; r[1][0] = 1.0; r[2][0] = 1.0;
  for (int d = 0; d < D; d++) {
    r[0][d + 1] = ut[d] - at * n[d];
    r[1][d + 1] = ut[d];
    r[2][d + 1] = ut[d] + at * n[d];
  }
  r[0][V - 1] = Ht - at * unt; r[1][V - 1] = 0.5 * q2; r[2][V - 1] = Ht + at * unt;
  real diss[V] = {0};
  for (int k = 0; k < 3; k++)
    for (int m = 0; m < V; m++) diss[m] += lam[k] * alpha[k] * r[k][m];
  real sh[D], ush = 0;
  for (int d = 0; d < D; d++) { sh[d] = (uR[d] - uL[d]) - dUn * n[d]; ush += ut[d] * sh[d]; }
  for (int d = 0; d < D; d++) diss[d + 1] += lam[1] * rhot * sh[d];
  diss[V - 1] += lam[1] * rhot * ush;
  for (int m = 0; m < V; m++) {
    real fl = 0, fr = 0;
    for (int d = 0; d < D; d++) { fl += FL[m][d] * n[d]; fr += FR[m][d] * n[d]; }
    Fn[m] = 0.5 * (fl + fr) - 0.5 * diss[m];
  }
}

// include/flux.hpp:75-202: subtract the viscous flux.
template <int V, int D>
void fvisc_add(const real *U, const real dU[V][D], real F[V][D], const Phys &ph) {
  real invrho = 1.0 / U[0];
  real vel[3] = {0, 0, 0};
  real ke2 = 0.0;
  for (int d = 0; d < D; d++) { vel[d] = U[d + 1] * invrho; ke2 += vel[d] * vel[d]; }
  real e_int = U[D + 1] * invrho - 0.5 * ke2;
  real mu;
  if (ph.fix_vis) mu = ph.mu;
  else {
    real rt_ratio = (ph.gamma - 1.0) * e_int / ph.rt;
    mu = ph.mu * std::pow(rt_ratio, 1.5) * (1. + ph.c_sth) / (rt_ratio + ph.c_sth);
  }
  // dvel[i][j] = d(vel_i)/dx_j
  real dvel[3][3] = {{0}};
  for (int i = 0; i < D; i++)
    for (int j = 0; j < D; j++) dvel[i][j] = (dU[i + 1][j] - dU[0][j] * vel[i]) * invrho;
  real de[3];
  for (int j = 0; j < D; j++) {
    real s = 0.0;
    for (int i = 0; i < D; i++) s += vel[i] * dvel[i][j];
    real dke = 0.5 * ke2 * dU[0][j] + U[0] * s;
    de[j] = (dU[D + 1][j] - dke - dU[0][j] * e_int) * invrho;
  }
  real div = 0.0;
  for (int i = 0; i < D; i++) div += dvel[i][i];
  real diag = div / 3.0;
  real tau[3][3];
  for (int i = 0; i < D; i++)
    for (int j = 0; j < D; j++)
      tau[i][j] = (i == j) ? 2.0 * mu * (dvel[i][i] - diag) : mu * (dvel[i][j] + dvel[j][i]);
  for (int j = 0; j < D; j++) {
    real w = 0.0;
    for (int i = 0; i < D; i++) { F[i + 1][j] -= tau[i][j]; w += vel[i] * tau[i][j]; }
    F[D + 1][j] -= (w + (mu / ph.prandtl) * ph.gamma * de[j]);
  }
}

// ---------------------------------------------------------------- the oracle solver
struct Oracle {
  Phys ph;
  int nDims = 0, order = 0, nVars = 0;
  // mesh
  int nNodes = 0;
  std::vector<real> coord_nodes;  // [nNodes][nDims]
  int etype = 0, nEles = 0, nNodesPerEle = 0;
  std::vector<int> ele2nodes;  // [nEles][nNodesPerEle]
  std::map<std::vector<unsigned>, int> bnd_faces;
  std::map<std::vector<unsigned>, std::vector<unsigned>> per_bnd_pairs;
  // connectivity
  unsigned nFptsPerFace = 0, nGfpts = 0, nGfpts_int = 0, nFaces = 0, nIntFaces = 0;
  std::vector<int> fpt2gfpt;          // [nFpts][nEles]
  std::vector<signed char> fpt2slot;  // [nFpts][nEles]
  std::vector<int> face2fpts;         // [nFptsPerFace][nFaces]
  std::vector<int> fpt2face;          // [nGfpts]
  std::vector<int> gf_fpt, gf_ele;    // [2][nGfpts] face views
  // element local data
  unsigned nSpts1D = 0, nSpts = 0, nFpts = 0, nFaceEle = 0;
  std::vector<real> loc1d, locDFR, wts1d;
  std::vector<real> loc_spts, loc_fpts, tnorm;  // [n][nDims]
  std::vector<int> idx_spts, idx_fpts;            // [n][nDims]
  std::vector<real> weights_spts;
  std::vector<real> shape_spts, shape_fpts, dshape_spts, dshape_fpts;
  std::vector<real> nodes;  // [nNodesPerEle][nDims][nEles]
  std::vector<real> coord_spts, coord_fpts;  // [n][nDims][nEles]
  std::vector<real> jaco_spts, jaco_fpts, inv_jaco_spts, inv_jaco_fpts;  // [D][n][D][E]
  std::vector<real> jaco_det_spts;  // [nSpts][nEles]
  std::vector<real> norm, dA;       // [D][nGfpts], [2][nGfpts]
  std::vector<real> oppE, oppD, oppD_fpts, oppDiv, oppDiv_fpts;
  // solution data (reference layouts, nElesPad == nEles on the CPU path)
  std::vector<real> U_spts, U_fpts, Ucomm, F_spts, Fcomm, dU_spts, dU_fpts, divF, U_ini;
  real rk_alpha[3] = {0.5, 0.5, 1.0};
  real rk_beta[4] = {1. / 6., 1. / 3., 1. / 3., 1. / 6.};

  // ------------------------------------------------------------ Gmsh reader
  void read_mesh(const std::string &file) {
    std::ifstream f(file);
    if (!f.is_open()) fail("oracle: cannot open mesh file " + file);
    std::string line;
    auto seek = [&](const char *tag) {
      f.clear(); f.seekg(0);
      while (std::getline(f, line))
        if (line.find(tag) != std::string::npos) return;
      fail(std::string("oracle: mesh file missing ") + tag);
    };
    // geometry.cpp:333-406: physical names -> periodic / fluid only
    seek("$PhysicalNames");
    unsigned nPhys; f >> nPhys; std::getline(f, line);
    std::map<int, int> bcid2type;  // gmsh physical id -> BC_PERIODIC or -1 (fluid)
    for (unsigned i = 0; i < nPhys; i++) {
      std::getline(f, line);
      std::stringstream ss(line);
      int bdim, bid; std::string name;
      ss >> bdim >> bid >> name;
      name.erase(std::remove(name.begin(), name.end(), '"'), name.end());
      std::transform(name.begin(), name.end(), name.begin(), ::tolower);
      if (name == "fluid") {
        if (bdim != nDims) fail("oracle: nDims in mesh file does not match");
        bcid2type[bid] = -1;
      } else if (name == "periodic") bcid2type[bid] = BC_PERIODIC;
      else fail("oracle: unsupported boundary '" + name + "' (periodic meshes only)");
    }
    // geometry.cpp:408-429
    seek("$Nodes");
    f >> nNodes;
    coord_nodes.assign((size_t)nNodes * nDims, 0.0);
    for (int n = 0; n < nNodes; n++) {
      unsigned id;
      double x[3];  // parsed as double: every build sees the same node coordinates
      f >> id >> x[0] >> x[1] >> x[2];
      for (int d = 0; d < nDims; d++) coord_nodes[(size_t)n * nDims + d] = x[d];
    }
    // geometry.cpp:431-840 (element pass) and 842-950 (boundary pass)
    seek("$Elements");
    unsigned nTot; f >> nTot;
    etype = (nDims == 2) ? QUAD_T : HEX_T;
    nNodesPerEle = (nDims == 2) ? 4 : 8;
    const int eleGmsh = (nDims == 2) ? 3 : 5, bndGmsh = (nDims == 2) ? 1 : 3;
    const unsigned nBndNodes = (nDims == 2) ? 2 : 4;
    nEles = 0;
    for (unsigned n = 0; n < nTot; n++) {
      unsigned id, type, nTags; f >> id >> type >> nTags;
      std::vector<int> tags(nTags);
      for (auto &t : tags) f >> t;
      if ((int)type == eleGmsh) {
        for (int k = 0; k < nNodesPerEle; k++) { int v; f >> v; ele2nodes.push_back(v - 1); }
        nEles++;
      } else if ((int)type == bndGmsh) {
        std::vector<unsigned> face(nBndNodes);
        for (auto &v : face) { f >> v; v--; }
        if (nTags < 1 || !bcid2type.count(tags[0])) fail("oracle: boundary face without physical tag");
        int bt = bcid2type[tags[0]];
        if (bt != BC_PERIODIC) fail("oracle: fluid-tagged boundary face");
        std::sort(face.begin(), face.end());
        bnd_faces[face] = bt;
      } else {
        fail("oracle: unsupported Gmsh element type " + std::to_string(type));
      }
      std::getline(f, line);
    }
    if (!bnd_faces.empty()) couple_periodic_bnds();
  }

  // geometry.cpp:1159-1255
  // Each periodic face is paired with the first other periodic face (in the map's order)
  // of the same node count whose centroid matches in nDims-1 coordinates that the face
  // itself does not lie flat in.  The centroids and the per-face "flat in d" flags are
  // the reference's per-pair quantities, computed once per face with the same arithmetic;
  // the scan per face is independent, so the faces run in parallel.
  void couple_periodic_bnds() {
    std::vector<const std::vector<unsigned> *> pf;
    for (auto &bf : bnd_faces)
      if (bf.second == BC_PERIODIC) pf.push_back(&bf.first);
    const long np = (long)pf.size();
    std::vector<real> cen(np * nDims, 0.0);
    std::vector<char> flat(np * nDims, 1);
    for (long i = 0; i < np; i++) {
      const auto &f = *pf[i];
      const unsigned nn = f.size();
      for (int d = 0; d < nDims; d++) {
        real c = 0.0;
        for (unsigned k = 0; k < nn; k++) c += coord_nodes[f[k] * nDims + d];
        cen[i * nDims + d] = c / nn;
        const real c0 = coord_nodes[f[0] * nDims + d];
        for (unsigned k = 1; k < nn; k++)
          if (std::abs(c0 - coord_nodes[f[k] * nDims + d]) > 1.e-6) flat[i * nDims + d] = 0;
      }
    }
    std::vector<long> match(np, -1);
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < np; i++) {
      const size_t nn = pf[i]->size();
      for (long j = 0; j < np; j++) {
        if (j == i || pf[j]->size() != nn) continue;
        int count = 0;
        for (int d = 0; d < nDims; d++)
          if (std::abs(cen[i * nDims + d] - cen[j * nDims + d]) < 1.e-6 && !flat[i * nDims + d]) count++;
        if (count == nDims - 1) { match[i] = j; break; }
      }
    }
    for (long i = 0; i < np; i++) {
      if (match[i] < 0) fail("oracle: unpaired periodic face");
      per_bnd_pairs[*pf[i]] = *pf[match[i]];
    }
  }

  std::vector<std::vector<unsigned>> face_nodes() const {
    if (nDims == 2) return {{0, 1}, {1, 2}, {2, 3}, {3, 0}};           // geometry.cpp:218-229
    return {{0, 1, 2, 3}, {5, 4, 7, 6}, {0, 3, 7, 4}, {2, 1, 5, 6}, {1, 0, 4, 5}, {3, 2, 6, 7}};  // :247-265
  }

  // ------------------------------------------------------------ setup_global_fpts (geometry.cpp:1257-1744)
  void setup_global_fpts() {
    nFptsPerFace = (nDims == 2) ? (order + 1) : (order + 1) * (order + 1);
    auto fnodes = face_nodes();
    nFaceEle = fnodes.size();
    auto is_periodic = [&](const std::vector<unsigned> &f) {
      auto it = bnd_faces.find(f);
      return it != bnd_faces.end() && it->second == BC_PERIODIC;
    };
    // counting pass
    std::set<std::vector<unsigned>> unique_faces;
    nGfpts_int = 0;
    for (int e = 0; e < nEles; e++) {
      for (auto &fn : fnodes) {
        std::vector<unsigned> face(fn.size());
        for (size_t i = 0; i < fn.size(); i++) face[i] = ele2nodes[(size_t)e * nNodesPerEle + fn[i]];
        std::sort(face.begin(), face.end());
        if (is_periodic(face) && unique_faces.count(per_bnd_pairs[face])) face = per_bnd_pairs[face];
        if (!unique_faces.count(face)) {
          if (bnd_faces.count(face) && !is_periodic(face)) fail("oracle: non-periodic boundary");
          nGfpts_int += nFptsPerFace;
        }
        unique_faces.insert(face);
      }
    }
    nGfpts = nGfpts_int;
    unsigned nFacesTot = unique_faces.size();
    fpt2face.assign(nGfpts, -1);
    face2fpts.assign((size_t)nFptsPerFace * nFacesTot, -1);
    const unsigned nFptsEle = nFptsPerFace * nFaceEle;
    std::vector<int> e2f((size_t)nEles * nFptsEle, -1);
    std::vector<signed char> e2s((size_t)nEles * nFptsEle, -1);
    std::map<std::vector<unsigned>, std::vector<unsigned>> face_fpts;
    std::map<std::vector<unsigned>, int> nodes_to_face;
    unsigned gfpt = 0;
    nFaces = 0; nIntFaces = 0;
    for (int e = 0; e < nEles; e++) {
      unsigned f0 = 0;
      for (auto &fn : fnodes) {
        std::vector<unsigned> face(fn.size());
        for (size_t i = 0; i < fn.size(); i++) face[i] = ele2nodes[(size_t)e * nNodesPerEle + fn[i]];
        std::sort(face.begin(), face.end());
        if (is_periodic(face) && face_fpts.count(per_bnd_pairs[face])) face = per_bnd_pairs[face];
        if (!face_fpts.count(face)) {
          std::vector<unsigned> fpts(nFptsPerFace);
          for (auto &fp : fpts) fp = gfpt++;
          face_fpts[face] = fpts;
          for (unsigned i = 0; i < nFptsPerFace; i++) {
            e2f[(size_t)e * nFptsEle + f0 + i] = fpts[i];
            e2s[(size_t)e * nFptsEle + f0 + i] = 0;
          }
          for (auto fp : fpts) fpt2face[fp] = nFaces;
          for (unsigned j = 0; j < nFptsPerFace; j++) face2fpts[(size_t)j * nFacesTot + nFaces] = fpts[j];
          nodes_to_face[face] = nFaces;
          nFaces++;
        } else {
          auto &fpts = face_fpts[face];
          for (unsigned i = 0; i < nFptsPerFace; i++) {
            e2f[(size_t)e * nFptsEle + f0 + i] = fpts[nFptsPerFace - i - 1];
            e2s[(size_t)e * nFptsEle + f0 + i] = 1;
          }
          nIntFaces++;
        }
        f0 += nFptsPerFace;
      }
    }
    fpt2gfpt.assign((size_t)nFptsEle * nEles, 0);
    fpt2slot.assign((size_t)nFptsEle * nEles, 0);
    for (int e = 0; e < nEles; e++)
      for (unsigned f = 0; f < nFptsEle; f++) {
        fpt2gfpt[(size_t)f * nEles + e] = e2f[(size_t)e * nFptsEle + f];
        fpt2slot[(size_t)f * nEles + e] = e2s[(size_t)e * nFptsEle + f];
      }
  }

  // ------------------------------------------------------------ element local data
  void set_locs() {
    nSpts1D = order + 1;
    loc1d = gl_points(nSpts1D);
    wts1d = gl_weights(nSpts1D);
    locDFR = loc1d;
    locDFR.insert(locDFR.begin(), -1.0);
    locDFR.push_back(1.0);
    const unsigned N = nSpts1D;
    if (nDims == 2) {
      nSpts = N * N; nFpts = 4 * N;
      loc_spts.assign(nSpts * 2, 0.0); idx_spts.assign(nSpts * 2, 0);
      weights_spts.assign(nSpts, 0.0);
      unsigned s = 0;
      for (unsigned i = 0; i < N; i++)  // quads.cpp:72-84
        for (unsigned j = 0; j < N; j++, s++) {
          loc_spts[s * 2] = loc1d[j]; loc_spts[s * 2 + 1] = loc1d[i];
          idx_spts[s * 2] = j; idx_spts[s * 2 + 1] = i;
          weights_spts[s] = wts1d[i] * wts1d[j];
        }
      loc_fpts.assign(nFpts * 2, 0.0); idx_fpts.assign(nFpts * 2, 0);
      unsigned fp = 0;
      for (unsigned face = 0; face < 4; face++)  // quads.cpp:86-123
        for (unsigned j = 0; j < N; j++, fp++) {
          real *l = &loc_fpts[fp * 2]; int *x = &idx_fpts[fp * 2];
          switch (face) {
            case 0: l[0] = loc1d[j]; l[1] = -1.0; x[0] = j; x[1] = -1; break;
            case 1: l[0] = 1.0; l[1] = loc1d[j]; x[0] = N; x[1] = j; break;
            case 2: l[0] = loc1d[N - j - 1]; l[1] = 1.0; x[0] = N - j - 1; x[1] = N; break;
            case 3: l[0] = -1.0; l[1] = loc1d[N - j - 1]; x[0] = -1; x[1] = N - j - 1; break;
          }
        }
      tnorm.assign(nFpts * 2, 0.0);  // quads.cpp:167-190
      for (unsigned f = 0; f < nFpts; f++) {
        switch (f / N) {
          case 0: tnorm[f * 2 + 1] = -1.0; break;
          case 1: tnorm[f * 2] = 1.0; break;
          case 2: tnorm[f * 2 + 1] = 1.0; break;
          case 3: tnorm[f * 2] = -1.0; break;
        }
      }
    } else {
      nSpts = N * N * N; nFpts = 6 * N * N;
      loc_spts.assign(nSpts * 3, 0.0); idx_spts.assign(nSpts * 3, 0);
      weights_spts.assign(nSpts, 0.0);
      unsigned s = 0;
      for (unsigned i = 0; i < N; i++)  // hexas.cpp:84-101
        for (unsigned j = 0; j < N; j++)
          for (unsigned k = 0; k < N; k++, s++) {
            loc_spts[s * 3] = loc1d[k]; loc_spts[s * 3 + 1] = loc1d[j]; loc_spts[s * 3 + 2] = loc1d[i];
            idx_spts[s * 3] = k; idx_spts[s * 3 + 1] = j; idx_spts[s * 3 + 2] = i;
            weights_spts[s] = wts1d[i] * wts1d[j] * wts1d[k];
          }
      loc_fpts.assign(nFpts * 3, 0.0); idx_fpts.assign(nFpts * 3, 0);
      unsigned fp = 0;
      for (unsigned face = 0; face < 6; face++)  // hexas.cpp:133-194
        for (unsigned j = 0; j < N; j++)
          for (unsigned k = 0; k < N; k++, fp++) {
            real *l = &loc_fpts[fp * 3]; int *x = &idx_fpts[fp * 3];
            const int r = N - k - 1;
            switch (face) {
              case 0: l[0] = loc1d[k]; l[1] = loc1d[j]; l[2] = -1.0; x[0] = k; x[1] = j; x[2] = -1; break;
              case 1: l[0] = loc1d[r]; l[1] = loc1d[j]; l[2] = 1.0; x[0] = r; x[1] = j; x[2] = N; break;
              case 2: l[0] = -1.0; l[1] = loc1d[k]; l[2] = loc1d[j]; x[0] = -1; x[1] = k; x[2] = j; break;
              case 3: l[0] = 1.0; l[1] = loc1d[r]; l[2] = loc1d[j]; x[0] = N; x[1] = r; x[2] = j; break;
              case 4: l[0] = loc1d[r]; l[1] = -1.0; l[2] = loc1d[j]; x[0] = r; x[1] = -1; x[2] = j; break;
              case 5: l[0] = loc1d[k]; l[1] = 1.0; l[2] = loc1d[j]; x[0] = k; x[1] = N; x[2] = j; break;
            }
          }
      tnorm.assign(nFpts * 3, 0.0);  // hexas.cpp:252-292
      const real tn[6][3] = {{0, 0, -1}, {0, 0, 1}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}};
      for (unsigned f = 0; f < nFpts; f++)
        for (int d = 0; d < 3; d++) tnorm[f * 3 + d] = tn[f / (N * N)][d];
    }
  }

  // Linear Lagrange shape functions on the 2-point grid {-1, 1}, node order from
  // structured_to_gmsh_{quad,hex} (funcs.cpp:1007-1091): ijk -> gmsh = {0,1,3,2[,4,5,7,6]}.
  void shape_at(const real *loc, real *val, real *dval /* [nNodes][nDims] */) const {
    const std::vector<real> xl = {-1.0, 1.0};
    const int ijk2gmsh[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    real lag[3][2], dlag[3][2];
    for (int d = 0; d < nDims; d++)
      for (int i = 0; i < 2; i++) { lag[d][i] = lagrange(xl, i, loc[d]); dlag[d][i] = lagrange_d1(xl, i, loc[d]); }
    int pt = 0;
    if (nDims == 2) {  // quads.cpp:520-560
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++, pt++) {
          int g = ijk2gmsh[pt];
          val[g] = lag[0][i] * lag[1][j];
          dval[g * 2] = dlag[0][i] * lag[1][j];
          dval[g * 2 + 1] = lag[0][i] * dlag[1][j];
        }
    } else {  // hexas.cpp:760-840
      for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
          for (int i = 0; i < 2; i++, pt++) {
            int g = ijk2gmsh[pt];
            val[g] = lag[0][i] * lag[1][j] * lag[2][k];
            dval[g * 3] = dlag[0][i] * lag[1][j] * lag[2][k];
            dval[g * 3 + 1] = lag[0][i] * dlag[1][j] * lag[2][k];
            dval[g * 3 + 2] = lag[0][i] * lag[1][j] * dlag[2][k];
          }
    }
  }

  // elements.cpp:214-388 (shape, coords) for spts and fpts
  void set_shape_coords() {
    const int nN = nNodesPerEle, D = nDims;
    shape_spts.assign((size_t)nSpts * nN, 0.0); dshape_spts.assign((size_t)D * nSpts * nN, 0.0);
    shape_fpts.assign((size_t)nFpts * nN, 0.0); dshape_fpts.assign((size_t)D * nFpts * nN, 0.0);
    std::vector<real> v(nN), dv(nN * D);
    for (unsigned s = 0; s < nSpts; s++) {
      shape_at(&loc_spts[s * D], v.data(), dv.data());
      for (int n = 0; n < nN; n++) {
        shape_spts[s * nN + n] = v[n];
        for (int d = 0; d < D; d++) dshape_spts[((size_t)d * nSpts + s) * nN + n] = dv[n * D + d];
      }
    }
    for (unsigned f = 0; f < nFpts; f++) {
      shape_at(&loc_fpts[f * D], v.data(), dv.data());
      for (int n = 0; n < nN; n++) {
        shape_fpts[f * nN + n] = v[n];
        for (int d = 0; d < D; d++) dshape_fpts[((size_t)d * nFpts + f) * nN + n] = dv[n * D + d];
      }
    }
    nodes.assign((size_t)nN * D * nEles, 0.0);
    for (int d = 0; d < D; d++)
      for (int e = 0; e < nEles; e++)
        for (int n = 0; n < nN; n++)
          nodes[((size_t)n * D + d) * nEles + e] = coord_nodes[(size_t)ele2nodes[(size_t)e * nN + n] * D + d];
    const int ncol = nEles * D;
    coord_spts.assign((size_t)nSpts * ncol, 0.0);
    coord_fpts.assign((size_t)nFpts * ncol, 0.0);
    dgemm_dense(nSpts, ncol, nN, shape_spts.data(), nodes.data(), coord_spts.data());
    dgemm_dense(nFpts, ncol, nN, shape_fpts.data(), nodes.data(), coord_fpts.data());
  }

  // cblas_dgemm(RowMajor, NoTrans, NoTrans, alpha=1, beta=0) as used for geometry
  static void dgemm_dense(int M, int N, int K, const real *A, const real *B, real *C) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; i++)
      for (int j = 0; j < N; j++) {
        real s = 0.0;
        for (int k = 0; k < K; k++) s += A[(size_t)i * K + k] * B[(size_t)k * N + j];
        C[(size_t)i * N + j] = s;
      }
  }

  // ------------------------------------------------------------ orient_fpts (solver.cpp:312-437)
  void orient_fpts() {
    const int D = nDims;
    std::vector<real> cL((size_t)D * nGfpts, 0.0), cR((size_t)D * nGfpts, 0.0);
    std::vector<size_t> ptrL(nGfpts, 0);
    for (int e = 0; e < nEles; e++)
      for (unsigned f = 0; f < nFpts; f++) {
        int g = fpt2gfpt[(size_t)f * nEles + e];
        int s = fpt2slot[(size_t)f * nEles + e];
        for (int d = 0; d < D; d++) {
          real c = coord_fpts[((size_t)f * D + d) * nEles + e];
          if (s == 0) cL[(size_t)d * nGfpts + g] = c; else cR[(size_t)d * nGfpts + g] = c;
        }
        // address of coord_fpts(fpt, 0, ele) in units of doubles
        if (s == 0) ptrL[g] = (size_t)f * D * nEles + e;
      }
    std::vector<unsigned> idxL(nGfpts_int), idxR(nGfpts_int), idxsort(nGfpts_int);
    for (unsigned i = 0; i < nGfpts_int; i++) idxL[i] = idxR[i] = idxsort[i] = i;
    // each face's points are sorted on their own (disjoint slices), so the faces run in parallel
#pragma omp parallel for schedule(static)
    for (long ff = 0; ff < (long)nIntFaces; ff++) {
      const size_t shift = (size_t)ff * nFptsPerFace;
      fuzzysort_ind(cL, nGfpts, idxL.data() + shift, nFptsPerFace, D);
      fuzzysort_ind(cR, nGfpts, idxR.data() + shift, nFptsPerFace, D);
    }
    std::sort(idxsort.begin(), idxsort.end(), [&](unsigned a, unsigned b) { return ptrL[a] < ptrL[b]; });
    auto cpL = idxL, cpR = idxR;
    for (unsigned i = 0; i < nGfpts_int; i++) { idxL[i] = cpL[idxsort[i]]; idxR[i] = cpR[idxsort[i]]; }
    cpL = idxL; cpR = idxR;
    for (unsigned i = 0; i < nGfpts_int; i++) { idxL[cpL[i]] = i; idxR[cpR[i]] = i; }
    auto old = fpt2gfpt;
    for (int e = 0; e < nEles; e++)
      for (unsigned f = 0; f < nFpts; f++) {
        size_t k = (size_t)f * nEles + e;
        int g = old[k];
        if ((unsigned)g >= nGfpts_int) continue;
        fpt2gfpt[k] = (fpt2slot[k] == 0) ? idxL[g] : idxR[g];
      }
    auto f2f = face2fpts;
    for (unsigned ff = 0; ff < nFaces; ff++)
      for (unsigned j = 0; j < nFptsPerFace; j++) {
        int g = f2f[(size_t)j * nFaces + ff];
        if ((unsigned)g >= nGfpts_int) continue;
        int gn = idxL[g];
        face2fpts[(size_t)j * nFaces + ff] = gn;
        fpt2face[gn] = ff;
      }
  }

  // ------------------------------------------------------------ metrics (elements.cpp:845-996)
  void inverse_transforms(const std::vector<real> &jac, std::vector<real> &inv, std::vector<real> *det,
                          unsigned nPts) {
    const int D = nDims;
    auto J = [&](int a, unsigned p, int b, int e) { return jac[(((size_t)a * nPts + p) * D + b) * nEles + e]; };
    auto I = [&](int a, unsigned p, int b, int e) -> real & { return inv[(((size_t)a * nPts + p) * D + b) * nEles + e]; };
    bool negative = false;
#pragma omp parallel for schedule(static) reduction(|| : negative)
    for (int e = 0; e < nEles; e++)
      for (unsigned p = 0; p < nPts; p++) {
        real dt;
        if (D == 2) {
          dt = J(0, p, 0, e) * J(1, p, 1, e) - J(0, p, 1, e) * J(1, p, 0, e);
          I(0, p, 0, e) = J(1, p, 1, e);  I(0, p, 1, e) = -J(1, p, 0, e);
          I(1, p, 0, e) = -J(0, p, 1, e); I(1, p, 1, e) = J(0, p, 0, e);
        } else {
          real xr = J(0, p, 0, e), xs = J(1, p, 0, e), xt = J(2, p, 0, e);
          real yr = J(0, p, 1, e), ys = J(1, p, 1, e), yt = J(2, p, 1, e);
          real zr = J(0, p, 2, e), zs = J(1, p, 2, e), zt = J(2, p, 2, e);
          dt = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
          I(0, p, 0, e) = ys * zt - yt * zs; I(0, p, 1, e) = xt * zs - xs * zt; I(0, p, 2, e) = xs * yt - xt * ys;
          I(1, p, 0, e) = yt * zr - yr * zt; I(1, p, 1, e) = xr * zt - xt * zr; I(1, p, 2, e) = xt * yr - xr * yt;
          I(2, p, 0, e) = yr * zs - ys * zr; I(2, p, 1, e) = xs * zr - xr * zs; I(2, p, 2, e) = xr * ys - xs * yr;
        }
        if (det) {
          (*det)[(size_t)p * nEles + e] = dt;
          if (dt < 0) negative = true;
        }
      }
    if (negative) fail("oracle: negative Jacobian at solution points");
  }

  void calc_transforms() {
    const int D = nDims, nN = nNodesPerEle, ncol = nEles * D;
    jaco_spts.assign((size_t)D * nSpts * ncol, 0.0);
    jaco_fpts.assign((size_t)D * nFpts * ncol, 0.0);
    dgemm_dense(nSpts * D, ncol, nN, dshape_spts.data(), nodes.data(), jaco_spts.data());
    dgemm_dense(nFpts * D, ncol, nN, dshape_fpts.data(), nodes.data(), jaco_fpts.data());
    inv_jaco_spts.assign(jaco_spts.size(), 0.0);
    inv_jaco_fpts.assign(jaco_fpts.size(), 0.0);
    jaco_det_spts.assign((size_t)nSpts * nEles, 0.0);
    inverse_transforms(jaco_spts, inv_jaco_spts, &jaco_det_spts, nSpts);
    inverse_transforms(jaco_fpts, inv_jaco_fpts, nullptr, nFpts);
    norm.assign((size_t)D * nGfpts, 0.0);
    dA.assign((size_t)2 * nGfpts, 0.0);
    for (int e = 0; e < nEles; e++)
      for (unsigned f = 0; f < nFpts; f++) {
        int g = fpt2gfpt[(size_t)f * nEles + e];
        int s = fpt2slot[(size_t)f * nEles + e];
        real n[3] = {0, 0, 0};
        for (int d1 = 0; d1 < D; d1++)
          for (int d2 = 0; d2 < D; d2++)
            n[d1] += inv_jaco_fpts[(((size_t)d2 * nFpts + f) * D + d1) * nEles + e] * tnorm[f * D + d2];
        if (s == 0)
          for (int d = 0; d < D; d++) norm[(size_t)d * nGfpts + g] = n[d];
        real a = 0;
        for (int d = 0; d < D; d++) a += n[d] * n[d];
        dA[(size_t)s * nGfpts + g] = std::sqrt(a);
        if (s == 0)
          for (int d = 0; d < D; d++) norm[(size_t)d * nGfpts + g] /= dA[g];
      }
  }

  // ------------------------------------------------------------ FR operators (elements.cpp:555-654)
  real nodal_basis(unsigned spt, const real *loc) const {  // hexas.cpp:337-348, quads.cpp:239-248
    real v = 1.0;
    for (int d = 0; d < nDims; d++) v *= lagrange(loc1d, idx_spts[spt * nDims + d], loc[d]);
    return v;
  }
  // DFR derivative basis on [-1, GL, 1] (hexas.cpp:384-435 / 464-490, quads.cpp:280-361)
  real d_nodal_basis(const int *idx, const real *loc, int dim) const {
    real v = 1.0;
    for (int d = 0; d < nDims; d++) {
      unsigned m = idx[d] + 1;
      real t = (d == dim) ? lagrange_d1(locDFR, m, loc[d]) : lagrange(locDFR, m, loc[d]);
      v = (d == 0) ? t : v * t;
    }
    return v;
  }

  void setup_FR() {
    const int D = nDims;
    oppE.assign((size_t)nFpts * nSpts, 0.0);
    for (unsigned s = 0; s < nSpts; s++)
      for (unsigned f = 0; f < nFpts; f++) oppE[(size_t)f * nSpts + s] = nodal_basis(s, &loc_fpts[f * D]);
    oppD.assign((size_t)D * nSpts * nSpts, 0.0);
    oppDiv.assign((size_t)nSpts * D * nSpts, 0.0);
    for (int dim = 0; dim < D; dim++)
      for (unsigned j = 0; j < nSpts; j++)
        for (unsigned i = 0; i < nSpts; i++) {
          real v = d_nodal_basis(&idx_spts[j * D], &loc_spts[i * D], dim);
          oppD[((size_t)dim * nSpts + i) * nSpts + j] = v;
          oppDiv[((size_t)i * D + dim) * nSpts + j] = v;
        }
    oppD_fpts.assign((size_t)D * nSpts * nFpts, 0.0);
    for (int dim = 0; dim < D; dim++)
      for (unsigned f = 0; f < nFpts; f++)
        for (unsigned s = 0; s < nSpts; s++)
          oppD_fpts[((size_t)dim * nSpts + s) * nFpts + f] = d_nodal_basis(&idx_fpts[f * D], &loc_spts[s * D], dim);
    oppDiv_fpts.assign((size_t)nSpts * nFpts, 0.0);
    for (int dim = 0; dim < D; dim++)
      for (unsigned f = 0; f < nFpts; f++) {
        int fac = 1;
        if (D == 2) { int face = f / nSpts1D; if (face == 0 || face == 3) fac = -1; }
        else { int face = f / (nSpts1D * nSpts1D); if (face % 2 == 0) fac = -1; }
        for (unsigned s = 0; s < nSpts; s++)
          oppDiv_fpts[(size_t)s * nFpts + f] += fac * oppD_fpts[((size_t)dim * nSpts + s) * nFpts + f];
      }
  }

  // ------------------------------------------------------------ setup (solver.cpp:55-180)
  void setup(const std::string &meshfile) {
    nVars = nDims + 2;
    read_mesh(meshfile);
    setup_global_fpts();
    set_locs();
    set_shape_coords();
    orient_fpts();
    calc_transforms();
    setup_FR();
    gf_fpt.assign((size_t)2 * nGfpts, -1);
    gf_ele.assign((size_t)2 * nGfpts, -1);
    for (int e = 0; e < nEles; e++)
      for (unsigned f = 0; f < nFpts; f++) {
        int g = fpt2gfpt[(size_t)f * nEles + e], s = fpt2slot[(size_t)f * nEles + e];
        gf_fpt[(size_t)s * nGfpts + g] = f;
        gf_ele[(size_t)s * nGfpts + g] = e;
      }
    for (unsigned g = 0; g < nGfpts; g++)
      if (gf_ele[g] < 0 || gf_ele[nGfpts + g] < 0) fail("oracle: flux point without two sides");
    const size_t E = nEles, V = nVars, D = nDims;
    U_spts.assign(nSpts * V * E, 0.0);
    U_ini.assign(nSpts * V * E, 0.0);
    U_fpts.assign(nFpts * V * E, 0.0);
    Fcomm.assign(nFpts * V * E, 0.0);
    F_spts.assign(D * nSpts * V * E, 0.0);
    divF.assign((size_t)ph.nStages * nSpts * V * E, 0.0);
    if (ph.viscous) {
      Ucomm.assign(nFpts * V * E, 0.0);
      dU_spts.assign(D * nSpts * V * E, 0.0);
      dU_fpts.assign(D * nFpts * V * E, 0.0);
    }
  }

  // face views (solver.cpp:2402-2430): element-local address of (slot, var, gfpt)
  size_t fview(const std::vector<real> &, int slot, int var, unsigned g) const {
    return ((size_t)gf_fpt[(size_t)slot * nGfpts + g] * nVars + var) * nEles + gf_ele[(size_t)slot * nGfpts + g];
  }

  // ------------------------------------------------------------ residual pieces
  template <int V, int D>
  void compute_F() {  // elements.cpp:1588-1699
    const size_t E = nEles;
#pragma omp parallel for schedule(static)
    for (int spt = 0; spt < (int)nSpts; spt++)
      for (size_t e = 0; e < E; e++) {
        real U[V], tdU[V][D], F[V][D], dU[V][D];
        for (int v = 0; v < V; v++) {
          U[v] = U_spts[((size_t)spt * V + v) * E + e];
          if (ph.viscous)
            for (int d = 0; d < D; d++) tdU[v][d] = dU_spts[(((size_t)d * nSpts + spt) * V + v) * E + e];
        }
        real ij[D][D];
        for (int a = 0; a < D; a++)
          for (int b = 0; b < D; b++) ij[a][b] = inv_jaco_spts[(((size_t)a * nSpts + spt) * D + b) * E + e];
        for (int v = 0; v < V; v++)
          for (int d = 0; d < D; d++) dU[v][d] = 0.0;
        if (ph.viscous) {
          real inv_det = 1.0 / jaco_det_spts[(size_t)spt * E + e];
          for (int v = 0; v < V; v++)
            for (int d1 = 0; d1 < D; d1++) {
              for (int d2 = 0; d2 < D; d2++) dU[v][d1] += tdU[v][d2] * ij[d2][d1];
              dU[v][d1] *= inv_det;
              dU_spts[(((size_t)d1 * nSpts + spt) * V + v) * E + e] = dU[v][d1];
            }
        }
        real P;
        fconv<V, D>(U, F, P, ph.gamma);
        if (ph.viscous) fvisc_add<V, D>(U, dU, F, ph);
        for (int v = 0; v < V; v++)
          for (int d1 = 0; d1 < D; d1++) {
            real t = 0.0;
            for (int d2 = 0; d2 < D; d2++) t += F[v][d2] * ij[d1][d2];
            F_spts[(((size_t)d1 * nSpts + spt) * V + v) * E + e] = t;
          }
      }
  }

  template <int V>
  void common_U() {  // faces.cpp:1366-1474, LDG branch, interior flux points
    const real beta = 1.0 * ph.ldg_b;  // flip_beta == 1 on one rank (geometry.cpp:3364-3392)
#pragma omp parallel for schedule(static)
    for (int g = 0; g < (int)nGfpts; g++) {
      real UL[V] = {}, UR[V] = {};
      if (beta == 0.5)
        for (int n = 0; n < V; n++) UR[n] = U_fpts[fview(U_fpts, 1, n, g)];
      else if (beta == -0.5)
        for (int n = 0; n < V; n++) UL[n] = U_fpts[fview(U_fpts, 0, n, g)];
      else
        for (int n = 0; n < V; n++) { UL[n] = U_fpts[fview(U_fpts, 0, n, g)]; UR[n] = U_fpts[fview(U_fpts, 1, n, g)]; }
      for (int n = 0; n < V; n++) {
        real UC = 0.5 * (UL[n] + UR[n]) - beta * (UL[n] - UR[n]);
        Ucomm[fview(Ucomm, 0, n, g)] = UC;
        Ucomm[fview(Ucomm, 1, n, g)] = UC;
      }
    }
  }

  template <int V, int D>
  void rusanov() {  // faces.cpp:1244-1364, rus_bias == 0 (interior)
#pragma omp parallel for schedule(static)
    for (int g = 0; g < (int)nGfpts; g++) {
      real UL[V], UR[V], FL[V][D], FR[V][D], PL, PR;
      for (int n = 0; n < V; n++) { UL[n] = U_fpts[fview(U_fpts, 0, n, g)]; UR[n] = U_fpts[fview(U_fpts, 1, n, g)]; }
      if (ph.fconv == 2) {
        real nv[D], F[V];
        for (int d = 0; d < D; d++) nv[d] = norm[(size_t)d * nGfpts + g];
        roe_flux<V, D>(UL, UR, nv, ph.gamma, F);
        for (int n = 0; n < V; n++) {
          Fcomm[fview(Fcomm, 0, n, g)] = F[n] * dA[g];
          Fcomm[fview(Fcomm, 1, n, g)] = -F[n] * dA[nGfpts + g];
        }
        continue;
      }
      fconv<V, D>(UL, FL, PL, ph.gamma);
      fconv<V, D>(UR, FR, PR, ph.gamma);
      real aL = std::sqrt(ph.gamma * PL / UL[0]);
      real aR = std::sqrt(ph.gamma * PR / UR[0]);
      real VnL = 0.0, VnR = 0.0;
      for (int d = 0; d < D; d++) {
        VnL += UL[d + 1] / UL[0] * norm[(size_t)d * nGfpts + g];
        VnR += UR[d + 1] / UR[0] * norm[(size_t)d * nGfpts + g];
      }
      real eig = std::max(std::abs(VnL) + aL, std::abs(VnR) + aR);
      real FnL[V] = {0}, FnR[V] = {0};
      for (int n = 0; n < V; n++)
        for (int d = 0; d < D; d++) {
          FnL[n] += FL[n][d] * norm[(size_t)d * nGfpts + g];
          FnR[n] += FR[n][d] * norm[(size_t)d * nGfpts + g];
        }
      for (int n = 0; n < V; n++) {
        real F = (0.5 * (FnR[n] + FnL[n]) - 0.5 * eig * (1.0 - ph.rus_k) * (UR[n] - UL[n]));
        Fcomm[fview(Fcomm, 0, n, g)] = F * dA[g];
        Fcomm[fview(Fcomm, 1, n, g)] = -F * dA[nGfpts + g];
      }
    }
  }

  template <int V, int D>
  void ldg_flux() {  // faces.cpp:1476-1614, LDG_bias == 0 (interior)
    const real tau = ph.ldg_tau;
    const real beta = 1.0 * ph.ldg_b;
#pragma omp parallel for schedule(static)
    for (int g = 0; g < (int)nGfpts; g++) {
      real UL[V], UR[V], dUL[V][D] = {}, dUR[V][D] = {};
      for (int n = 0; n < V; n++) { UL[n] = U_fpts[fview(U_fpts, 0, n, g)]; UR[n] = U_fpts[fview(U_fpts, 1, n, g)]; }
      auto dUv = [&](int slot, int d, int n) {
        size_t f = gf_fpt[(size_t)slot * nGfpts + g], e = gf_ele[(size_t)slot * nGfpts + g];
        return dU_fpts[(((size_t)d * nFpts + f) * V + n) * nEles + e];
      };
      for (int d = 0; d < D; d++)
        for (int n = 0; n < V; n++) {
          if (beta == 0.5) dUL[n][d] = dUv(0, d, n);
          else if (beta == -0.5) dUR[n][d] = dUv(1, d, n);
          else { dUL[n][d] = dUv(0, d, n); dUR[n][d] = dUv(1, d, n); }
        }
      real FL[V][D] = {}, FR[V][D] = {};
      fvisc_add<V, D>(UL, dUL, FL, ph);
      fvisc_add<V, D>(UR, dUR, FR, ph);
      real FnL[V] = {0}, FnR[V] = {0};
      for (int n = 0; n < V; n++)
        for (int d = 0; d < D; d++) {
          FnL[n] += FL[n][d] * norm[(size_t)d * nGfpts + g];
          FnR[n] += FR[n][d] * norm[(size_t)d * nGfpts + g];
        }
      for (int n = 0; n < V; n++) {
        real Fc = 0.5 * (FnL[n] + FnR[n]) + tau * (UL[n] - UR[n]) + beta * (FnL[n] - FnR[n]);
        Fcomm[fview(Fcomm, 0, n, g)] += Fc * dA[g];
        Fcomm[fview(Fcomm, 1, n, g)] -= Fc * dA[nGfpts + g];
      }
    }
  }

  template <int V, int D>
  void residual_t(int stage) {  // solver.cpp:1494-1645
    const int E = nEles, n = E * V;
    op_mm(nFpts, n, nSpts, 1.0, oppE.data(), U_spts.data(), 0.0, U_fpts.data());   // extrapolate_U
    if (ph.viscous) {
      op_mm(nSpts * D, n, nSpts, 1.0, oppD.data(), U_spts.data(), 0.0, dU_spts.data());  // compute_dU_spts
      common_U<V>();
      op_mm(nSpts * D, n, nFpts, 1.0, oppD_fpts.data(), Ucomm.data(), 1.0, dU_spts.data());  // compute_dU_fpts
    }
    compute_F<V, D>();
    if (ph.viscous)
      for (int d = 0; d < D; d++)  // extrapolate_dU
        op_mm(nFpts, n, nSpts, 1.0, oppE.data(), dU_spts.data() + (size_t)d * nSpts * n, 0.0,
              dU_fpts.data() + (size_t)d * nFpts * n);
    real *dv = divF.data() + (size_t)stage * nSpts * n;
    op_mm(nSpts, n, nSpts * D, 1.0, oppDiv.data(), F_spts.data(), 0.0, dv);  // compute_divF_spts
    rusanov<V, D>();
    if (ph.viscous) ldg_flux<V, D>();
    op_mm(nSpts, n, nFpts, 1.0, oppDiv_fpts.data(), Fcomm.data(), 1.0, dv);  // compute_divF_fpts
  }

  void compute_residual(int stage) {
    if (stage < 0 || stage >= ph.nStages) fail("oracle: stage out of range");
    if (nDims == 2) residual_t<4, 2>(stage); else residual_t<5, 3>(stage);
  }

  // step_RK (solver.cpp:2865-3097) with dt_type 0; src (reference layout [nS][V][E]) is
  // the p-multigrid source term of the sourceBT branches (solver.cpp:2946-2962, 3061-3073)
  void step_rk(const real *src = nullptr) {
    const size_t E = nEles, V = nVars, n = V * E;
    const real dt = ph.dt;
    if (ph.nStages > 1) U_ini = U_spts;
    for (int stage = 0; stage < ph.nStages - 1; stage++) {
      compute_residual(stage);
      const real *dv = divF.data() + (size_t)stage * nSpts * n;
#pragma omp parallel for schedule(static)
      for (int s = 0; s < (int)nSpts; s++)
        for (size_t v = 0; v < V; v++)
          for (size_t e = 0; e < E; e++) {
            size_t k = ((size_t)s * V + v) * E + e;
            if (src)
              U_spts[k] = U_ini[k] - rk_alpha[stage] * dt / jaco_det_spts[(size_t)s * E + e] * (dv[k] + src[k]);
            else
              U_spts[k] = U_ini[k] - rk_alpha[stage] * dt / jaco_det_spts[(size_t)s * E + e] * dv[k];
          }
    }
    compute_residual(ph.nStages - 1);
    if (ph.nStages > 1) U_spts = U_ini;
    const real *beta = (ph.nStages > 1) ? rk_beta : nullptr;
    for (int stage = 0; stage < ph.nStages; stage++) {
      const real b = beta ? beta[stage] : 1.0;
      const real *dv = divF.data() + (size_t)stage * nSpts * n;
#pragma omp parallel for schedule(static)
      for (int s = 0; s < (int)nSpts; s++)
        for (size_t v = 0; v < V; v++)
          for (size_t e = 0; e < E; e++) {
            size_t k = ((size_t)s * V + v) * E + e;
            if (src)
              U_spts[k] -= b * dt / jaco_det_spts[(size_t)s * E + e] * (dv[k] + src[k]);
            else
              U_spts[k] -= b * dt / jaco_det_spts[(size_t)s * E + e] * dv[k];
          }
    }
  }

  // report_residuals (solver.cpp:6990-7137): per-variable norm of divF(0)/detJ, / nDoF
  void report_residuals(real *res) const {
    const size_t E = nEles, V = nVars;
    std::vector<real> r(V, 0.0);
    for (size_t s = 0; s < nSpts; s++)
      for (size_t v = 0; v < V; v++)
        for (size_t e = 0; e < E; e++) {
          real x = divF[(s * V + v) * E + e], j = jaco_det_spts[s * E + e];
          if (ph.res_type == 0) r[v] = std::max(r[v], std::abs(x / j));
          else if (ph.res_type == 1) r[v] += std::abs(x / j);
          else r[v] += x * x / (j * j);
          if (std::isnan(r[v])) fail("oracle: NaN in residual");
        }
    const real nDoF = (real)(nSpts * E);
    for (size_t v = 0; v < V; v++) {
      real x = (ph.res_type == 2) ? std::sqrt(r[v]) : r[v];
      res[v] = x / nDoF;
    }
  }
};

// ---------------------------------------------------------------- p-multigrid
// PMGrid (multigrid.cpp:18-518) with the dense tensor operators of Quads/Hexas::setup_PMG
// (quads.cpp:363-465, hexas.cpp:586-704): oppRes/oppPro [nSpts_to][nSpts_from], built as
// the reference builds them (single-order Lagrange steps between Gauss-Legendre point sets,
// chained by dense matrix products in the reference's association) and applied as the
// reference's cblas_dgemm over the [nSpts][nVars*nEles] arrays.
struct OraclePMG {
  Oracle *fine = nullptr;
  std::vector<std::unique_ptr<Oracle>> grids;  // grids[0] empty: level 0 is `fine`
  std::vector<int> levels, steps;
  real rel_fac = 1.0;
  std::string cycle_type = "V";
  int fmg_vcycles = 1;
  std::vector<std::vector<real>> oppRes, oppPro;           // per level
  std::vector<std::vector<real>> sources, solutions, corrections;  // [nS][V][E] per level

  int nLevels() const { return (int)levels.size(); }
  Oracle &level(int n) { return n == 0 ? *fine : *grids[n]; }

  // dense single-order step between orders: rows = points of order `to`, columns = order `from`
  static std::vector<real> step_op(int D, int from, int to) {
    const unsigned nf = from + 1, nt = to + 1;
    const auto xf = gl_points(nf), xt = gl_points(nt);
    const unsigned Nf = D == 2 ? nf * nf : nf * nf * nf, Nt = D == 2 ? nt * nt : nt * nt * nt;
    std::vector<real> A((size_t)Nt * Nf);
    for (unsigned spt = 0; spt < Nf; spt++) {
      const unsigned i1 = spt % nf, j1 = (spt / nf) % nf, k1 = spt / (nf * nf);
      for (unsigned r = 0; r < Nt; r++) {
        const unsigned i2 = r % nt, j2 = (r / nt) % nt, k2 = r / (nt * nt);
        real v = lagrange(xf, i1, xt[i2]) * lagrange(xf, j1, xt[j2]);
        if (D == 3) v *= lagrange(xf, k1, xt[k2]);
        A[(size_t)r * Nf + spt] = v;
      }
    }
    return A;
  }
  static std::vector<real> dense_mm(const std::vector<real> &A, const std::vector<real> &B, int m, int k, int n) {
    std::vector<real> C((size_t)m * n, 0.0);
    for (int i = 0; i < m; i++)
      for (int j = 0; j < n; j++) {
        real s = 0.0;
        for (int l = 0; l < k; l++) s += A[(size_t)i * k + l] * B[(size_t)l * n + j];
        C[(size_t)i * n + j] = s;
      }
    return C;
  }
  static int npts(int D, int p) { return D == 2 ? (p + 1) * (p + 1) : (p + 1) * (p + 1) * (p + 1); }
  // hexas.cpp:646-697: ((opps[res] opps[res+1]) opps[res+2]) ... opps[order-1]
  static std::vector<real> restriction(int D, int order, int res) {
    std::vector<real> R = step_op(D, res + 1, res);
    for (int P = res + 1; P < order; P++) R = dense_mm(R, step_op(D, P + 1, P), npts(D, res), npts(D, P), npts(D, P + 1));
    return R;
  }
  // hexas.cpp:594-644: ((opps[pro] opps[pro-1]) ...) opps[order+1]
  static std::vector<real> prolongation(int D, int order, int pro) {
    std::vector<real> Pm = step_op(D, pro - 1, pro);
    for (int P = pro - 1; P > order; P--) Pm = dense_mm(Pm, step_op(D, P - 1, P), npts(D, pro), npts(D, P), npts(D, P - 1));
    return Pm;
  }

  // multigrid.cpp:18-103
  void setup(Oracle *f, const std::string &meshfile) {
    fine = f;
    const int D = f->nDims, L = nLevels();
    if (L < 2 || levels[0] != f->order) fail("oracle: mg_levels must start at the fine order with >= 2 levels");
    grids.clear();
    grids.emplace_back();
    oppRes.assign(L, {});
    oppPro.assign(L, {});
    sources.assign(L, {});
    solutions.assign(L, {});
    corrections.assign(L, {});
    oppRes[0] = restriction(D, levels[0], levels[1]);
    for (int n = 1; n < L; n++) {
      auto g = std::make_unique<Oracle>();
      g->nDims = D;
      g->order = levels[n];
      g->ph = f->ph;
      g->setup(meshfile);
      oppPro[n] = prolongation(D, levels[n], levels[n - 1]);
      if (n + 1 < L) oppRes[n] = restriction(D, levels[n], levels[n + 1]);
      const size_t sz = g->U_spts.size();
      sources[n].assign(sz, 0.0);
      solutions[n].assign(sz, 0.0);
      corrections[n].assign(sz, 0.0);
      grids.push_back(std::move(g));
    }
  }

  // C[nSc][V*E] = alpha A[nSc][nSf] B[nSf][V*E] + beta C
  static void gemm(const std::vector<real> &A, int m, int k, const real *B, int n, real alpha, real beta,
                   real *C) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) {
        real s = 0.0;
        for (int l = 0; l < k; l++) s += A[(size_t)i * k + l] * B[(size_t)l * n + j];
        C[(size_t)i * n + j] = (beta == 0.0) ? alpha * s : alpha * s + beta * C[(size_t)i * n + j];
      }
  }

  // multigrid.cpp:335-392
  void restrict_pmg(int nf) {
    Oracle &f = level(nf), &c = level(nf + 1);
    const int n = f.nVars * f.nEles;
    gemm(oppRes[nf], c.nSpts, f.nSpts, f.U_spts.data(), n, 1.0, 0.0, c.U_spts.data());
    gemm(oppRes[nf], c.nSpts, f.nSpts, f.divF.data(), n, 1.0, 0.0, c.divF.data());
  }
  // multigrid.cpp:394-413
  void prolong_err(int nc, const std::vector<real> &corr) {
    Oracle &c = level(nc), &f = level(nc - 1);
    gemm(oppPro[nc], f.nSpts, c.nSpts, corr.data(), c.nVars * c.nEles, rel_fac, 1.0, f.U_spts.data());
  }
  // multigrid.cpp:435-457
  void prolong_U(int nc) {
    Oracle &c = level(nc), &f = level(nc - 1);
    gemm(oppPro[nc], f.nSpts, c.nSpts, c.U_spts.data(), c.nVars * c.nEles, 1.0, 0.0, f.U_spts.data());
  }
  // multigrid.cpp:474-502
  void compute_source_term(int n) {
    Oracle &g = level(n);
    const size_t sz = g.U_spts.size();
    for (size_t i = 0; i < sz; i++) sources[n][i] = g.divF[i];
    g.compute_residual(0);
    for (size_t i = 0; i < sz; i++) sources[n][i] -= g.divF[i];
  }
  // multigrid.cpp:105-261
  void v_cycle(int lvl) {
    const int L = nLevels();
    if (lvl != L - 1) {
      level(lvl).compute_residual(0);
      restrict_pmg(lvl);
    }
    for (int n = lvl + 1; n < L; n++) {
      compute_source_term(n);
      solutions[n] = grids[n]->U_spts;
      for (int s = 0; s < steps[n]; s++) grids[n]->step_rk(sources[n].data());
      if (n + 1 < L) {
        grids[n]->compute_residual(0);
        for (size_t i = 0; i < sources[n].size(); i++) grids[n]->divF[i] += sources[n][i];
        restrict_pmg(n);
      }
    }
    for (int n = L - 1; n > lvl; n--) {
      if (n != L - 1)
        for (int s = 0; s < steps[n]; s++) grids[n]->step_rk(sources[n].data());
      for (size_t i = 0; i < corrections[n].size(); i++) corrections[n][i] = grids[n]->U_spts[i] - solutions[n][i];
      if (n > lvl + 1) prolong_err(n, corrections[n]);
    }
    if (lvl != L - 1) prolong_err(lvl + 1, corrections[lvl + 1]);
  }
  // multigrid.cpp:263-333
  void cycle() {
    if (cycle_type == "V") {
      for (int s = 0; s < steps[0]; s++) fine->step_rk();
      v_cycle(0);
    } else if (cycle_type == "FMG") {
      for (int n = nLevels() - 1; n > 0; n--) {
        for (int c = 0; c < fmg_vcycles; c++) {
          grids[n]->step_rk();
          v_cycle(n);
        }
        grids[n]->step_rk();
        if (n > 1) prolong_U(n);
      }
      prolong_U(1);
      cycle_type = "V";
      cycle();
    } else {
      fail("oracle: Multigrid cycle type unknown!");
    }
  }
};

thread_local std::string g_err;

template <typename Fn>
int guarded(Fn &&fn) {
  try { fn(); return 0; }
  catch (const std::exception &ex) { g_err = ex.what(); return 1; }
}

}  // namespace

// ---------------------------------------------------------------- C API (ctypes)
extern "C" {

struct zo_phys {
  int viscous;
  double gamma, prandtl, mu, rt, c_sth;
  int fix_vis;
  double rus_k, ldg_b, ldg_tau, dt;
  int nStages;
  int res_type;
  int fconv;
};

const char *zo_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {
Phys phys_from(const zo_phys *p) {
  Phys ph;
  ph.viscous = p->viscous; ph.gamma = p->gamma; ph.prandtl = p->prandtl; ph.mu = p->mu; ph.rt = p->rt;
  ph.c_sth = p->c_sth; ph.fix_vis = p->fix_vis; ph.rus_k = p->rus_k; ph.ldg_b = p->ldg_b;
  ph.ldg_tau = p->ldg_tau; ph.dt = p->dt; ph.nStages = p->nStages; ph.res_type = p->res_type; ph.fconv = p->fconv;
  return ph;
}
}  // namespace

extern "C" {

void *zo_create(const char *meshfile, int nDims, int order, const zo_phys *p) {
  Oracle *o = new Oracle();
  int rc = guarded([&] {
    if (nDims != 2 && nDims != 3) fail("oracle: nDims must be 2 or 3");
    if (order < 0 || order > 9) fail("oracle: order must be in 0..9");
    if (p->nStages != 1 && p->nStages != 4) fail("oracle: nStages must be 1 or 4");
    if (p->fconv != 0 && p->fconv != 2) fail("oracle: fconv must be 0 (Rusanov) or 2 (Roe)");
    o->nDims = nDims; o->order = order;
    o->ph = phys_from(p);
    o->setup(meshfile);
  });
  if (rc) { delete o; return nullptr; }
  return o;
}

void zo_destroy(void *h) { delete static_cast<Oracle *>(h); }

// out: nEles nSpts nFpts nVars nDims nGfpts nGfpts_int nStages nFaces nIntFaces
int zo_dims(void *h, int *out) {
  auto *o = static_cast<Oracle *>(h);
  int v[10] = {o->nEles, (int)o->nSpts, (int)o->nFpts, o->nVars, o->nDims, (int)o->nGfpts,
               (int)o->nGfpts_int, o->ph.nStages, (int)o->nFaces, (int)o->nIntFaces};
  std::memcpy(out, v, sizeof(v));
  return 0;
}

int zo_get_maps(void *h, int *fpt2gfpt, signed char *slot) {
  auto *o = static_cast<Oracle *>(h);
  std::memcpy(fpt2gfpt, o->fpt2gfpt.data(), o->fpt2gfpt.size() * sizeof(int));
  std::memcpy(slot, o->fpt2slot.data(), o->fpt2slot.size());
  return 0;
}

// name: one of the double arrays below; n: capacity in doubles; returns count or -1
long zo_get_array(void *h, const char *name, double *out, long n) {
  auto *o = static_cast<Oracle *>(h);
  const std::map<std::string, const std::vector<real> *> arrs = {
      {"oppE", &o->oppE}, {"oppD", &o->oppD}, {"oppD_fpts", &o->oppD_fpts}, {"oppDiv", &o->oppDiv},
      {"oppDiv_fpts", &o->oppDiv_fpts}, {"coord_spts", &o->coord_spts}, {"coord_fpts", &o->coord_fpts},
      {"jaco_det_spts", &o->jaco_det_spts}, {"inv_jaco_spts", &o->inv_jaco_spts},
      {"inv_jaco_fpts", &o->inv_jaco_fpts}, {"norm", &o->norm}, {"dA", &o->dA},
      {"weights_spts", &o->weights_spts}, {"loc_spts", &o->loc_spts}, {"loc_fpts", &o->loc_fpts},
      {"U_spts", &o->U_spts}, {"U_fpts", &o->U_fpts}, {"Ucomm", &o->Ucomm}, {"F_spts", &o->F_spts},
      {"Fcomm", &o->Fcomm}, {"dU_spts", &o->dU_spts}, {"dU_fpts", &o->dU_fpts}, {"divF", &o->divF},
      {"coord_nodes", &o->coord_nodes}};
  auto it = arrs.find(name);
  if (it == arrs.end()) { g_err = std::string("oracle: unknown array ") + name; return -1; }
  const auto &a = *it->second;
  if ((long)a.size() > n) { g_err = "oracle: output buffer too small"; return -1; }
  std::copy(a.begin(), a.end(), out);
  return (long)a.size();
}

int zo_set_U(void *h, const double *U) {
  auto *o = static_cast<Oracle *>(h);
  std::copy(U, U + o->U_spts.size(), o->U_spts.begin());
  return 0;
}

int zo_compute_residual(void *h, int stage) {
  return guarded([&] { static_cast<Oracle *>(h)->compute_residual(stage); });
}

int zo_step_rk(void *h, int nsteps) {
  return guarded([&] {
    for (int i = 0; i < nsteps; i++) static_cast<Oracle *>(h)->step_rk();
  });
}

int zo_report_residuals(void *h, double *res) {
  return guarded([&] {
    auto *o = static_cast<Oracle *>(h);
    std::vector<real> r(o->nVars);
    o->report_residuals(r.data());
    std::copy(r.begin(), r.end(), res);
  });
}

// Stand-alone pieces for closed-form checks of the restatement.
int zo_gl_points(int n, double *pts, double *wts) {
  return guarded([&] {
    auto p = gl_points(n), w = gl_weights(n);
    std::copy(p.begin(), p.end(), pts);
    std::copy(w.begin(), w.end(), wts);
  });
}

void *zo_pmg_create(void *h, const char *meshfile, int nLevels, const int *levels, const int *steps, double rel_fac,
                    const char *cycle_type, int fmg_vcycles) {
  OraclePMG *p = new OraclePMG();
  int rc = guarded([&] {
    p->levels.assign(levels, levels + nLevels);
    p->steps.assign(steps, steps + nLevels);
    for (int n = 1; n < nLevels; n++)
      if (levels[n] >= levels[n - 1] || levels[n] < 0) fail("oracle: mg_levels must decrease strictly");
    p->rel_fac = rel_fac;
    p->cycle_type = cycle_type;
    p->fmg_vcycles = fmg_vcycles;
    p->setup(static_cast<Oracle *>(h), meshfile);
  });
  if (rc) { delete p; return nullptr; }
  return p;
}

void zo_pmg_destroy(void *p) { delete static_cast<OraclePMG *>(p); }

int zo_pmg_cycle(void *p, int ncycles) {
  return guarded([&] {
    for (int i = 0; i < ncycles; i++) static_cast<OraclePMG *>(p)->cycle();
  });
}

// which: 0 U_spts, 2 divF(0), 3 source, 4 solution, 5 correction; reference layout [nS][V][E]
long zo_pmg_array(void *h, int level, int which, double *out, long n, int set) {
  long count = -1;
  int rc = guarded([&] {
    auto *p = static_cast<OraclePMG *>(h);
    if (level < 1 || level >= p->nLevels()) fail("oracle: pmg level out of range");
    Oracle &g = *p->grids[level];
    std::vector<real> *v = nullptr;
    switch (which) {
      case 0: v = &g.U_spts; break;
      case 2: v = &g.divF; break;
      case 3: v = &p->sources[level]; break;
      case 4: v = &p->solutions[level]; break;
      case 5: v = &p->corrections[level]; break;
      default: fail("oracle: which must be 0, 2, 3, 4 or 5");
    }
    const long sz = (long)g.U_spts.size();
    count = sz;
    if (!out) return;
    if (n < sz) fail("oracle: buffer too small");
    if (set) std::copy(out, out + sz, v->begin());
    else std::copy(v->begin(), v->begin() + sz, out);
  });
  return rc ? -1 : count;
}

// the dense PMG operators: kind 0 = restriction of level `level`, 1 = prolongation
long zo_pmg_operator(void *h, int level, int kind, double *out, long n) {
  long count = -1;
  int rc = guarded([&] {
    auto *p = static_cast<OraclePMG *>(h);
    if (level < 0 || level >= p->nLevels()) fail("oracle: pmg level out of range");
    const auto &A = kind == 0 ? p->oppRes[level] : p->oppPro[level];
    count = (long)A.size();
    if (out) {
      if (n < count) fail("oracle: buffer too small");
      std::copy(A.begin(), A.end(), out);
    }
  });
  return rc ? -1 : count;
}

double zo_lagrange(const double *grid, int n, int mode, double x) {
  return (double)lagrange(std::vector<real>(grid, grid + n), mode, x);
}

double zo_lagrange_d1(const double *grid, int n, int mode, double x) {
  return (double)lagrange_d1(std::vector<real>(grid, grid + n), mode, x);
}

// Pointwise fluxes on one state (nDims 2 or 3): F[V][D] convective (+ viscous if dU given).
// Convective common normal flux for the unit normal n (no area): Rusanov or Roe.
int zo_common_flux(int nDims, const zo_phys *p, const double *UL, const double *UR, const double *n, double *F) {
  return guarded([&] {
    const Phys ph = phys_from(p);
    if (ph.fconv != 2) fail("oracle: zo_common_flux is the Roe flux (fconv 2)");
    const int V = nDims + 2;
    real ul[5], ur[5], nn[3], f[5];
    std::copy(UL, UL + V, ul);
    std::copy(UR, UR + V, ur);
    std::copy(n, n + nDims, nn);
    if (nDims == 2) roe_flux<4, 2>(ul, ur, nn, ph.gamma, f);
    else roe_flux<5, 3>(ul, ur, nn, ph.gamma, f);
    std::copy(f, f + V, F);
  });
}

int zo_point_flux(int nDims, const zo_phys *p, const double *U, const double *dU, double *F) {
  return guarded([&] {
    const Phys ph = phys_from(p);
    real u[5];
    std::copy(U, U + nDims + 2, u);
    if (nDims == 2) {
      real f[4][2], P, du[4][2];
      fconv<4, 2>(u, f, P, ph.gamma);
      if (dU) { std::copy(dU, dU + 8, &du[0][0]); fvisc_add<4, 2>(u, du, f, ph); }
      std::copy(&f[0][0], &f[0][0] + 8, F);
    } else {
      real f[5][3], P, du[5][3];
      fconv<5, 3>(u, f, P, ph.gamma);
      if (dU) { std::copy(dU, dU + 15, &du[0][0]); fvisc_add<5, 3>(u, du, f, ph); }
      std::copy(&f[0][0], &f[0][0] + 15, F);
    }
  });
}

}  // extern "C"
