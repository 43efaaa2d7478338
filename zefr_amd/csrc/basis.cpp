// basis.cpp -- Gauss-Legendre points, Lagrange bases and the reference element tables.
//
// Point sets and orderings are those of the reference (src/points.cpp:9-209,
// src/quads.cpp:45-190, src/hexas.cpp:51-292) because the solution layout and the flux
// point numbering are part of the interface.  The DFR operators are not stored as dense
// matrices: on tensor-product elements every one of them is a 1D operator applied along
// element lines, so the device only receives
//   lE[side][m]  = l_m(-1 | +1)                      (oppE, src/elements.cpp:565-575)
//   Dss[i][m]    = dL^DFR_{m+1}/dx (x_i)              (oppD, oppDiv, :577-622)
//   Dsf[i][side] = dL^DFR_{0 | N+1}/dx (x_i)          (oppD_fpts, oppDiv_fpts, :594-654)
// where L^DFR are the Lagrange polynomials on [-1, x_1..x_N, 1].  dense_operator()
// expands these tables back into the reference's dense operators for checking.
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"

namespace zfr {

// A symmetric rule of n values from its half at the -1 end, listed inwards (for odd n the
// last entry is the middle value): magnitudes, with 'sign' on the -1 side (points -1,
// weights +1).
static std::vector<double> mirror(const std::vector<double> &half, unsigned n, double sign) {
  std::vector<double> r(n);
  for (unsigned i = 0; i < n / 2; i++) {
    r[i] = sign * half[i];
    r[n - 1 - i] = half[i];
  }
  if (n % 2) r[n / 2] = half[n / 2];
  return r;
}

// src/points.cpp:9-98 (closed forms up to 5 points; for 6..10 points the tables are the
// reference's 15-digit ones, needed to reproduce its operators at p = 5..9: the 7-point
// outer node ends in ...758, one unit below the correctly rounded value).
std::vector<double> gauss_legendre_points(unsigned n) {
  switch (n) {
    case 1: return {0.0};
    case 2: { const double a = std::sqrt(1. / 3.); return {-a, a}; }
    case 3: { const double a = std::sqrt(3. / 5.); return {-a, 0.0, a}; }
    case 4: {
      const double o = std::sqrt((3. + 2. * std::sqrt(6. / 5.)) / 7.);
      const double i = std::sqrt((3. - 2. * std::sqrt(6. / 5.)) / 7.);
      return {-o, -i, i, o};
    }
    case 5: {
      const double o = (1. / 3.) * std::sqrt(5. + 2. * std::sqrt(10. / 7.));
      const double i = (1. / 3.) * std::sqrt(5. - 2. * std::sqrt(10. / 7.));
      return {-o, -i, 0.0, i, o};
    }
    case 6: return mirror({0.932469514203152, 0.661209386466264, 0.238619186083197}, n, -1.0);
    case 7: return mirror({0.949107912342758, 0.741531185599394, 0.405845151377397, 0.0}, n, -1.0);
    case 8: return mirror({0.960289856497536, 0.796666477413627, 0.525532409916329, 0.183434642495650}, n, -1.0);
    case 9:
      return mirror({0.968160239507626, 0.836031107326636, 0.613371432700590, 0.324253423403809, 0.0}, n, -1.0);
    case 10:
      return mirror({0.973906528517172, 0.865063366688985, 0.679409568299024, 0.433895394129247,
                        0.148874338981631}, n, -1.0);
    default: raise("Gauss-Legendre points are available for 1..10 points (order 0..9)");
  }
}

// src/points.cpp:123-209
std::vector<double> gauss_legendre_weights(unsigned n) {
  switch (n) {
    case 1: return {2.0};
    case 2: return {1.0, 1.0};
    case 3: return {5. / 9., 8. / 9., 5. / 9.};
    case 4: {
      const double a = 1. / 36. * (18. - std::sqrt(30.)), b = 1. / 36. * (18. + std::sqrt(30.));
      return {a, b, b, a};
    }
    case 5: {
      const double a = 1. / 900. * (322 - 13. * std::sqrt(70.)), b = 1. / 900. * (322 + 13. * std::sqrt(70.));
      return {a, b, 128. / 225., b, a};
    }
    case 6: return mirror({0.171324492379171, 0.360761573048139, 0.467913934572691}, n, 1.0);
    case 7: return mirror({0.129484966168870, 0.279705391489277, 0.381830050505119, 0.417959183673469}, n, 1.0);
    case 8: return mirror({0.101228536290377, 0.222381034453374, 0.313706645877887, 0.362683783378362}, n, 1.0);
    case 9:
      return mirror({0.081274388361575, 0.180648160694857, 0.260610696402936, 0.312347077040003,
                        0.330239355001260}, n, 1.0);
    case 10:
      return mirror({0.066671344308688, 0.149451349150581, 0.219086362515982, 0.269266719309996,
                        0.295524224714753}, n, 1.0);
    default: raise("Gauss-Legendre weights are available for 1..10 points");
  }
}

// Product form of the Lagrange basis (include/polynomials.hpp:15-28).
double lagrange(const std::vector<double> &g, unsigned mode, double x) {
  double v = 1.0;
  for (unsigned i = 0; i < g.size(); i++)
    if (i != mode) v *= (x - g[i]) / (g[mode] - g[i]);
  return v;
}

// First derivative as a sum of products over the dropped factor
// (include/polynomials.hpp:47-76).
double lagrange_d1(const std::vector<double> &g, unsigned mode, double x) {
  double den = 1.0;
  for (unsigned i = 0; i < g.size(); i++)
    if (i != mode) den *= (g[mode] - g[i]);
  double sum = 0.0;
  for (unsigned j = 0; j < g.size(); j++) {
    if (j == mode) continue;
    double t = 1.0;
    for (unsigned i = 0; i < g.size(); i++)
      if (i != mode && i != j) t *= (x - g[i]);
    sum += t;
  }
  return sum / den;
}

void RefElement::build(int nd, int p) {
  if (nd != 2 && nd != 3) raise("nDims must be 2 or 3");
  // quads to order 9 (the reference's Gauss-Legendre tables end at 10 points); hexes to
  // order 8: the kernels run one thread per solution point, and at order 9 the Euler
  // divergence's LDS per element (180 KiB) exceeds what a workgroup can allocate (at
  // order 8 the viscous adjoint's, 199 KiB, does: its launch refuses)
  if (p < 0 || p > max_order(nd)) raise("order must be in 0.." + std::to_string(max_order(nd)) + (nd == 2 ? " for quads" : " for hexes"));
  nDims = nd; order = p; N = p + 1;
  loc1d = gauss_legendre_points(N);
  wts1d = gauss_legendre_weights(N);
  locDFR = loc1d;
  locDFR.insert(locDFR.begin(), -1.0);
  locDFR.push_back(1.0);
  const int D = nd;
  if (D == 2) {
    nSpts = N * N; nFptsPerFace = N; nFaces = 4; nNodes = 4;
  } else {
    nSpts = N * N * N; nFptsPerFace = N * N; nFaces = 6; nNodes = 8;
  }
  nFpts = nFaces * nFptsPerFace;
  if (nFpts > 512) raise("flux points per element must fit the 9-bit partner index");
  loc_spts.assign((size_t)nSpts * D, 0.0); idx_spts.assign((size_t)nSpts * D, 0);
  weights_spts.assign(nSpts, 0.0);
  // solution points: x index fastest (src/quads.cpp:72-84, src/hexas.cpp:84-101)
  for (int s = 0; s < nSpts; s++) {
    int r = s;
    double w = 1.0;
    for (int d = 0; d < D; d++) {
      const int i = r % N;
      r /= N;
      loc_spts[s * D + d] = loc1d[i];
      idx_spts[s * D + d] = i;
    }
    // weight product in the reference's order (outer index first)
    if (D == 2) w = wts1d[idx_spts[s * D + 1]] * wts1d[idx_spts[s * D]];
    else w = wts1d[idx_spts[s * D + 2]] * wts1d[idx_spts[s * D + 1]] * wts1d[idx_spts[s * D]];
    weights_spts[s] = w;
  }
  // flux points face by face (src/quads.cpp:86-123, src/hexas.cpp:133-194)
  loc_fpts.assign((size_t)nFpts * D, 0.0); idx_fpts.assign((size_t)nFpts * D, 0);
  tnorm.assign((size_t)nFpts * D, 0.0);
  for (int f = 0; f < nFpts; f++) {
    int ix[3] = {0, 0, 0};
    int dim = 0, side = 0;
    if (D == 2) {
      const int face = f / N, j = f % N;
      switch (face) {
        case 0: ix[0] = j; ix[1] = -1; dim = 1; side = 0; break;
        case 1: ix[0] = N; ix[1] = j; dim = 0; side = 1; break;
        case 2: ix[0] = N - j - 1; ix[1] = N; dim = 1; side = 1; break;
        default: ix[0] = -1; ix[1] = N - j - 1; dim = 0; side = 0; break;
      }
    } else {
      const int face = f / (N * N), j = (f % (N * N)) / N, k = f % N;
      switch (face) {
        case 0: ix[0] = k; ix[1] = j; ix[2] = -1; dim = 2; side = 0; break;
        case 1: ix[0] = N - k - 1; ix[1] = j; ix[2] = N; dim = 2; side = 1; break;
        case 2: ix[0] = -1; ix[1] = k; ix[2] = j; dim = 0; side = 0; break;
        case 3: ix[0] = N; ix[1] = N - k - 1; ix[2] = j; dim = 0; side = 1; break;
        case 4: ix[0] = N - k - 1; ix[1] = -1; ix[2] = j; dim = 1; side = 0; break;
        default: ix[0] = k; ix[1] = N; ix[2] = j; dim = 1; side = 1; break;
      }
    }
    for (int d = 0; d < D; d++) {
      idx_fpts[f * D + d] = ix[d];
      loc_fpts[f * D + d] = (ix[d] < 0) ? -1.0 : (ix[d] >= N ? 1.0 : loc1d[ix[d]]);
    }
    tnorm[f * D + dim] = side ? 1.0 : -1.0;
  }
  // 1D tables
  ops1d.assign(2 * N + N * N + 2 * N, 0.0);
  for (int m = 0; m < N; m++) {
    ops1d[m] = lagrange(loc1d, m, -1.0);
    ops1d[N + m] = lagrange(loc1d, m, 1.0);
  }
  for (int i = 0; i < N; i++) {
    for (int m = 0; m < N; m++) ops1d[2 * N + i * N + m] = lagrange_d1(locDFR, m + 1, loc1d[i]);
    ops1d[2 * N + N * N + 2 * i] = lagrange_d1(locDFR, 0, loc1d[i]);
    ops1d[2 * N + N * N + 2 * i + 1] = lagrange_d1(locDFR, N + 1, loc1d[i]);
  }
}

// Bi/trilinear shape functions on the corner grid {-1, 1}; node numbering is Gmsh's
// (corner ijk -> gmsh index {0,1,3,2,4,5,7,6}, src/funcs.cpp:1007-1091).
void RefElement::shape(const double *loc, double *val, double *dval) const {
  static const int ijk2gmsh[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  static const std::vector<double> xl = {-1.0, 1.0};
  double l[3][2], dl[3][2];
  for (int d = 0; d < nDims; d++)
    for (int i = 0; i < 2; i++) {
      l[d][i] = lagrange(xl, i, loc[d]);
      dl[d][i] = lagrange_d1(xl, i, loc[d]);
    }
  const int nc = (nDims == 2) ? 4 : 8;
  for (int pt = 0; pt < nc; pt++) {
    const int i = pt & 1, j = (pt >> 1) & 1, k = (pt >> 2) & 1;
    const int g = ijk2gmsh[pt];
    if (nDims == 2) {
      val[g] = l[0][i] * l[1][j];
      dval[g * 2] = dl[0][i] * l[1][j];
      dval[g * 2 + 1] = l[0][i] * dl[1][j];
    } else {
      val[g] = l[0][i] * l[1][j] * l[2][k];
      dval[g * 3] = dl[0][i] * l[1][j] * l[2][k];
      dval[g * 3 + 1] = l[0][i] * dl[1][j] * l[2][k];
      dval[g * 3 + 2] = l[0][i] * l[1][j] * dl[2][k];
    }
  }
}

// Expand the 1D tables into the reference's dense operators (layouts of
// src/elements.cpp:558-563).  Used only for inspection and tests.
std::vector<double> RefElement::dense_operator(const std::string &name) const {
  const int D = nDims, nS = nSpts, nF = nFpts;
  auto lE = [&](int side, int m) { return ops1d[side * N + m]; };
  auto Dss = [&](int i, int m) { return ops1d[2 * N + i * N + m]; };
  auto Dsf = [&](int i, int side) { return ops1d[2 * N + N * N + 2 * i + side]; };
  auto coord = [&](int s, int d) { return idx_spts[s * D + d]; };
  // does spt s lie on the line through spt/fpt coordinates 'ix' along 'dim'?
  auto on_line = [&](int s, const int *ix, int dim) {
    for (int d = 0; d < D; d++)
      if (d != dim && coord(s, d) != ix[d]) return false;
    return true;
  };
  auto fdim = [&](int f) {
    for (int d = 0; d < D; d++)
      if (idx_fpts[f * D + d] < 0 || idx_fpts[f * D + d] >= N) return d;
    return -1;
  };
  std::vector<double> out;
  if (name == "oppE") {
    out.assign((size_t)nF * nS, 0.0);
    for (int f = 0; f < nF; f++) {
      const int dim = fdim(f), side = idx_fpts[f * D + dim] < 0 ? 0 : 1;
      for (int s = 0; s < nS; s++)
        if (on_line(s, &idx_fpts[f * D], dim)) out[(size_t)f * nS + s] = lE(side, coord(s, dim));
    }
  } else if (name == "oppD" || name == "oppDiv") {
    out.assign((size_t)D * nS * nS, 0.0);
    for (int dim = 0; dim < D; dim++)
      for (int i = 0; i < nS; i++)
        for (int j = 0; j < nS; j++) {
          if (!on_line(j, &idx_spts[i * D], dim)) continue;
          const double v = Dss(coord(i, dim), coord(j, dim));
          if (name == "oppD") out[((size_t)dim * nS + i) * nS + j] = v;
          else out[((size_t)i * D + dim) * nS + j] = v;
        }
  } else if (name == "oppD_fpts" || name == "oppDiv_fpts") {
    const bool div = name == "oppDiv_fpts";
    out.assign(div ? (size_t)nS * nF : (size_t)D * nS * nF, 0.0);
    for (int f = 0; f < nF; f++) {
      const int dim = fdim(f), side = idx_fpts[f * D + dim] < 0 ? 0 : 1;
      for (int s = 0; s < nS; s++) {
        if (!on_line(s, &idx_fpts[f * D], dim)) continue;
        const double v = Dsf(coord(s, dim), side);
        if (div) out[(size_t)s * nF + f] = side ? v : -v;
        else out[((size_t)dim * nS + s) * nF + f] = v;
      }
    }
  } else {
    raise("unknown operator " + name);
  }
  return out;
}

}  // namespace zfr
