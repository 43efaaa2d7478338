// physics.hpp -- pointwise Euler / Navier-Stokes physics of the residual path and its
// reverse-mode (vector-Jacobian) counterparts, shared by the HIP kernels and by a host
// build used to check the derivatives on the CPU.
//
// The reference forms the full flux tensor F[var][dim] (include/flux.hpp:1-202) and then
// contracts it with a vector: the face normal (src/faces.cpp:1313-1322, 1580-1590) or a
// row of the adjugate metric (src/elements.cpp:1676-1687).  Here the contraction is done
// analytically: Fn = F . n is assembled from primitive quantities without materialising
// F, which keeps register pressure low.  Terms are the same; only FP64 rounding order
// differs.
//
// The parameter object `a` of every function is any struct with the members gamma,
// prandtl, mu, rt, c_sth, fix_vis, rus_k, ldg_b, ldg_tau (the kernels' argument block or
// the host-side zfr_phys_t).
//
// Reverse mode: each *_vjp function takes the cotangent of its forward function's output
// and ADDS the cotangents of its inputs to the output arrays (Ub, ...), so contributions
// from several uses of one input accumulate.  Derivations are by hand from the forward
// formulas below; nothing here has a counterpart in the reference's explicit path (its
// implicit solver has separate flux Jacobians, include/flux.hpp:205-721).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ZFR_HD __host__ __device__ __forceinline__
#define ZFR_UNROLL _Pragma("unroll")
#else
#define ZFR_HD inline
#define ZFR_UNROLL
#endif

namespace zphys {

template <int D>
struct Prim {
  double invrho, P, H;
  double vel[D];
};

// include/flux.hpp:11-63: invrho, pressure, total enthalpy, velocity
template <int D>
ZFR_HD Prim<D> primitives(const double *U, double gamma) {
  constexpr int V = D + 2;
  Prim<D> p;
  p.invrho = 1. / U[0];
  double momF = U[1] * U[1];
  ZFR_UNROLL
  for (int d = 1; d < D; d++) momF += U[d + 1] * U[d + 1];
  momF *= p.invrho;
  p.P = (gamma - 1.0) * (U[V - 1] - 0.5 * momF);
  p.H = (U[V - 1] + p.P) * p.invrho;
  ZFR_UNROLL
  for (int d = 0; d < D; d++) p.vel[d] = U[d + 1] * p.invrho;
  return p;
}

// Inviscid flux contracted with n: Fn[0] = m.n, Fn[1+k] = m_k (m.n)/rho + P n_k,
// Fn[V-1] = H (m.n), with m the momentum.
template <int D>
ZFR_HD void conv_proj(const double *U, const Prim<D> &p, const double *n, double *Fn) {
  constexpr int V = D + 2;
  double mn = U[1] * n[0];
  ZFR_UNROLL
  for (int d = 1; d < D; d++) mn += U[d + 1] * n[d];
  Fn[0] = mn;
  const double un = mn * p.invrho;
  ZFR_UNROLL
  for (int k = 0; k < D; k++) Fn[k + 1] = U[k + 1] * un + p.P * n[k];
  Fn[V - 1] = p.H * mn;
}

// Ub += (d conv_proj / dU)^T Fb.  With un = m.n / rho, A = sum_k Fb[1+k] u_k,
// B = sum_k Fb[1+k] n_k and P_rho = (g-1)|u|^2/2, P_m = -(g-1) u, P_E = g-1:
//   rho: -un A + B P_rho + Fb[V-1] (P_rho - H) un
//   m_j: Fb[0] n_j + Fb[1+j] un + n_j A + P_m_j B + Fb[V-1] (P_m_j un + H n_j)
//   E  : (g-1) B + g un Fb[V-1]
template <int D>
ZFR_HD void conv_proj_vjp(const double *U, const Prim<D> &p, const double *n, const double *Fb, double gamma,
                          double *Ub) {
  constexpr int V = D + 2;
  double mn = U[1] * n[0];
  ZFR_UNROLL
  for (int d = 1; d < D; d++) mn += U[d + 1] * n[d];
  const double un = mn * p.invrho;
  double A = 0.0, B = 0.0, u2 = 0.0;
  ZFR_UNROLL
  for (int k = 0; k < D; k++) {
    A += Fb[k + 1] * p.vel[k];
    B += Fb[k + 1] * n[k];
    u2 += p.vel[k] * p.vel[k];
  }
  const double gm1 = gamma - 1.0;
  const double Prho = 0.5 * gm1 * u2;
  const double FE = Fb[V - 1];
  Ub[0] += -un * A + B * Prho + FE * (Prho - p.H) * un;
  ZFR_UNROLL
  for (int j = 0; j < D; j++) {
    const double Pm = -gm1 * p.vel[j];
    Ub[j + 1] += Fb[0] * n[j] + Fb[j + 1] * un + n[j] * A + Pm * B + FE * (Pm * un + p.H * n[j]);
  }
  Ub[V - 1] += gm1 * B + gamma * un * FE;
}

// Viscous stress tau and heat + work flux q of include/flux.hpp:75-202:
// F_visc[1+i][j] = -tau[i][j], F_visc[V-1][j] = -q[j].
template <int D>
struct Visc {
  double tau[D][D];
  double q[D];
};

// Dynamic viscosity: constant, or Sutherland's law (include/flux.hpp:95-99) with r^1.5
// evaluated as r * sqrt(r).  The reference's pow(r, 1.5) differs from it by an ulp or
// two; a general pow() on the GPU would also hoist its series constants into ~40
// registers across the face loops.
template <typename P>
ZFR_HD double viscosity(double e_int, const P &a) {
  if (a.fix_vis) return a.mu;
  const double r = (a.gamma - 1.0) * e_int / a.rt;
  return a.mu * (r * sqrt(r)) * (1. + a.c_sth) / (r + a.c_sth);
}

// The gradient enters only through its derivative directions one at a time: direction j
// needs dU[0..V-1][j] and yields dvel[:][j] (velocity gradient) and de[j] (internal energy
// gradient).  fetch(j, dUj) supplies dU[.][j]; a compiler memory barrier after each
// direction keeps the next direction's loads from being hoisted, so only one direction's
// V values are live (the gradient is often gathered straight from LDS or global memory).
template <int D, typename Fetch, typename P>
ZFR_HD Visc<D> viscous_dir(const double *U, const Prim<D> &p, Fetch &&fetch, const P &a) {
  constexpr int V = D + 2;
  Visc<D> s;
  double ke2 = p.vel[0] * p.vel[0];
  ZFR_UNROLL
  for (int d = 1; d < D; d++) ke2 += p.vel[d] * p.vel[d];
  const double e_int = U[V - 1] * p.invrho - 0.5 * ke2;
  const double mu = viscosity(e_int, a);
  double dvel[D][D], de[D];
  ZFR_UNROLL
  for (int j = 0; j < D; j++) {
    double dUj[V];
    fetch(j, dUj);
    ZFR_UNROLL
    for (int i = 0; i < D; i++) dvel[i][j] = (dUj[i + 1] - dUj[0] * p.vel[i]) * p.invrho;
    double w = p.vel[0] * dvel[0][j];
    ZFR_UNROLL
    for (int i = 1; i < D; i++) w += p.vel[i] * dvel[i][j];
    const double dke = 0.5 * ke2 * dUj[0] + U[0] * w;
    de[j] = (dUj[V - 1] - dke - dUj[0] * e_int) * p.invrho;
    asm volatile("" ::: "memory");
  }
  double div = dvel[0][0];
  ZFR_UNROLL
  for (int i = 1; i < D; i++) div += dvel[i][i];
  const double diag = div / 3.0;
  ZFR_UNROLL
  for (int i = 0; i < D; i++)
    ZFR_UNROLL
    for (int j = 0; j < D; j++)
      s.tau[i][j] = (i == j) ? 2.0 * mu * (dvel[i][i] - diag) : mu * (dvel[i][j] + dvel[j][i]);
  const double kc = (mu / a.prandtl) * a.gamma;
  ZFR_UNROLL
  for (int j = 0; j < D; j++) {
    double work = p.vel[0] * s.tau[0][j];
    ZFR_UNROLL
    for (int i = 1; i < D; i++) work += p.vel[i] * s.tau[i][j];
    s.q[j] = work + kc * de[j];
  }
  return s;
}

template <int D, typename P>
ZFR_HD Visc<D> viscous(const double *U, const Prim<D> &p, const double dU[][D], const P &a) {
  return viscous_dir<D>(U, p, [&](int j, double *dUj) {
    ZFR_UNROLL
    for (int v = 0; v < D + 2; v++) dUj[v] = dU[v][j];
  }, a);
}

// Fn -= F_visc . n
template <int D>
ZFR_HD void visc_proj_sub(const Visc<D> &s, const double *n, double *Fn) {
  constexpr int V = D + 2;
  ZFR_UNROLL
  for (int i = 0; i < D; i++) {
    double t = s.tau[i][0] * n[0];
    ZFR_UNROLL
    for (int j = 1; j < D; j++) t += s.tau[i][j] * n[j];
    Fn[i + 1] -= t;
  }
  double qn = s.q[0] * n[0];
  ZFR_UNROLL
  for (int j = 1; j < D; j++) qn += s.q[j] * n[j];
  Fn[V - 1] -= qn;
}

// Viscous flux of one side contracted with n (zero-initialised output, as the reference's
// FL / FR in src/faces.cpp:1555-1560).
template <int D, typename Fetch, typename P>
ZFR_HD void visc_normal(const double *U, Fetch &&fetch, const double *n, const P &a, double *G) {
  constexpr int V = D + 2;
  ZFR_UNROLL
  for (int v = 0; v < V; v++) G[v] = 0.0;
  const Prim<D> p = primitives<D>(U, a.gamma);
  visc_proj_sub<D>(viscous_dir<D>(U, p, fetch, a), n, G);
}

// Reverse of the viscous flux T[1+i][j] = tau_ij, T[V-1][j] = q_j (the tensor that is
// subtracted from the convective flux) for a cotangent Phi[V][D] on T:
//   Ub  += (dT/dU)^T Phi,   dUb(j, .) receives (dT/d dU[.][j])^T Phi direction by direction.
// fetch(j, dUj) supplies the gradient as in viscous_dir; sink(j, dUbj) receives V values.
// phi(v, j) returns Phi[v][j].
template <int D, typename Fetch, typename Phi, typename Sink, typename P>
ZFR_HD void viscous_vjp(const double *U, const Prim<D> &p, Fetch &&fetch, Phi &&phi, Sink &&sink, const P &a,
                        double *Ub) {
  constexpr int V = D + 2;
  // forward intermediates
  double ke2 = 0.0;
  ZFR_UNROLL
  for (int d = 0; d < D; d++) ke2 += p.vel[d] * p.vel[d];
  const double e_int = U[V - 1] * p.invrho - 0.5 * ke2;
  const double mu = viscosity(e_int, a);
  double dvel[D][D], de[D], d0[D];
  ZFR_UNROLL
  for (int j = 0; j < D; j++) {
    double dUj[V];
    fetch(j, dUj);
    ZFR_UNROLL
    for (int i = 0; i < D; i++) dvel[i][j] = (dUj[i + 1] - dUj[0] * p.vel[i]) * p.invrho;
    double w = 0.0;
    ZFR_UNROLL
    for (int i = 0; i < D; i++) w += p.vel[i] * dvel[i][j];
    const double dke = 0.5 * ke2 * dUj[0] + U[0] * w;
    de[j] = (dUj[V - 1] - dke - dUj[0] * e_int) * p.invrho;
    d0[j] = dUj[0];
    asm volatile("" ::: "memory");
  }
  double div = 0.0;
  ZFR_UNROLL
  for (int i = 0; i < D; i++) div += dvel[i][i];
  double S[D][D];
  ZFR_UNROLL
  for (int i = 0; i < D; i++)
    ZFR_UNROLL
    for (int j = 0; j < D; j++) S[i][j] = dvel[i][j] + dvel[j][i] - ((i == j) ? (2.0 / 3.0) * div : 0.0);
  const double kc = (mu / a.prandtl) * a.gamma;
  // reverse: q_j = sum_i u_i tau_ij + kc de_j,  tau = mu S
  double tb[D][D], ub[D], deb[D], kcb = 0.0, mub = 0.0;
  ZFR_UNROLL
  for (int i = 0; i < D; i++) ub[i] = 0.0;
  ZFR_UNROLL
  for (int j = 0; j < D; j++) {
    const double qb = phi(V - 1, j);
    ZFR_UNROLL
    for (int i = 0; i < D; i++) {
      tb[i][j] = phi(i + 1, j) + qb * p.vel[i];
      ub[i] += qb * mu * S[i][j];
    }
    deb[j] = qb * kc;
    kcb += qb * de[j];
  }
  double trb = 0.0;
  ZFR_UNROLL
  for (int i = 0; i < D; i++) {
    trb += tb[i][i];
    ZFR_UNROLL
    for (int j = 0; j < D; j++) mub += tb[i][j] * S[i][j];
  }
  mub += kcb * a.gamma / a.prandtl;
  double velb[D][D];  // d tau / d dvel
  ZFR_UNROLL
  for (int i = 0; i < D; i++)
    ZFR_UNROLL
    for (int j = 0; j < D; j++) velb[i][j] = mu * (tb[i][j] + tb[j][i]) - ((i == j) ? (2.0 / 3.0) * mu * trb : 0.0);
  // The reverse sweep needs the gradient only through dU[0][j] (kept in d0) and through
  // X_j = de_j / invrho and dU[1+i][j] - dU[0][j] u_i = dvel_ij / invrho, so it does not
  // fetch the gradient again.
  double eib = 0.0, ke2b = 0.0, irb = 0.0, rhob = 0.0;
  ZFR_UNROLL
  for (int j = 0; j < D; j++) {
    double dUb[V];
    // de_j = X_j invrho, X_j = dU[V-1][j] - dke_j - dU[0][j] e_int
    double w = 0.0;
    ZFR_UNROLL
    for (int i = 0; i < D; i++) w += p.vel[i] * dvel[i][j];
    const double Xb = deb[j] * p.invrho;
    irb += deb[j] * de[j] * U[0];
    dUb[V - 1] = Xb;
    const double dkeb = -Xb;
    dUb[0] = -Xb * e_int;
    eib += -Xb * d0[j];
    // dke = ke2 dU0 / 2 + rho w
    ke2b += 0.5 * dkeb * d0[j];
    dUb[0] += 0.5 * ke2 * dkeb;
    rhob += dkeb * w;
    const double wb = dkeb * U[0];
    // dvel_ij = (dU[1+i][j] - dU[0][j] u_i) invrho, w_j = sum_i u_i dvel_ij
    ZFR_UNROLL
    for (int i = 0; i < D; i++) {
      const double vb = velb[i][j] + wb * p.vel[i];
      ub[i] += wb * dvel[i][j] - vb * d0[j] * p.invrho;
      dUb[i + 1] = vb * p.invrho;
      dUb[0] -= vb * p.vel[i] * p.invrho;
      irb += vb * dvel[i][j] * U[0];
    }
    sink(j, dUb);
  }
  // mu(e_int)
  if (!a.fix_vis) {
    const double r = (a.gamma - 1.0) * e_int / a.rt;
    const double dmu_dr = mu * (1.5 / r - 1.0 / (r + a.c_sth));
    eib += mub * dmu_dr * (a.gamma - 1.0) / a.rt;
  }
  // e_int = E invrho - ke2 / 2, ke2 = |u|^2, u = m invrho
  Ub[V - 1] += eib * p.invrho;
  irb += eib * U[V - 1];
  ke2b += -0.5 * eib;
  ZFR_UNROLL
  for (int i = 0; i < D; i++) {
    ub[i] += 2.0 * ke2b * p.vel[i];
    Ub[i + 1] += ub[i] * p.invrho;
    irb += ub[i] * U[i + 1];
  }
  Ub[0] += rhob - irb * p.invrho * p.invrho;
}

// Reverse of primitives' P and of the wave speed lambda = |m.n|/rho + sqrt(g P / rho):
// Ub += (d lambda / dU) lb.
template <int D>
ZFR_HD void wavespeed_vjp(const double *U, const Prim<D> &p, const double *n, double gamma, double lb,
                          double *Ub) {
  constexpr int V = D + 2;
  double mn = U[1] * n[0];
  ZFR_UNROLL
  for (int d = 1; d < D; d++) mn += U[d + 1] * n[d];
  const double un = mn * p.invrho;
  const double snd = sqrt(gamma * p.P * p.invrho);
  // |un|
  const double unb = (un >= 0.0) ? lb : -lb;
  ZFR_UNROLL
  for (int j = 0; j < D; j++) Ub[j + 1] += unb * n[j] * p.invrho;
  double rhob = -unb * un * p.invrho;
  // a = sqrt(s), s = g P invrho
  const double sb = lb / (2.0 * snd);
  const double Pb = sb * gamma * p.invrho;
  rhob += -sb * gamma * p.P * p.invrho * p.invrho;
  // P = (g-1)(E - |m|^2 invrho / 2)
  const double gm1 = gamma - 1.0;
  double u2 = 0.0;
  ZFR_UNROLL
  for (int j = 0; j < D; j++) {
    u2 += p.vel[j] * p.vel[j];
    Ub[j + 1] += -Pb * gm1 * p.vel[j];
  }
  rhob += Pb * gm1 * 0.5 * u2;
  Ub[V - 1] += Pb * gm1;
  Ub[0] += rhob;
}

// Roe's approximate Riemann solver (P. L. Roe, J. Comput. Phys. 43 (1981) 357-372) for the
// Euler flux across the unit normal n: the dissipation |A~| (UR - UL) of the Roe matrix at
// the Roe-averaged state, by wave strengths and right eigenvectors (the acoustic waves
// un -+ a, the entropy wave and the shear waves at un).  Harten's entropy fix (A. Harten,
// J. Comput. Phys. 49 (1983) 357-393) smooths the acoustic |lambda| below
// delta = kRoeEntropyFix * a~ to (lambda^2 + delta^2) / (2 delta).  The reference has
// no Roe flux; this is the textbook solver.
constexpr double kRoeEntropyFix = 0.1;

ZFR_HD double harten_abs(double lam, double delta) {
  const double l = fabs(lam);
  return (l < delta) ? 0.5 * (lam * lam + delta * delta) / delta : l;
}

template <int D>
ZFR_HD void roe_dissipation(const double *UL, const double *UR, const Prim<D> &pL, const Prim<D> &pR,
                            const double *n, double gamma, double *diss) {
  constexpr int V = D + 2;
  const double sL = sqrt(UL[0]), sR = sqrt(UR[0]);
  const double wden = 1.0 / (sL + sR);
  double u[D], un = 0.0, q2 = 0.0, unL = 0.0, unR = 0.0;
  ZFR_UNROLL
  for (int d = 0; d < D; d++) {
    u[d] = (sL * pL.vel[d] + sR * pR.vel[d]) * wden;
    un += u[d] * n[d];
    q2 += u[d] * u[d];
    unL += pL.vel[d] * n[d];
    unR += pR.vel[d] * n[d];
  }
  const double H = (sL * pL.H + sR * pR.H) * wden;
  const double rho = sL * sR;
  const double a2 = (gamma - 1.0) * (H - 0.5 * q2);
  const double a = sqrt(a2);
  const double dr = UR[0] - UL[0], dp = pR.P - pL.P, dun = unR - unL;
  const double delta = kRoeEntropyFix * a;
  const double l1 = harten_abs(un - a, delta), l2 = fabs(un), l3 = harten_abs(un + a, delta);
  const double a1 = 0.5 * (dp - rho * a * dun) / a2;  // strength of the un - a wave
  const double a3 = 0.5 * (dp + rho * a * dun) / a2;  // un + a
  const double a2e = dr - dp / a2;                     // entropy wave
  const double w1 = l1 * a1, w3 = l3 * a3, w2 = l2 * a2e;
  // shear waves: rho (du - dun n) at speed un
  double du_t[D], ut_du = 0.0;
  ZFR_UNROLL
  for (int d = 0; d < D; d++) {
    du_t[d] = (pR.vel[d] - pL.vel[d]) - dun * n[d];
    ut_du += u[d] * du_t[d];
  }
  diss[0] = w1 + w2 + w3;
  ZFR_UNROLL
  for (int d = 0; d < D; d++)
    diss[d + 1] = w1 * (u[d] - a * n[d]) + w2 * u[d] + w3 * (u[d] + a * n[d]) + l2 * rho * du_t[d];
  diss[V - 1] = w1 * (H - a * un) + w2 * 0.5 * q2 + w3 * (H + a * un) + l2 * rho * ut_du;
}

// Reverse of roe_dissipation: ULb += (d diss / d UL)^T db, URb likewise.  Derived by hand
// from the forward statements above, in reverse order; the kinks (|lambda| at the entropy
// fix threshold and at 0) take the branch the forward evaluation took.
template <int D>
ZFR_HD void roe_dissipation_vjp(const double *UL, const double *UR, const Prim<D> &pL, const Prim<D> &pR,
                                const double *n, double gamma, const double *db, double *ULb, double *URb) {
  constexpr int V = D + 2;
  // forward recompute
  const double sL = sqrt(UL[0]), sR = sqrt(UR[0]);
  const double wden = 1.0 / (sL + sR);
  double u[D], un = 0.0, q2 = 0.0, unL = 0.0, unR = 0.0;
  ZFR_UNROLL
  for (int d = 0; d < D; d++) {
    u[d] = (sL * pL.vel[d] + sR * pR.vel[d]) * wden;
    un += u[d] * n[d];
    q2 += u[d] * u[d];
    unL += pL.vel[d] * n[d];
    unR += pR.vel[d] * n[d];
  }
  const double H = (sL * pL.H + sR * pR.H) * wden;
  const double rho = sL * sR;
  const double a2 = (gamma - 1.0) * (H - 0.5 * q2);
  const double a = sqrt(a2);
  const double dr = UR[0] - UL[0], dp = pR.P - pL.P, dun = unR - unL;
  const double delta = kRoeEntropyFix * a;
  const double lam1 = un - a, lam3 = un + a;
  const double l1 = harten_abs(lam1, delta), l2 = fabs(un), l3 = harten_abs(lam3, delta);
  const double A1 = 0.5 * (dp - rho * a * dun) / a2, A3 = 0.5 * (dp + rho * a * dun) / a2;
  const double A2 = dr - dp / a2;
  const double w1 = l1 * A1, w3 = l3 * A3, w2 = l2 * A2;
  double dt[D], ut = 0.0;
  ZFR_UNROLL
  for (int d = 0; d < D; d++) {
    dt[d] = (pR.vel[d] - pL.vel[d]) - dun * n[d];
    ut += u[d] * dt[d];
  }
  // reverse: outputs
  const double dbE = db[V - 1];
  double bw1 = db[0] + dbE * (H - a * un), bw2 = db[0] + dbE * 0.5 * q2, bw3 = db[0] + dbE * (H + a * un);
  double sh = dbE * ut, bu[D], bdt[D];
  double ba = dbE * (w3 - w1) * un, bun = dbE * (w3 - w1) * a, bH = dbE * (w1 + w3), bq2 = dbE * 0.5 * w2;
  ZFR_UNROLL
  for (int d = 0; d < D; d++) {
    const double b = db[d + 1];
    bw1 += b * (u[d] - a * n[d]);
    bw2 += b * u[d];
    bw3 += b * (u[d] + a * n[d]);
    sh += b * dt[d];
    ba += b * (w3 - w1) * n[d];
    bu[d] = b * (w1 + w2 + w3) + dbE * l2 * rho * dt[d];
    bdt[d] = l2 * rho * (b + dbE * u[d]);
  }
  double bl2 = rho * sh, brho = l2 * sh;
  // w = l A
  const double bl1 = bw1 * A1, bA1 = bw1 * l1, bl3 = bw3 * A3, bA3 = bw3 * l3;
  bl2 += bw2 * A2;
  const double bA2 = bw2 * l2;
  // wave strengths
  double bdp = 0.5 * (bA1 + bA3) / a2 - bA2 / a2;
  double bdun = 0.5 * rho * a * (bA3 - bA1) / a2;
  brho += 0.5 * a * dun * (bA3 - bA1) / a2;
  ba += 0.5 * rho * dun * (bA3 - bA1) / a2;
  double ba2 = -(A1 * bA1 + A3 * bA3) / a2 + dp * bA2 / (a2 * a2);
  const double bdr = bA2;
  // |lambda| with the entropy fix
  auto dabs = [&](double lam, double &bdel) {  // d habs / d lam; accumulates d / d delta
    if (fabs(lam) < delta) {
      bdel += (delta * delta - lam * lam) / (2.0 * delta * delta);
      return lam / delta;
    }
    return lam >= 0.0 ? 1.0 : -1.0;
  };
  double dd1 = 0.0, dd3 = 0.0;
  const double g1 = dabs(lam1, dd1), g3 = dabs(lam3, dd3);
  bun += bl1 * g1 + bl3 * g3 + bl2 * (un >= 0.0 ? 1.0 : -1.0);
  ba += -bl1 * g1 + bl3 * g3 + kRoeEntropyFix * (bl1 * dd1 + bl3 * dd3);
  // shear jumps
  double bvL[D], bvR[D];
  ZFR_UNROLL
  for (int d = 0; d < D; d++) {
    bvR[d] = bdt[d];
    bvL[d] = -bdt[d];
    bdun -= bdt[d] * n[d];
  }
  // a, a2, H, q2, un, u
  ba2 += ba / (2.0 * a);
  bH += (gamma - 1.0) * ba2;
  bq2 -= 0.5 * (gamma - 1.0) * ba2;
  double bsL = sR * brho, bsR = sL * brho;
  double bw = (sL * pL.H + sR * pR.H) * bH;
  const double bHL = sL * wden * bH, bHR = sR * wden * bH;
  bsL += pL.H * wden * bH;
  bsR += pR.H * wden * bH;
  ZFR_UNROLL
  for (int d = 0; d < D; d++) {
    bu[d] += 2.0 * u[d] * bq2 + n[d] * bun;
    bvL[d] += n[d] * (-bdun) + sL * wden * bu[d];
    bvR[d] += n[d] * bdun + sR * wden * bu[d];
    bsL += pL.vel[d] * wden * bu[d];
    bsR += pR.vel[d] * wden * bu[d];
    bw += (sL * pL.vel[d] + sR * pR.vel[d]) * bu[d];
  }
  bsL -= wden * wden * bw;
  bsR -= wden * wden * bw;
  // primitives -> conservative, each side: rho, v = m/rho, P = (g-1)(E - |m|^2/(2 rho)), H = (E+P)/rho
  auto side = [&](const double *U, const Prim<D> &p, double bs, double brho_dir, double bP, const double *bv,
                  double bHs, double *Ub) {
    const double bPt = bP + bHs * p.invrho;
    double vv = 0.0;
    ZFR_UNROLL
    for (int d = 0; d < D; d++) vv += p.vel[d] * p.vel[d];
    double br = brho_dir + bs / (2.0 * sqrt(U[0])) - bHs * p.H * p.invrho + bPt * (gamma - 1.0) * 0.5 * vv;
    ZFR_UNROLL
    for (int d = 0; d < D; d++) {
      br -= bv[d] * p.vel[d] * p.invrho;
      Ub[d + 1] += bv[d] * p.invrho - bPt * (gamma - 1.0) * p.vel[d];
    }
    Ub[0] += br;
    Ub[V - 1] += bHs * p.invrho + bPt * (gamma - 1.0);
  };
  side(UL, pL, bsL, -bdr, -bdp, bvL, bHL, ULb);
  side(UR, pR, bsR, bdr, bdp, bvR, bHR, URb);
}

// LDG beta branches of src/faces.cpp:1520-1545, fixed at compile time: BM = 0 any beta
// (both sides' gradients), 1 beta = +1/2 (left side only), 2 beta = -1/2 (right only).
// Common normal flux for this element's side of a flux point: Rusanov
// (src/faces.cpp:1244-1364) plus LDG when viscous (src/faces.cpp:1476-1614), scaled by
// the signed area of the side (Fcomm(0) = F dA0, Fcomm(1) = -F dA1).
// G1 / G2: viscous normal fluxes already formed (G1 = left for BM 0 / 1, right for BM 2;
// G2 = right for BM 0).
// The convective part of common_flux alone (no area), for the kernels that form it before
// the viscous part: Roe here, whose temporaries are then dead before the gradient is read.
template <int D, int FC, typename P>
ZFR_HD void conv_common_flux(const double *UL, const double *UR, const double *n, const P &a, double *F) {
  constexpr int V = D + 2;
  const Prim<D> pL = primitives<D>(UL, a.gamma), pR = primitives<D>(UR, a.gamma);
  double FnL[V], FnR[V];
  conv_proj<D>(UL, pL, n, FnL);
  conv_proj<D>(UR, pR, n, FnR);
  if (FC == 1) {
    double dis[V];
    roe_dissipation<D>(UL, UR, pL, pR, n, a.gamma, dis);
    ZFR_UNROLL
    for (int v = 0; v < V; v++) F[v] = 0.5 * (FnR[v] + FnL[v]) - 0.5 * dis[v];
  } else {
    const double aL = sqrt(a.gamma * pL.P * pL.invrho);
    const double aR = sqrt(a.gamma * pR.P * pR.invrho);
    const double eig = fmax(fabs(FnL[0] * pL.invrho) + aL, fabs(FnR[0] * pR.invrho) + aR);
    const double diss = 0.5 * eig * (1.0 - a.rus_k);
    ZFR_UNROLL
    for (int v = 0; v < V; v++) F[v] = 0.5 * (FnR[v] + FnL[v]) - diss * (UR[v] - UL[v]);
  }
}

// The LDG part of common_flux added to F (no area): G1 / G2 as in common_flux.
template <int D, int BM, typename P>
ZFR_HD void ldg_common_add(const double *UL, const double *UR, const double *G1, const double *G2, const P &a,
                           double *F) {
  constexpr int V = D + 2;
  const double beta = a.ldg_b, tau = a.ldg_tau;
  ZFR_UNROLL
  for (int v = 0; v < V; v++) {
    const double GL = (BM == 2) ? 0.0 : G1[v];
    const double GR = (BM == 0) ? G2[v] : ((BM == 2) ? G1[v] : 0.0);
    F[v] += 0.5 * (GL + GR) + tau * (UL[v] - UR[v]) + beta * (GL - GR);
  }
}

// FC selects the convective flux: 0 Rusanov (the reference's), 1 Roe (roe_dissipation).
template <int D, bool VISC, int BM, int FC = 0, typename P>
ZFR_HD void common_flux(const double *UL, const double *UR, const double *G1, const double *G2, const double *n,
                        double dAs, const P &a, double *Fout) {
  constexpr int V = D + 2;
  const Prim<D> pL = primitives<D>(UL, a.gamma), pR = primitives<D>(UR, a.gamma);
  double FnL[V], FnR[V];
  conv_proj<D>(UL, pL, n, FnL);
  conv_proj<D>(UR, pR, n, FnR);
  if (FC == 1) {
    double dis[V];
    roe_dissipation<D>(UL, UR, pL, pR, n, a.gamma, dis);
    ZFR_UNROLL
    for (int v = 0; v < V; v++) Fout[v] = (0.5 * (FnR[v] + FnL[v]) - 0.5 * dis[v]) * dAs;
  } else {
    const double aL = sqrt(a.gamma * pL.P * pL.invrho);
    const double aR = sqrt(a.gamma * pR.P * pR.invrho);
    const double eig = fmax(fabs(FnL[0] * pL.invrho) + aL, fabs(FnR[0] * pR.invrho) + aR);
    const double diss = 0.5 * eig * (1.0 - a.rus_k);
    ZFR_UNROLL
    for (int v = 0; v < V; v++) Fout[v] = (0.5 * (FnR[v] + FnL[v]) - diss * (UR[v] - UL[v])) * dAs;
  }
  if (VISC) {
    const double beta = a.ldg_b, tau = a.ldg_tau;
    ZFR_UNROLL
    for (int v = 0; v < V; v++) {
      const double GL = (BM == 2) ? 0.0 : G1[v];
      const double GR = (BM == 0) ? G2[v] : ((BM == 2) ? G1[v] : 0.0);
      const double Fc = 0.5 * (GL + GR) + tau * (UL[v] - UR[v]) + beta * (GL - GR);
      Fout[v] += Fc * dAs;
    }
  }
}

// Reverse of the LDG part of common_flux for the area-scaled cotangent Fs.
template <int D, int BM, typename P>
ZFR_HD void ldg_vjp_add(const double *Fs, const P &a, double *ULb, double *URb, double *G1b, double *G2b) {
  constexpr int V = D + 2;
  const double beta = a.ldg_b, tau = a.ldg_tau;
  ZFR_UNROLL
  for (int v = 0; v < V; v++) {
    ULb[v] += tau * Fs[v];
    URb[v] -= tau * Fs[v];
    const double GLb = (0.5 + beta) * Fs[v], GRb = (0.5 - beta) * Fs[v];
    if (BM == 0) {
      G1b[v] += GLb;
      G2b[v] += GRb;
    } else if (BM == 1) {
      G1b[v] += GLb;
    } else {
      G1b[v] += GRb;
    }
  }
}

// Reverse of common_flux for the cotangent Fb of Fout: adds to ULb, URb and, when
// viscous, to G1b (and G2b for BM 0).  The wave-speed maximum passes its cotangent to
// the larger side (the left one on a tie, as fmax keeps its first argument).
template <int D, bool VISC, int BM, int FC = 0, typename P>
ZFR_HD void common_flux_vjp(const double *UL, const double *UR, const double *n, double dAs, const P &a,
                            const double *Fb, double *ULb, double *URb, double *G1b, double *G2b) {
  constexpr int V = D + 2;
  const Prim<D> pL = primitives<D>(UL, a.gamma), pR = primitives<D>(UR, a.gamma);
  double Fs[V], Fh[V];
  ZFR_UNROLL
  for (int v = 0; v < V; v++) {
    Fs[v] = Fb[v] * dAs;
    Fh[v] = 0.5 * Fs[v];
  }
  conv_proj_vjp<D>(UL, pL, n, Fh, a.gamma, ULb);
  conv_proj_vjp<D>(UR, pR, n, Fh, a.gamma, URb);
  if (FC == 1) {
    double db[V];
    ZFR_UNROLL
    for (int v = 0; v < V; v++) db[v] = -Fh[v];
    roe_dissipation_vjp<D>(UL, UR, pL, pR, n, a.gamma, db, ULb, URb);
    if (VISC) ldg_vjp_add<D, BM>(Fs, a, ULb, URb, G1b, G2b);
    return;
  }
  double mnL = UL[1] * n[0], mnR = UR[1] * n[0];
  ZFR_UNROLL
  for (int d = 1; d < D; d++) {
    mnL += UL[d + 1] * n[d];
    mnR += UR[d + 1] * n[d];
  }
  const double lamL = fabs(mnL * pL.invrho) + sqrt(a.gamma * pL.P * pL.invrho);
  const double lamR = fabs(mnR * pR.invrho) + sqrt(a.gamma * pR.P * pR.invrho);
  const double eig = fmax(lamL, lamR);
  const double diss = 0.5 * eig * (1.0 - a.rus_k);
  double dissb = 0.0;
  ZFR_UNROLL
  for (int v = 0; v < V; v++) {
    ULb[v] += diss * Fs[v];
    URb[v] -= diss * Fs[v];
    dissb -= Fs[v] * (UR[v] - UL[v]);
  }
  const double eigb = 0.5 * (1.0 - a.rus_k) * dissb;
  if (lamL >= lamR) wavespeed_vjp<D>(UL, pL, n, a.gamma, eigb, ULb);
  else wavespeed_vjp<D>(UR, pR, n, a.gamma, eigb, URb);
  if (VISC) ldg_vjp_add<D, BM>(Fs, a, ULb, URb, G1b, G2b);
}

}  // namespace zphys
