// kernels.hip -- HIP kernels (gfx950) for the DFR residual hot path of ZEFR on quads and hexes.
//
// Design (MI355X-first, not a translation of the reference's kernel set):
//  * Element-major storage: element e owns the contiguous block U[e][var][spt].  A
//    workgroup takes EB consecutive elements, so its global loads are one contiguous
//    chunk, staged into LDS.
//  * The reference's constant-operator GEMMs (GiMMiK, src/gimmik_gpu.cu) are replaced by
//    sum-factorised line contractions: every DFR operator on tensor-product elements is
//    a 1D operator applied along the lines of the element (oppE -> lE, oppD -> Dss + Dsf,
//    oppDiv -> the same, oppDiv_fpts -> signed Dsf).  The 1D tables are staged in LDS.
//  * Face coupling is element-centric: each element reads its partner's trace through an
//    integer map (no pointer tables), and computes the common solution / common flux for
//    its own side of the face.  Both sides evaluate the same formula on the same L/R
//    states, so the interface flux stays single-valued.
//  * Navier-Stokes needs two grid-wide dependencies (U traces, then gradient traces), so
//    a residual is three launches: extrapolate -> gradient -> divergence.  The divergence
//    kernel recomputes the corrected gradient instead of storing it: re-reading U and the
//    U trace costs less HBM traffic than writing and re-reading the gradient.
//  * With LDG beta = +-1/2 only one side's gradient enters each common flux, so only that
//    side's gradient trace is written and read.
//
// All arithmetic is FP64.  Reference call sites are cited at each kernel.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "zefr_amd.h"
#include "common.hpp"

namespace zk {

constexpr int kMaxThreads = 256;

template <int D, int N>
struct Shape {
  static constexpr int nS = (D == 2) ? N * N : N * N * N;
  static constexpr int nFF = (D == 2) ? N : N * N;
  static constexpr int nF = 2 * D * nFF;
  static constexpr int V = D + 2;
  static constexpr int nM = D * D + 1;
  static constexpr int FPT = (nF + nS - 1) / nS;  // flux points per thread
  static constexpr int NOPS = 2 * N + N * N + 2 * N;
  static constexpr int NOPS_PAD = (NOPS + 7) & ~7;
  static constexpr int defaultEB() { return (kMaxThreads / nS) > 0 ? kMaxThreads / nS : 1; }
};

// Offsets of the 1D tables inside ops1d.
template <int N> __device__ __forceinline__ const double *tab_lE(const double *o, int side) { return o + side * N; }
template <int N> __device__ __forceinline__ const double *tab_Dss(const double *o, int i) { return o + 2 * N + i * N; }
template <int N> __device__ __forceinline__ const double *tab_Dsf(const double *o, int i) { return o + 2 * N + N * N + 2 * i; }

struct KArgs {
  int E, EB, viscous;
  int fix_vis;
  double gamma, prandtl, mu, rt, c_sth;
  double rus_k, ldg_b, ldg_tau;
  const double *ops;
  const double *metric;
  const double *fgeo;
  const unsigned int *nbr;
  double *Utr;
  double *Gtr;
};

// Flux point f of the reference ordering (src/quads.cpp:86-123, src/hexas.cpp:133-194):
// normal direction, side (0: the -1 end, 1: the +1 end), first spt of the line through
// it along 'dim', and that line's spt stride.
template <int D, int N>
__device__ __forceinline__ void fpt_info(int f, int &dim, int &side, int &base, int &str) {
  if constexpr (D == 2) {
    const int face = f / N, j = f - face * N;
    switch (face) {
      case 0: dim = 1; side = 0; base = j; break;
      case 1: dim = 0; side = 1; base = N * j; break;
      case 2: dim = 1; side = 1; base = N - 1 - j; break;
      default: dim = 0; side = 0; base = N * (N - 1 - j); break;
    }
    str = (dim == 0) ? 1 : N;
  } else {
    constexpr int NN = N * N;
    const int face = f / NN, r = f - face * NN, j = r / N, k = r - j * N;
    switch (face) {
      case 0: dim = 2; side = 0; base = k + N * j; break;
      case 1: dim = 2; side = 1; base = (N - 1 - k) + N * j; break;
      case 2: dim = 0; side = 0; base = N * k + NN * j; break;
      case 3: dim = 0; side = 1; base = N * (N - 1 - k) + NN * j; break;
      case 4: dim = 1; side = 0; base = (N - 1 - k) + NN * j; break;
      default: dim = 1; side = 1; base = k + NN * j; break;
    }
    str = (dim == 0) ? 1 : ((dim == 1) ? N : NN);
  }
}

// The flux point at the 'side' end of the line through spt (a, b, c) along 'dim'.
template <int D, int N>
__device__ __forceinline__ int fpt_end(int a, int b, int c, int dim, int side) {
  if constexpr (D == 2) {
    if (dim == 0) return side ? (N + b) : (3 * N + (N - 1 - b));
    return side ? (2 * N + (N - 1 - a)) : a;
  } else {
    constexpr int NN = N * N;
    if (dim == 0) return side ? (3 * NN + c * N + (N - 1 - b)) : (2 * NN + c * N + b);
    if (dim == 1) return side ? (5 * NN + c * N + a) : (4 * NN + c * N + (N - 1 - a));
    return side ? (NN + b * N + (N - 1 - a)) : (b * N + a);
  }
}

// ---------------------------------------------------------------- pointwise physics
// include/flux.hpp:11-63 (zero grid velocity)
template <int D>
__device__ __forceinline__ void conv_flux(const double *U, double F[][D], double &P, double gamma) {
  constexpr int V = D + 2;
  const double invrho = 1. / U[0];
  double momF = U[1] * U[1];
#pragma unroll
  for (int d = 1; d < D; d++) momF += U[d + 1] * U[d + 1];
  momF *= invrho;
  P = (gamma - 1.0) * (U[V - 1] - 0.5 * momF);
  const double H = (U[V - 1] + P) * invrho;
#pragma unroll
  for (int d = 0; d < D; d++) {
    F[0][d] = U[d + 1];
#pragma unroll
    for (int m = 0; m < D; m++) F[m + 1][d] = U[d + 1] * U[m + 1] * invrho;
    F[d + 1][d] += P;
    F[V - 1][d] = U[d + 1] * H;
  }
}

// include/flux.hpp:75-202: F -= viscous flux(U, grad U)
template <int D>
__device__ __forceinline__ void visc_flux_sub(const double *U, const double dU[][D], double F[][D],
                                              const KArgs &a) {
  constexpr int V = D + 2;
  const double invrho = 1.0 / U[0];
  double vel[D];
  double ke2 = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) { vel[d] = U[d + 1] * invrho; ke2 += vel[d] * vel[d]; }
  const double e_int = U[V - 1] * invrho - 0.5 * ke2;
  double mu;
  if (a.fix_vis) {
    mu = a.mu;
  } else {
    const double rt_ratio = (a.gamma - 1.0) * e_int / a.rt;
    mu = a.mu * pow(rt_ratio, 1.5) * (1. + a.c_sth) / (rt_ratio + a.c_sth);
  }
  double dvel[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) dvel[i][j] = (dU[i + 1][j] - dU[0][j] * vel[i]) * invrho;
  double de[D];
#pragma unroll
  for (int j = 0; j < D; j++) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) s += vel[i] * dvel[i][j];
    const double dke = 0.5 * ke2 * dU[0][j] + U[0] * s;
    de[j] = (dU[V - 1][j] - dke - dU[0][j] * e_int) * invrho;
  }
  double div = 0.0;
#pragma unroll
  for (int i = 0; i < D; i++) div += dvel[i][i];
  const double diag = div / 3.0;
  const double kc = (mu / a.prandtl) * a.gamma;
#pragma unroll
  for (int j = 0; j < D; j++) {
    double w = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++) {
      const double tau = (i == j) ? 2.0 * mu * (dvel[i][i] - diag) : mu * (dvel[i][j] + dvel[j][i]);
      F[i + 1][j] -= tau;
      w += vel[i] * tau;
    }
    F[V - 1][j] -= (w + kc * de[j]);
  }
}

// Common normal flux at one flux point for this element's side: Rusanov
// (src/faces.cpp:1244-1364) plus, when viscous, LDG (src/faces.cpp:1476-1614); the
// result is scaled by the signed area of this side (Fcomm(0) = F dA0, Fcomm(1) = -F dA1).
template <int D>
__device__ __forceinline__ void common_flux(const double *UL, const double *UR, const double dUL[][D],
                                            const double dUR[][D], bool hasL, bool hasR,
                                            const double *n, double dAs, const KArgs &a,
                                            bool viscous, double *Fout) {
  constexpr int V = D + 2;
  double FL[V][D], FR[V][D], PL, PR;
  conv_flux<D>(UL, FL, PL, a.gamma);
  conv_flux<D>(UR, FR, PR, a.gamma);
  const double aL = sqrt(a.gamma * PL / UL[0]);
  const double aR = sqrt(a.gamma * PR / UR[0]);
  double VnL = 0.0, VnR = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    VnL += UL[d + 1] / UL[0] * n[d];
    VnR += UR[d + 1] / UR[0] * n[d];
  }
  const double eig = fmax(fabs(VnL) + aL, fabs(VnR) + aR);
  const double diss = 0.5 * eig * (1.0 - a.rus_k);
#pragma unroll
  for (int v = 0; v < V; v++) {
    double FnL = 0.0, FnR = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) { FnL += FL[v][d] * n[d]; FnR += FR[v][d] * n[d]; }
    Fout[v] = (0.5 * (FnR + FnL) - diss * (UR[v] - UL[v])) * dAs;
  }
  if (viscous) {
    double GL[V][D], GR[V][D];
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int d = 0; d < D; d++) { GL[v][d] = 0.0; GR[v][d] = 0.0; }
    if (hasL) visc_flux_sub<D>(UL, dUL, GL, a);
    if (hasR) visc_flux_sub<D>(UR, dUR, GR, a);
    const double beta = a.ldg_b, tau = a.ldg_tau;
#pragma unroll
    for (int v = 0; v < V; v++) {
      double FnL = 0.0, FnR = 0.0;
#pragma unroll
      for (int d = 0; d < D; d++) { FnL += GL[v][d] * n[d]; FnR += GR[v][d] * n[d]; }
      const double Fc = 0.5 * (FnL + FnR) + tau * (UL[v] - UR[v]) + beta * (FnL - FnR);
      Fout[v] += Fc * dAs;
    }
  }
}

__device__ __forceinline__ void unpack_nbr(unsigned int p, long &ep, int &slot, int &fp) {
  ep = (long)(p >> 9);
  slot = (int)((p >> 8) & 1u);
  fp = (int)(p & 255u);
}

// Which sides' gradients enter the LDG flux (src/faces.cpp:1520-1545).
__device__ __forceinline__ bool need_left_grad(double beta) { return beta != -0.5; }
__device__ __forceinline__ bool need_right_grad(double beta) { return beta != 0.5; }

// ---------------------------------------------------------------- kernel 1: traces
// Elements::extrapolate_U (src/elements.cpp:1337-1360): U_fpts = oppE U_spts.
template <int D, int N>
__global__ __launch_bounds__(kMaxThreads) void k_extrapolate(const double *__restrict__ U, KArgs a) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  extern __shared__ double lds[];
  double *lops = lds;
  double *lU = lds + S::NOPS_PAD;
  const int tid = threadIdx.x, NT = a.EB * nS;
  const long e0 = (long)blockIdx.x * a.EB;
  const long neb = min((long)a.EB, (long)a.E - e0);
  for (int i = tid; i < S::NOPS; i += NT) lops[i] = a.ops[i];
  const double *Ub = U + e0 * V * nS;
  for (long i = tid; i < neb * V * nS; i += NT) lU[i] = Ub[i];
  __syncthreads();
  const int el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  if (el >= neb) return;
  const double *u = lU + el * V * nS;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    if (f < nF) {
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
#pragma unroll
      for (int v = 0; v < V; v++) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < N; m++) s += lE[m] * u[v * nS + base + m * str];
        a.Utr[(e * V + v) * nF + f] = s;
      }
    }
  }
}

// Shared by the gradient and divergence kernels: own + partner U at this thread's flux
// points, and (viscous) the LDG common solution written to lC (src/faces.cpp:1366-1474).
template <int D, int N>
__device__ __forceinline__ void face_states(const KArgs &a, const double *lops, const double *u, double *lC,
                                            long e, int t, bool valid, bool viscous,
                                            double UL[][D + 2], double UR[][D + 2], int *slot) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    slot[k] = 0;
    if (!valid || f >= nF) continue;
    int dim, side, base, str;
    fpt_info<D, N>(f, dim, side, base, str);
    const double *lE = tab_lE<N>(lops, side);
    long ep; int sl, fp;
    unpack_nbr(a.nbr[e * nF + f], ep, sl, fp);
    slot[k] = sl;
    double own[V], par[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; m++) s += lE[m] * u[v * nS + base + m * str];
      own[v] = s;
      par[v] = a.Utr[(ep * V + v) * nF + fp];
    }
#pragma unroll
    for (int v = 0; v < V; v++) {
      UL[k][v] = sl ? par[v] : own[v];
      UR[k][v] = sl ? own[v] : par[v];
    }
    if (viscous) {
      const double beta = a.ldg_b;
#pragma unroll
      for (int v = 0; v < V; v++) {
        const double l = (beta == 0.5) ? 0.0 : UL[k][v];
        const double r = (beta == -0.5) ? 0.0 : UR[k][v];
        lC[v * nF + f] = 0.5 * (l + r) - beta * (l - r);
      }
    }
  }
}

// Corrected reference-space gradient at spt t (oppD U + oppD_fpts Ucomm), mapped to
// physical space with the adjugate metric (src/elements.cpp:1393-1447, 1630-1650).
template <int D, int N>
__device__ __forceinline__ void spt_gradient(const double *lops, const double *u, const double *lC, int t,
                                             const double ij[][D], double detJ, double g[][D]) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  const int ia = t % N, ib = (t / N) % N, ic = (D == 3) ? t / (N * N) : 0;
  double gr[D][V];
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
    const int str = (d == 0) ? 1 : ((d == 1) ? N : N * N);
    const int base = t - id * str;
    const double *Dr = tab_Dss<N>(lops, id);
    const double *Df = tab_Dsf<N>(lops, id);
    const int f0 = fpt_end<D, N>(ia, ib, ic, d, 0), f1 = fpt_end<D, N>(ia, ib, ic, d, 1);
#pragma unroll
    for (int v = 0; v < V; v++) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; m++) s += Dr[m] * u[v * nS + base + m * str];
      s += Df[0] * lC[v * nF + f0];
      s += Df[1] * lC[v * nF + f1];
      gr[d][v] = s;
    }
  }
  const double inv_det = 1.0 / detJ;
#pragma unroll
  for (int v = 0; v < V; v++)
#pragma unroll
    for (int d1 = 0; d1 < D; d1++) {
      double s = 0.0;
#pragma unroll
      for (int d2 = 0; d2 < D; d2++) s += gr[d2][v] * ij[d2][d1];
      g[v][d1] = s * inv_det;
    }
}

template <int D>
__device__ __forceinline__ void load_metric(const KArgs &a, long e, int t, int nS, double ij[][D], double &detJ,
                                            bool want_det) {
  constexpr int nM = D * D + 1;
  const double *m = a.metric + e * nM * nS + t;
#pragma unroll
  for (int d1 = 0; d1 < D; d1++)
#pragma unroll
    for (int d2 = 0; d2 < D; d2++) ij[d1][d2] = m[(d1 * D + d2) * nS];
  detJ = want_det ? m[D * D * nS] : 1.0;
}

// ---------------------------------------------------------------- kernel 2: gradient
// Common solution + corrected gradient + physical transform + gradient extrapolation.
template <int D, int N>
__global__ __launch_bounds__(kMaxThreads) void k_gradient(const double *__restrict__ U, double *__restrict__ dU_out,
                                                          KArgs a) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  extern __shared__ double lds[];
  const int EB = a.EB;
  double *lops = lds;
  double *lU = lops + S::NOPS_PAD;
  double *lC = lU + EB * V * nS;
  double *lG = lC + EB * V * nF;
  const int tid = threadIdx.x, NT = EB * nS;
  const long e0 = (long)blockIdx.x * EB;
  const long neb = min((long)EB, (long)a.E - e0);
  for (int i = tid; i < S::NOPS; i += NT) lops[i] = a.ops[i];
  const double *Ub = U + e0 * V * nS;
  for (long i = tid; i < neb * V * nS; i += NT) lU[i] = Ub[i];
  __syncthreads();
  const int el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  const double *u = lU + el * V * nS;
  double *c = lC + el * V * nF;
  double *g_l = lG + el * D * V * nS;
  double UL[S::FPT][V], UR[S::FPT][V];
  int slot[S::FPT];
  face_states<D, N>(a, lops, u, c, e, t, valid, true, UL, UR, slot);
  __syncthreads();
  if (valid) {
    double ij[D][D], detJ;
    load_metric<D>(a, e, t, nS, ij, detJ, true);
    double g[V][D];
    spt_gradient<D, N>(lops, u, c, t, ij, detJ, g);
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = g[v][d];
    if (dU_out) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) dU_out[((e * D + d) * V + v) * nS + t] = g[v][d];
    }
  }
  __syncthreads();
  if (!valid) return;
  const double beta = a.ldg_b;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    if (f >= nF) continue;
    // the partner needs this side's gradient iff this side is L and dUL is used, or R and dUR is used
    const bool write = slot[k] ? need_right_grad(beta) : need_left_grad(beta);
    if (!write) continue;
    int dim, side, base, str;
    fpt_info<D, N>(f, dim, side, base, str);
    const double *lE = tab_lE<N>(lops, side);
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < V; v++) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < N; m++) s += lE[m] * g_l[(d * V + v) * nS + base + m * str];
        a.Gtr[((e * D + d) * V + v) * nF + f] = s;
      }
  }
}

// ---------------------------------------------------------------- kernel 3: divergence
// compute_F + compute_divF_spts + compute_common_F + compute_divF_fpts.
template <int D, int N, bool VISC>
__global__ __launch_bounds__(kMaxThreads) void k_divergence(const double *__restrict__ U, double *__restrict__ divF,
                                                            KArgs a) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  extern __shared__ double lds[];
  const int EB = a.EB;
  double *lops = lds;
  double *lU = lops + S::NOPS_PAD;
  double *lC = lU + EB * V * nS;
  double *lG = lC + EB * V * nF;  // gradient (viscous), then the transformed flux
  const int tid = threadIdx.x, NT = EB * nS;
  const long e0 = (long)blockIdx.x * EB;
  const long neb = min((long)EB, (long)a.E - e0);
  for (int i = tid; i < S::NOPS; i += NT) lops[i] = a.ops[i];
  const double *Ub = U + e0 * V * nS;
  for (long i = tid; i < neb * V * nS; i += NT) lU[i] = Ub[i];
  __syncthreads();
  const int el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  const double *u = lU + el * V * nS;
  double *c = lC + el * V * nF;
  double *g_l = lG + el * D * V * nS;
  double UL[S::FPT][V], UR[S::FPT][V];
  int slot[S::FPT];
  face_states<D, N>(a, lops, u, c, e, t, valid, VISC, UL, UR, slot);
  if (VISC) __syncthreads();

  // pointwise flux at the spt, transformed to reference space (src/elements.cpp:1588-1699)
  double tF[V][D];
  if (valid) {
    double ij[D][D], detJ;
    load_metric<D>(a, e, t, nS, ij, detJ, VISC);
    double Us[V];
#pragma unroll
    for (int v = 0; v < V; v++) Us[v] = u[v * nS + t];
    double F[V][D], P;
    conv_flux<D>(Us, F, P, a.gamma);
    if (VISC) {
      double g[V][D];
      spt_gradient<D, N>(lops, u, c, t, ij, detJ, g);
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = g[v][d];
      visc_flux_sub<D>(Us, g, F, a);
    }
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int d1 = 0; d1 < D; d1++) {
        double s = 0.0;
#pragma unroll
        for (int d2 = 0; d2 < D; d2++) s += F[v][d2] * ij[d1][d2];
        tF[v][d1] = s;
      }
    if (!VISC) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = tF[v][d];
    }
  }
  if (VISC) __syncthreads();

  // common flux at this thread's flux points -> lC (Ucomm is no longer needed)
  if (valid) {
    const double beta = a.ldg_b;
#pragma unroll
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      const double *fg = a.fgeo + e * (D + 1) * nF + f;
      double n[D];
#pragma unroll
      for (int d = 0; d < D; d++) n[d] = fg[d * nF];
      const double dAs = fg[D * nF];
      double dUL[V][D], dUR[V][D];
      bool hasL = false, hasR = false;
      if (VISC) {
        int dim, side, base, str;
        fpt_info<D, N>(f, dim, side, base, str);
        const double *lE = tab_lE<N>(lops, side);
        long ep; int sl, fp;
        unpack_nbr(a.nbr[e * nF + f], ep, sl, fp);
        hasL = need_left_grad(beta);
        hasR = need_right_grad(beta);
        const bool ownNeed = sl ? hasR : hasL;
        const bool parNeed = sl ? hasL : hasR;
        double og[V][D], pg[V][D];
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int v = 0; v < V; v++) {
            double s = 0.0;
            if (ownNeed) {
#pragma unroll
              for (int m = 0; m < N; m++) s += lE[m] * g_l[(d * V + v) * nS + base + m * str];
            }
            og[v][d] = s;
            pg[v][d] = parNeed ? a.Gtr[((ep * D + d) * V + v) * nF + fp] : 0.0;
          }
#pragma unroll
        for (int v = 0; v < V; v++)
#pragma unroll
          for (int d = 0; d < D; d++) {
            dUL[v][d] = sl ? pg[v][d] : og[v][d];
            dUR[v][d] = sl ? og[v][d] : pg[v][d];
          }
      }
      double Fo[V];
      common_flux<D>(UL[k], UR[k], dUL, dUR, hasL, hasR, n, dAs, a, VISC, Fo);
#pragma unroll
      for (int v = 0; v < V; v++) c[v * nF + f] = Fo[v];
    }
  }
  __syncthreads();
  if (VISC) {
    if (valid) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = tF[v][d];
    }
    __syncthreads();
  }
  if (!valid) return;

  // divergence: oppDiv tF + oppDiv_fpts Fcomm (src/elements.cpp:1449-1501)
  const int ia = t % N, ib = (t / N) % N, ic = (D == 3) ? t / (N * N) : 0;
  double acc[V];
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
    const int str = (d == 0) ? 1 : ((d == 1) ? N : N * N);
    const int base = t - id * str;
    const double *Dr = tab_Dss<N>(lops, id);
#pragma unroll
    for (int v = 0; v < V; v++) {
      double s = acc[v];
#pragma unroll
      for (int m = 0; m < N; m++) s += Dr[m] * g_l[(d * V + v) * nS + base + m * str];
      acc[v] = s;
    }
  }
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
    const double *Df = tab_Dsf<N>(lops, id);
    const int f0 = fpt_end<D, N>(ia, ib, ic, d, 0), f1 = fpt_end<D, N>(ia, ib, ic, d, 1);
#pragma unroll
    for (int v = 0; v < V; v++) {
      acc[v] += (-Df[0]) * c[v * nF + f0];
      acc[v] += Df[1] * c[v * nF + f1];
    }
  }
#pragma unroll
  for (int v = 0; v < V; v++) divF[(e * V + v) * nS + t] = acc[v];
}

// ---------------------------------------------------------------- RK + norm
// src/solver.cpp:2897-2941 (stage), 2997-3062 (final), element-major layout.
template <int D>
__global__ __launch_bounds__(kMaxThreads) void k_rk_stage(double *__restrict__ U, const double *__restrict__ Uini,
                                                          const double *__restrict__ divF,
                                                          const double *__restrict__ metric, long total, int V,
                                                          int nS, double alpha, double dt) {
  constexpr int nM = D * D + 1;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long e = i / (V * nS);
    const int s = (int)(i % nS);
    const double detJ = metric[(e * nM + D * D) * nS + s];
    U[i] = Uini[i] - alpha * dt / detJ * divF[i];
  }
}

template <int D>
__global__ __launch_bounds__(kMaxThreads) void k_rk_final(double *U, const double *Uini,  // may alias (Euler)
                                                          const double *__restrict__ divF,
                                                          const double *__restrict__ metric, long total, int V,
                                                          int nS, double b0, double b1, double b2, double b3,
                                                          int nStages, double dt) {
  constexpr int nM = D * D + 1;
  const double b[4] = {b0, b1, b2, b3};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long e = i / (V * nS);
    const int s = (int)(i % nS);
    const double detJ = metric[(e * nM + D * D) * nS + s];
    double u = Uini[i];
    for (int st = 0; st < nStages; st++) u -= b[st] * dt / detJ * divF[st * total + i];
    U[i] = u;
  }
}

__device__ __forceinline__ double wave_reduce(double x, int type) {
  for (int off = warpSize / 2; off > 0; off >>= 1) {
    const double y = __shfl_down(x, off);
    x = (type == 0) ? fmax(x, y) : x + y;
  }
  return x;
}

// Block partials of the residual norm (src/solver.cpp:7005-7026).
template <int D>
__global__ __launch_bounds__(kMaxThreads) void k_norm_partial(const double *__restrict__ divF,
                                                              const double *__restrict__ metric, long E, int nS,
                                                              int type, double *__restrict__ part) {
  constexpr int V = D + 2, nM = D * D + 1;
  __shared__ double sh[kMaxThreads / 32][V];
  double acc[V];
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = 0.0;
  const long npts = E * nS;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < npts; i += (long)gridDim.x * blockDim.x) {
    const long e = i / nS;
    const int s = (int)(i % nS);
    const double detJ = metric[(e * nM + D * D) * nS + s];
#pragma unroll
    for (int v = 0; v < V; v++) {
      const double x = divF[(e * V + v) * nS + s];
      if (type == 0) acc[v] = fmax(acc[v], fabs(x / detJ));
      else if (type == 1) acc[v] += fabs(x / detJ);
      else acc[v] += x * x / (detJ * detJ);
    }
  }
  const int lane = threadIdx.x % warpSize, w = threadIdx.x / warpSize;
#pragma unroll
  for (int v = 0; v < V; v++) {
    const double r = wave_reduce(acc[v], type);
    if (lane == 0) sh[w][v] = r;
  }
  __syncthreads();
  if (threadIdx.x < V) {
    const int v = threadIdx.x;
    const int nw = blockDim.x / warpSize;
    double r = sh[0][v];
    for (int k = 1; k < nw; k++) r = (type == 0) ? fmax(r, sh[k][v]) : r + sh[k][v];
    part[blockIdx.x * V + v] = r;
  }
}

__global__ __launch_bounds__(kMaxThreads) void k_copy(double *__restrict__ dst, const double *__restrict__ src,
                                                      long n) {
  const long stride = (long)gridDim.x * blockDim.x * 2;
  for (long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 2; i < n; i += stride) {
    if (i + 1 < n) {
      const double2 v = *reinterpret_cast<const double2 *>(src + i);
      *reinterpret_cast<double2 *>(dst + i) = v;
    } else {
      dst[i] = src[i];
    }
  }
}

__global__ void k_norm_final(const double *__restrict__ part, int nblk, int V, int type, double *__restrict__ out) {
  const int v = threadIdx.x;
  if (v >= V) return;
  double r = 0.0;
  for (int b = 0; b < nblk; b++) r = (type == 0) ? fmax(r, part[b * V + v]) : r + part[b * V + v];
  out[v] = r;
}

// ---------------------------------------------------------------- dispatch
// LDS budget per workgroup; kept at 64 KiB so several workgroups share a CU.
constexpr size_t kMaxLds = 64 * 1024;

// Elements per workgroup: the requested tile, limited by the thread count and the LDS
// the kernel needs per element (per_ele doubles).
inline int eles_per_block(const zfr_ele_desc_t *d, int defEB, size_t fixed, size_t per_ele) {
  int eb = d->eles_per_block > 0 ? d->eles_per_block : defEB;
  while (eb > 1 && (eb * d->nSpts > kMaxThreads || sizeof(double) * (fixed + eb * per_ele) > kMaxLds)) eb--;
  return eb < 1 ? 1 : eb;
}

inline KArgs make_args(const zfr_ele_desc_t *d, int EB) {
  KArgs a;
  a.E = d->nEles; a.EB = EB; a.viscous = d->viscous; a.fix_vis = d->fix_vis;
  a.gamma = d->gamma; a.prandtl = d->prandtl; a.mu = d->mu; a.rt = d->rt; a.c_sth = d->c_sth;
  a.rus_k = d->rus_k; a.ldg_b = d->ldg_b; a.ldg_tau = d->ldg_tau;
  a.ops = d->ops1d; a.metric = d->metric; a.fgeo = d->fgeo; a.nbr = d->nbr; a.Utr = d->Utr; a.Gtr = d->Gtr;
  return a;
}

enum class Op { Extrap, Grad, Div };

template <int D, int N>
int launch(Op op, const zfr_ele_desc_t *d, const double *U, double *out, hipStream_t st) {
  using S = Shape<D, N>;
  if (d->nSpts != S::nS || d->nFpts != S::nF || d->nVars != S::V)
    return zfr::fail("kernel launch: element descriptor does not match (nDims, order)");
  if (d->nEles <= 0) return 0;
  const size_t per_ele = (op == Op::Extrap) ? (size_t)S::V * S::nS
                                            : (size_t)(S::V * S::nS + S::V * S::nF + D * S::V * S::nS);
  const int EB = eles_per_block(d, S::defaultEB(), S::NOPS_PAD, per_ele);
  const KArgs a = make_args(d, EB);
  const dim3 grid((unsigned)((d->nEles + EB - 1) / EB)), block((unsigned)(EB * S::nS));
  const size_t lds = sizeof(double) * (S::NOPS_PAD + EB * per_ele);
  switch (op) {
    case Op::Extrap:
      hipLaunchKernelGGL((k_extrapolate<D, N>), grid, block, lds, st, U, a);
      break;
    case Op::Grad:
      hipLaunchKernelGGL((k_gradient<D, N>), grid, block, lds, st, U, out, a);
      break;
    case Op::Div:
      if (d->viscous) hipLaunchKernelGGL((k_divergence<D, N, true>), grid, block, lds, st, U, out, a);
      else hipLaunchKernelGGL((k_divergence<D, N, false>), grid, block, lds, st, U, out, a);
      break;
  }
  return zfr::check_hip(hipGetLastError(), "kernel launch");
}

int dispatch(Op op, const zfr_ele_desc_t *d, const double *U, double *out, hipStream_t st) {
  const int N = d->order + 1;
  if (d->nDims == 2) {
    switch (N) {
      case 1: return launch<2, 1>(op, d, U, out, st);
      case 2: return launch<2, 2>(op, d, U, out, st);
      case 3: return launch<2, 3>(op, d, U, out, st);
      case 4: return launch<2, 4>(op, d, U, out, st);
      case 5: return launch<2, 5>(op, d, U, out, st);
      case 6: return launch<2, 6>(op, d, U, out, st);
    }
  } else if (d->nDims == 3) {
    switch (N) {
      case 1: return launch<3, 1>(op, d, U, out, st);
      case 2: return launch<3, 2>(op, d, U, out, st);
      case 3: return launch<3, 3>(op, d, U, out, st);
      case 4: return launch<3, 4>(op, d, U, out, st);
      case 5: return launch<3, 5>(op, d, U, out, st);
      case 6: return launch<3, 6>(op, d, U, out, st);
    }
  }
  return zfr::fail("kernel launch: supported shapes are quads and hexes with order 0..5");
}

inline unsigned grid_for(long total) {
  long g = (total + kMaxThreads - 1) / kMaxThreads;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace zk

extern "C" {

int zfr_extrapolate(const zfr_ele_desc_t *d, const double *U, void *stream) {
  return zk::dispatch(zk::Op::Extrap, d, U, nullptr, (hipStream_t)stream);
}

int zfr_gradient(const zfr_ele_desc_t *d, const double *U, double *dU_out, void *stream) {
  if (!d->viscous) return zfr::fail("zfr_gradient: descriptor is inviscid");
  return zk::dispatch(zk::Op::Grad, d, U, dU_out, (hipStream_t)stream);
}

int zfr_divergence(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream) {
  return zk::dispatch(zk::Op::Div, d, U, divF, (hipStream_t)stream);
}

int zfr_rk_stage(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF, double alpha,
                 double dt, void *stream) {
  const long total = (long)d->nEles * d->nVars * d->nSpts;
  if (total <= 0) return 0;
  if (d->nDims == 2)
    hipLaunchKernelGGL(zk::k_rk_stage<2>, dim3(zk::grid_for(total)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream,
                       U, U_ini, divF, d->metric, total, d->nVars, d->nSpts, alpha, dt);
  else
    hipLaunchKernelGGL(zk::k_rk_stage<3>, dim3(zk::grid_for(total)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream,
                       U, U_ini, divF, d->metric, total, d->nVars, d->nSpts, alpha, dt);
  return zfr::check_hip(hipGetLastError(), "zfr_rk_stage");
}

int zfr_rk_final(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF, const double *beta,
                 int nStages, double dt, void *stream) {
  if (nStages < 1 || nStages > 4) return zfr::fail("zfr_rk_final: nStages must be 1..4");
  const long total = (long)d->nEles * d->nVars * d->nSpts;
  if (total <= 0) return 0;
  double b[4] = {0, 0, 0, 0};
  for (int i = 0; i < nStages; i++) b[i] = beta[i];
  if (d->nDims == 2)
    hipLaunchKernelGGL(zk::k_rk_final<2>, dim3(zk::grid_for(total)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream,
                       U, U_ini, divF, d->metric, total, d->nVars, d->nSpts, b[0], b[1], b[2], b[3], nStages, dt);
  else
    hipLaunchKernelGGL(zk::k_rk_final<3>, dim3(zk::grid_for(total)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream,
                       U, U_ini, divF, d->metric, total, d->nVars, d->nSpts, b[0], b[1], b[2], b[3], nStages, dt);
  return zfr::check_hip(hipGetLastError(), "zfr_rk_final");
}

int zfr_residual_norm(const zfr_ele_desc_t *d, const double *divF, int res_type, double *work, double *res_out,
                      void *stream) {
  if (res_type < 0 || res_type > 2) return zfr::fail("zfr_residual_norm: res_type must be 0, 1 or 2");
  const long npts = (long)d->nEles * d->nSpts;
  unsigned nblk = zk::grid_for(npts);
  if (nblk > 1024) nblk = 1024;
  hipStream_t st = (hipStream_t)stream;
  if (d->nDims == 2)
    hipLaunchKernelGGL(zk::k_norm_partial<2>, dim3(nblk), dim3(zk::kMaxThreads), 0, st, divF, d->metric,
                       (long)d->nEles, d->nSpts, res_type, work);
  else
    hipLaunchKernelGGL(zk::k_norm_partial<3>, dim3(nblk), dim3(zk::kMaxThreads), 0, st, divF, d->metric,
                       (long)d->nEles, d->nSpts, res_type, work);
  hipLaunchKernelGGL(zk::k_norm_final, dim3(1), dim3(64), 0, st, work, (int)nblk, d->nVars, res_type, res_out);
  return zfr::check_hip(hipGetLastError(), "zfr_residual_norm");
}

int zfr_stream_copy(double *dst, const double *src, long n, void *stream) {
  if (n <= 0) return 0;
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return zfr::fail("zfr_stream_copy: pointers must be 16-byte aligned");
  hipLaunchKernelGGL(zk::k_copy, dim3(4096), dim3(zk::kMaxThreads), 0, (hipStream_t)stream, dst, src, n);
  return zfr::check_hip(hipGetLastError(), "zfr_stream_copy");
}

}  // extern "C"
