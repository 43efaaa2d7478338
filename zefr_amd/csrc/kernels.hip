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
//  * Navier-Stokes with a general LDG beta needs two grid-wide dependencies (U traces,
//    then gradient traces), so a residual is three launches: extrapolate -> gradient ->
//    divergence.  The divergence kernel recomputes the corrected gradient instead of
//    storing it: re-reading U and the U trace costs less HBM traffic than writing and
//    re-reading the gradient.
//  * With LDG beta = +-1/2 (the usual setting) only one side's gradient enters each
//    common flux and the common solution is the other side's trace, so the "owner" side
//    can finish the whole common flux right after its own gradient, and the only other-
//    side data it needs (the partner's trace at its owned points) it extrapolates from the
//    partner's U itself: gradient_flux (gradient, volume divergence, owned common fluxes)
//    -> face_correction, two launches.  No gradient trace is stored at all.
//  * Euler: divergence alone, extrapolating the partner traces from the partner's U (with
//    a fused RK update: extrapolate -> divergence, since the update overwrites U).
//
// All arithmetic is FP64.  Reference call sites are cited at each kernel.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "zefr_amd.h"
#include "common.hpp"
#include "physics.hpp"

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
  double *Ftr;
  // k_extrapolate: 0 = the trace at every flux point; 1 = only where the partner owns the
  // face (the one-sided LDG path reads nothing else)
  int trace_sel;
  // RK update fused into the last kernel of a residual (mode 0: none, 1: stage, 2: final)
  struct Rk {
    double *U;
    const double *Uini, *divF0;  // divF0: stage 0 of the [nStages][E][V][nS] residuals
    const double *src;           // p-multigrid source term added to every divF_s, or null
    long total;
    double alpha, dt, beta[4];
    int nStages, stage, mode;
    double *trace;  // mode 1: also write the new state's trace here (the next stage's Utr), or null
  } rk;
};

// One case of a multi-case launch: its kernel arguments and arrays.  A multi-case launch
// runs the same kernel over several independent solvers of one shape in one grid:
// workgroup b does block b - starts[c] of case c, where starts[c] <= b < starts[c + 1].
struct MultiCase {
  KArgs a;
  const double *U;
  double *out;
};

// The case of workgroup b: binary search over the cases' first workgroups (uniform across
// the workgroup, so the loads are scalar).
__device__ __forceinline__ int find_case(const long *__restrict__ starts, int nc, long b) {
  int lo = 0, hi = nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

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
// primitives, conv_proj, viscous_dir, visc_proj_sub, visc_normal, common_flux and their
// reverse-mode counterparts live in physics.hpp (shared with a host build).
using namespace zphys;

__device__ __forceinline__ void unpack_nbr(unsigned int p, long &ep, int &slot, int &fp) {
  ep = (long)(p >> 9);
  slot = (int)((p >> 8) & 1u);
  fp = (int)(p & 255u);
}

// Which sides' gradients enter the LDG flux (src/faces.cpp:1520-1545).
__device__ __forceinline__ bool need_left_grad(double beta) { return beta != -0.5; }
__device__ __forceinline__ bool need_right_grad(double beta) { return beta != 0.5; }

// Stage EB consecutive elements' U blocks and the 1D tables into LDS.
template <int D, int N>
__device__ __forceinline__ void stage_U(const double *__restrict__ U, const KArgs &a, double *lops, double *lU,
                                        long e0, long neb, int Ustride) {
  using S = Shape<D, N>;
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int i = tid; i < S::NOPS; i += NT) lops[i] = a.ops[i];
  const double *Ub = U + e0 * S::V * S::nS;
  for (long i = tid; i < neb * S::V * S::nS; i += NT) {
    const long el = i / (S::V * S::nS), r = i - el * (S::V * S::nS);
    lU[el * Ustride + r] = Ub[i];
  }
}

// Value of var v along the line (base, str) of a [var][spt] block, extrapolated to one end.
template <int N>
__device__ __forceinline__ double line_extrap(const double *lE, const double *blk, int base, int str) {
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < N; m++) s += lE[m] * blk[base + m * str];
  return s;
}

// The partner's trace at its flux point fp (all V variables), extrapolated from the
// partner's solution line in global memory with the 1D table (global copy of ops1d): what
// k_extrapolate writes there, for kernels that run without a trace launch before them.
template <int D, int N>
__device__ __forceinline__ void partner_trace(const double *__restrict__ U, const double *ops, long ep, int fp,
                                              double *out) {
  using S = Shape<D, N>;
  int dim, side, base, str;
  fpt_info<D, N>(fp, dim, side, base, str);
  const double *lE = tab_lE<N>(ops, side);
  const double *Up = U + ep * S::V * S::nS + base;
#pragma unroll
  for (int v = 0; v < S::V; v++) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < N; m++) s += lE[m] * Up[v * S::nS + m * str];
    out[v] = s;
  }
}

// RK update of one spt straight after its residual, in the last kernel of the residual
// (the stage update of src/solver.cpp:2897-2941 and the final combination of
// :2997-3062, with the same arithmetic as k_rk_stage / k_rk_final).  acc holds this
// stage's divF at the spt; the other stages' come from memory.  With a source term
// (src/solver.cpp:2946-2962, 3061-3073) every divF_s is replaced by divF_s + src.
template <int D>
__device__ __forceinline__ void rk_epilogue(const KArgs &a, long e, int t, int nS, const double *acc,
                                            double *unew = nullptr) {
  constexpr int V = D + 2, nM = D * D + 1;
  if (a.rk.mode == 0) return;
  const double detJ = a.metric[(e * nM + D * D) * nS + t];
#pragma unroll
  for (int v = 0; v < V; v++) {
    const long i = (e * V + v) * nS + t;
    if (a.rk.mode == 1) {
      const double r = a.rk.src ? acc[v] + a.rk.src[i] : acc[v];
      const double u = a.rk.Uini[i] - a.rk.alpha * a.rk.dt / detJ * r;
      a.rk.U[i] = u;
      if (unew) unew[v] = u;
    } else {
      double u = a.rk.Uini[i];
      for (int s = 0; s < a.rk.nStages; s++) {
        double r = (s == a.rk.stage) ? acc[v] : a.rk.divF0[s * a.rk.total + i];
        if (a.rk.src) r += a.rk.src[i];
        u -= a.rk.beta[s] * a.rk.dt / detJ * r;
      }
      a.rk.U[i] = u;
    }
  }
}

// After a fused RK stage update: the trace of the new state at this element's flux points,
// written to a.rk.trace ([E][V][nF], the next stage's Utr), so that the next stage's
// extrapolation launch is skipped.  owner < 0 writes every point (k_extrapolate with
// trace_sel 0); otherwise only the points whose partner owns the face (trace_sel 1 with
// that owner slot).  The new state goes through the element's LDS block w (V * nS
// doubles), which no thread may still read; every thread of the workgroup calls this (it
// holds barriers).
template <int D, int N>
__device__ __forceinline__ void rk_trace(const KArgs &a, const double *lops, double *w, long e, int t, bool valid,
                                         const double *unew, int owner) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  __syncthreads();
  if (valid) {
#pragma unroll
    for (int v = 0; v < V; v++) w[v * nS + t] = unew[v];
  }
  __syncthreads();
  if (!valid) return;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    if (f >= nF) continue;
    if (owner >= 0 && (int)((a.nbr[e * nF + f] >> 8) & 1u) == owner) continue;
    int dim, side, base, str;
    fpt_info<D, N>(f, dim, side, base, str);
    const double *lE = tab_lE<N>(lops, side);
#pragma unroll
    for (int v = 0; v < V; v++) a.rk.trace[(e * V + v) * nF + f] = line_extrap<N>(lE, w + v * nS, base, str);
  }
}

// The k-th owned flux point, owned faces taken in increasing order (mask bit q = face q).
template <int D, int N>
__device__ __forceinline__ int owned_fpt(unsigned mask, int k) {
  using S = Shape<D, N>;
  int q = k / S::nFF;
  const int r = k - q * S::nFF;
#pragma unroll
  for (int b = 0; b < 2 * D; b++) {
    if ((mask >> b) & 1u) {
      if (q == 0) return b * S::nFF + r;
      q--;
    }
  }
  return 0;  // not reached for k < popcount(mask) * nFF
}

// ---------------------------------------------------------------- kernel 1: traces
// Elements::extrapolate_U (src/elements.cpp:1337-1360): U_fpts = oppE U_spts.  With
// trace_sel = 1 only the points of faces the partner owns are written: on the one-sided
// LDG path the owner reads the other side's trace and nothing reads its own.
template <int D, int N>
__device__ __forceinline__ void extrapolate_block(const double *__restrict__ U, const KArgs &a, long bid) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  extern __shared__ double lds[];
  double *lops = lds;
  double *lU = lds + S::NOPS_PAD;
  const long e0 = bid * a.EB;
  const long neb = min((long)a.EB, (long)a.E - e0);
  const int tid = threadIdx.x, el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  // this thread's slots, fetched before the staging barrier so the loads overlap it
  int slot[S::FPT];
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    slot[k] = (a.trace_sel && el < neb && f < nF) ? (int)((a.nbr[e * nF + f] >> 8) & 1u) : -1;
  }
  stage_U<D, N>(U, a, lops, lU, e0, neb, V * nS);
  __syncthreads();
  if (el >= neb) return;
  const double *u = lU + el * V * nS;
  const int owner = (a.ldg_b == 0.5) ? 0 : 1;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    if (f < nF) {
      if (a.trace_sel && slot[k] == owner) continue;
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
#pragma unroll
      for (int v = 0; v < V; v++) a.Utr[(e * V + v) * nF + f] = line_extrap<N>(lE, u + v * nS, base, str);
    }
  }
}

template <int D, int N>
__global__ __launch_bounds__(kMaxThreads) void k_extrapolate(const double *__restrict__ U, KArgs a) {
  extrapolate_block<D, N>(U, a, blockIdx.x);
}

// The same kernel for workgroups of one wave (see k_visc_onesided_w64).
template <int D, int N>
__global__ __launch_bounds__(64) void k_extrapolate_w64(const double *__restrict__ U, KArgs a) {
  extrapolate_block<D, N>(U, a, blockIdx.x);
}

template <int D, int N>
__global__ __launch_bounds__(kMaxThreads) void k_extrapolate_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  extrapolate_block<D, N>(cs[c].U, cs[c].a, (long)blockIdx.x - starts[c]);
}

// LDG common solution at this thread's flux points (src/faces.cpp:1366-1474):
// Ucomm = 1/2 (UL + UR) - beta (UL - UR), with the beta = +-1/2 one-sided branches.
template <int D, int N>
__device__ __forceinline__ void common_solution(const KArgs &a, const double *lops, const double *u, double *lC,
                                                long e, int t) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  const double beta = a.ldg_b;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    if (f >= nF) continue;
    int dim, side, base, str;
    fpt_info<D, N>(f, dim, side, base, str);
    const double *lE = tab_lE<N>(lops, side);
    long ep; int sl, fp;
    unpack_nbr(a.nbr[e * nF + f], ep, sl, fp);
#pragma unroll
    for (int v = 0; v < V; v++) {
      const double own = line_extrap<N>(lE, u + v * nS, base, str);
      const double par = a.Utr[(ep * V + v) * nF + fp];
      const double uL = sl ? par : own, uR = sl ? own : par;
      const double l = (beta == 0.5) ? 0.0 : uL;
      const double r = (beta == -0.5) ? 0.0 : uR;
      lC[v * nF + f] = 0.5 * (l + r) - beta * (l - r);
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
// Common solution + corrected gradient + physical transform + gradient extrapolation to
// the flux points whose gradient the partner element needs.  LDS per element holds U and
// Ucomm, then (aliased) the physical gradient.
template <int D, int N>
__device__ __forceinline__ void gradient_block(const double *__restrict__ U, double *__restrict__ dU_out,
                                               const KArgs &a, long bid) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  constexpr int W = (V * nS + V * nF > D * V * nS) ? V * nS + V * nF : D * V * nS;
  extern __shared__ double lds[];
  double *lops = lds;
  double *lW = lds + S::NOPS_PAD;
  const long e0 = bid * a.EB;
  const long neb = min((long)a.EB, (long)a.E - e0);
  stage_U<D, N>(U, a, lops, lW, e0, neb, W);
  __syncthreads();
  const int tid = threadIdx.x, el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  double *u = lW + el * W;
  double *c = u + V * nS;
  if (valid) common_solution<D, N>(a, lops, u, c, e, t);
  __syncthreads();
  double g[V][D];
  if (valid) {
    double ij[D][D], detJ;
    load_metric<D>(a, e, t, nS, ij, detJ, true);
    spt_gradient<D, N>(lops, u, c, t, ij, detJ, g);
    if (dU_out) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) dU_out[((e * D + d) * V + v) * nS + t] = g[v][d];
    }
  }
  __syncthreads();  // U and Ucomm are dead: the region now holds the gradient
  if (valid) {
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < V; v++) u[(d * V + v) * nS + t] = g[v][d];
  }
  __syncthreads();
  if (!valid) return;
  const double beta = a.ldg_b;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    if (f >= nF) continue;
    const int sl = (int)((a.nbr[e * nF + f] >> 8) & 1u);
    // the partner needs this side's gradient iff this side is L and dUL is used, or R and dUR is used
    const bool write = sl ? need_right_grad(beta) : need_left_grad(beta);
    if (!write) continue;
    int dim, side, base, str;
    fpt_info<D, N>(f, dim, side, base, str);
    const double *lE = tab_lE<N>(lops, side);
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < V; v++)
        a.Gtr[((e * D + d) * V + v) * nF + f] = line_extrap<N>(lE, u + (d * V + v) * nS, base, str);
  }
}

template <int D, int N>
__global__ __launch_bounds__(kMaxThreads) void k_gradient(const double *__restrict__ U, double *__restrict__ dU_out,
                                                          KArgs a) {
  gradient_block<D, N>(U, dU_out, a, blockIdx.x);
}

template <int D, int N>
__global__ __launch_bounds__(kMaxThreads) void k_gradient_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  gradient_block<D, N>(cs[c].U, nullptr, cs[c].a, (long)blockIdx.x - starts[c]);
}

// ---------------------------------------------------------------- kernel 3: divergence
// compute_F + compute_divF_spts + compute_common_F + compute_divF_fpts.
// Phases (viscous): U -> LDS | Ucomm -> lC | gradient -> lG | common flux -> lC |
// transformed flux -> lG | divergence.  Inviscid: U -> LDS | flux + common flux | divergence.
// SELF (Euler only): the partner traces are extrapolated here from the partner's U, so an
// Euler residual is this one launch; otherwise they are read from Utr.
template <int D, int N, bool VISC, int BM, int FC, bool SELF>
__device__ __forceinline__ void divergence_block(const double *__restrict__ U, double *__restrict__ divF,
                                                 const KArgs &a, long bid) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  extern __shared__ double lds[];
  const int EB = a.EB;
  double *lops = lds;
  double *lU = lops + S::NOPS_PAD;
  double *lC = lU + EB * V * nS;
  double *lG = lC + EB * V * nF;  // gradient (viscous), then the transformed flux
  const long e0 = bid * EB;
  const long neb = min((long)EB, (long)a.E - e0);
  stage_U<D, N>(U, a, lops, lU, e0, neb, V * nS);
  __syncthreads();
  const int tid = threadIdx.x, el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  const double *u = lU + el * V * nS;
  double *c = lC + el * V * nF;
  double *g_l = lG + el * D * V * nS;
  if (VISC) {
    if (valid) common_solution<D, N>(a, lops, u, c, e, t);
    __syncthreads();
    if (valid) {
      double ij[D][D], detJ, g[V][D];
      load_metric<D>(a, e, t, nS, ij, detJ, true);
      spt_gradient<D, N>(lops, u, c, t, ij, detJ, g);
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = g[v][d];
    }
    __syncthreads();
  }

  // common flux at this thread's flux points -> lC (Ucomm has been consumed)
  if (valid) {
#pragma unroll 1
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
      long ep; int sl, fp;
      unpack_nbr(a.nbr[e * nF + f], ep, sl, fp);
      const double *fg = a.fgeo + e * (D + 1) * nF + f;
      double n[D];
#pragma unroll
      for (int d = 0; d < D; d++) n[d] = fg[d * nF];
      const double dAs = fg[D * nF];
      // own side's state / gradient from LDS, the partner's from its traces
      auto own_state = [&](double *Us) {
#pragma unroll
        for (int v = 0; v < V; v++) Us[v] = line_extrap<N>(lE, u + v * nS, base, str);
      };
      auto par_state = [&](double *Us) {
        if constexpr (SELF) {
          partner_trace<D, N>(U, a.ops, ep, fp, Us);
        } else {
#pragma unroll
          for (int v = 0; v < V; v++) Us[v] = a.Utr[(ep * V + v) * nF + fp];
        }
      };
      // viscous normal fluxes first (one side's state + gradient live at a time):
      // G1 = left (BM 0, 1) or right (BM 2) side, G2 = right side (BM 0)
      double G1[VISC ? V : 1], G2[(VISC && BM == 0) ? V : 1];
      if (VISC) {
        auto side_visc = [&](bool own, double *G) {
          double Us[V];
          if (own) {
            own_state(Us);
            visc_normal<D>(Us, [&](int j, double *dUj) {
#pragma unroll
              for (int v = 0; v < V; v++) dUj[v] = line_extrap<N>(lE, g_l + (j * V + v) * nS, base, str);
            }, n, a, G);
          } else {
            par_state(Us);
            visc_normal<D>(Us, [&](int j, double *dUj) {
#pragma unroll
              for (int v = 0; v < V; v++) dUj[v] = a.Gtr[((ep * D + j) * V + v) * nF + fp];
            }, n, a, G);
          }
        };
        if (BM == 0) {
          side_visc(sl == 0, G1);
          side_visc(sl == 1, G2);
        } else {
          side_visc((BM == 1) ? (sl == 0) : (sl == 1), G1);
        }
      }
      double UL[V], UR[V];
      if (sl) { par_state(UL); own_state(UR); } else { own_state(UL); par_state(UR); }
      double Fo[V];
      common_flux<D, VISC, BM, FC>(UL, UR, G1, G2, n, dAs, a, Fo);
#pragma unroll
      for (int v = 0; v < V; v++) c[v * nF + f] = Fo[v];
    }
  }
  if (VISC) __syncthreads();  // the gradient field in lG is consumed

  // pointwise flux at the spt, contracted with the adjugate rows (src/elements.cpp:1588-1699);
  // each thread reads and overwrites only its own spt column of lG
  if (valid) {
    double ij[D][D], detJ;
    load_metric<D>(a, e, t, nS, ij, detJ, false);
    double Us[V];
#pragma unroll
    for (int v = 0; v < V; v++) Us[v] = u[v * nS + t];
    const Prim<D> p = primitives<D>(Us, a.gamma);
    double tF[D][V];
#pragma unroll
    for (int d1 = 0; d1 < D; d1++) conv_proj<D>(Us, p, ij[d1], tF[d1]);
    if (VISC) {
      double g[V][D];
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) g[v][d] = g_l[(d * V + v) * nS + t];
      const Visc<D> vs = viscous<D>(Us, p, g, a);
#pragma unroll
      for (int d1 = 0; d1 < D; d1++) visc_proj_sub<D>(vs, ij[d1], tF[d1]);
    }
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = tF[d][v];
  }
  __syncthreads();
  double unew[V];
  if (valid) {
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
    rk_epilogue<D>(a, e, t, nS, acc, unew);
  }
  // the new state's full trace for the next stage, into a buffer other than the Utr this
  // launch reads; the element's U block is no longer read
  if (a.rk.mode == 1 && a.rk.trace) rk_trace<D, N>(a, lops, lU + el * V * nS, e, t, valid, unew, -1);
}

template <int D, int N, bool VISC, int BM, int FC, bool SELF = !VISC>
__global__ __launch_bounds__(kMaxThreads) void k_divergence(const double *__restrict__ U, double *__restrict__ divF,
                                                            KArgs a) {
  static_assert(!(VISC && SELF), "the viscous divergence reads the gradient traces, hence the traces too");
  divergence_block<D, N, VISC, BM, FC, SELF>(U, divF, a, blockIdx.x);
}

template <int D, int N, bool VISC, int BM, int FC>
__global__ __launch_bounds__(kMaxThreads) void k_divergence_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  divergence_block<D, N, VISC, BM, FC, !VISC>(cs[c].U, cs[c].out, cs[c].a, (long)blockIdx.x - starts[c]);
}

// ---------------------------------------------------------------- one-sided LDG path
// With LDG beta = +1/2 (BM 1) the common viscous flux uses only the left (slot-0)
// side's gradient, with beta = -1/2 (BM 2) only the right side's (src/faces.cpp:1520-1545),
// and the common solution is the other side's state.  The element on that "owner" side
// therefore has everything for the complete common flux (Rusanov + LDG) right after its
// own corrected gradient: it writes the unscaled common normal flux to Ftr at its owned
// flux points, and no gradient trace is needed.  The partner's solution trace at an owned
// point is extrapolated from the partner's U inside the kernel.  A residual is then
//   k_visc_onesided (gradient, volume flux divergence, owned common fluxes)
//   -> k_face_correction (oppDiv_fpts Fcomm, both sides).

// Register budget: for hex p=3 four waves per SIMD (<= 128 VGPRs) is what the kernel's
// LDS allows per CU and the Rusanov face phase fits it without spilling, so the allocator
// is held to it.  The Roe face phase needs more: three waves (<= 168 VGPRs) measured
// fastest (four spill, 2.76 ms; three 1.52 ms; uncapped, two waves, 1.80 ms at 64^3).
// Other shapes would spill under a cap and keep the default (quads at p=4: three waves,
// 168 VGPRs with 15 spilled, measured 0.165 ms against 0.128 ms uncapped at 256^2).
template <int D, int N, int FC>
struct OneSidedWaves {
  static constexpr int value = (D == 3 && N == 4) ? (FC == 1 ? 3 : 4) : 1;
};

// SELF: extrapolate the partner traces this element needs from the partner's U (no trace
// launch before this kernel; they are parked in Utr at this element's own owned points
// for the face phase); otherwise read them from Utr at the partner's points (written by a
// trace launch or by the previous RK stage's update).  A template parameter: as a
// run-time branch the two prologues together spill the hex p=3 kernel.
template <int D, int N, int BM, int FC, bool SELF>
__device__ __forceinline__ void visc_onesided_block(const double *__restrict__ U, double *__restrict__ divF,
                                                    const KArgs &a, long bid) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  constexpr int owner = (BM == 1) ? 0 : 1;  // slot whose gradient enters the common flux
  // one LDS region per element holds the common solution, then (after a barrier) the
  // gradient, then the transformed flux
  constexpr int W = (V * nF > D * V * nS) ? V * nF : D * V * nS;
  extern __shared__ double lds[];
  const int EB = a.EB;
  double *lops = lds;
  double *lU = lops + S::NOPS_PAD;
  double *lW = lU + EB * V * nS;
  unsigned *lmask = reinterpret_cast<unsigned *>(lW + EB * W);  // owned faces per element
  const long e0 = bid * EB;
  const long neb = min((long)EB, (long)a.E - e0);
  const int tid = threadIdx.x, el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  // the slots of this thread's flux points and, where this element owns the face, the
  // partner's trace: with SELF extrapolated here from the partner's solution line
  // (no trace launch runs before this kernel), else read from the partner's Utr.  Issued
  // before the staging barrier so the loads overlap it.
  int slot[S::FPT];
  double cp[S::FPT][V];
  constexpr bool self = SELF;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    slot[k] = -1;
    if (valid && f < nF) {
      long ep; int sl, fp;
      unpack_nbr(a.nbr[e * nF + f], ep, sl, fp);
      slot[k] = sl;
      if (sl == owner) {
        if (self) {
          partner_trace<D, N>(U, a.ops, ep, fp, cp[k]);
        } else {
#pragma unroll
          for (int v = 0; v < V; v++) cp[k][v] = a.Utr[(ep * V + v) * nF + fp];
        }
      }
    }
  }
  stage_U<D, N>(U, a, lops, lU, e0, neb, V * nS);
  if (threadIdx.x < EB) lmask[threadIdx.x] = 0u;
  __syncthreads();
  const double *u = lU + el * V * nS;
  double *c = lW + el * W;
  double *g_l = lW + el * W;
  // common solution (= the non-owner side's state) and the mask of owned faces (all flux
  // points of a face share its slot, so ownership is per face)
  if (valid) {
#pragma unroll
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      if (slot[k] == owner) {
#pragma unroll
        for (int v = 0; v < V; v++) {
          c[v * nF + f] = cp[k][v];
          if (self) a.Utr[(e * V + v) * nF + f] = cp[k][v];  // own slot: read back in the face phase
        }
        if (f % S::nFF == 0) atomicOr(&lmask[el], 1u << (f / S::nFF));
      } else {
        int dim, side, base, str;
        fpt_info<D, N>(f, dim, side, base, str);
        const double *lE = tab_lE<N>(lops, side);
#pragma unroll
        for (int v = 0; v < V; v++) c[v * nF + f] = line_extrap<N>(lE, u + v * nS, base, str);
      }
    }
  }
  __syncthreads();
  {
    double g[V][D];
    if (valid) {
      double ij[D][D], detJ;
      load_metric<D>(a, e, t, nS, ij, detJ, true);
      spt_gradient<D, N>(lops, u, c, t, ij, detJ, g);
    }
    __syncthreads();  // the common solution is consumed
    if (valid) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = g[v][d];
    }
  }
  __syncthreads();
  // common normal flux (unscaled) at the owned flux points -> Ftr
  if (valid) {
    const unsigned mask = lmask[el];
    const int cnt = __popc(mask) * S::nFF;
#pragma unroll 1
    for (int k = t; k < cnt; k += nS) {
      const int f = owned_fpt<D, N>(mask, k);  // face-major: neighbouring lanes, neighbouring points
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
      const double *fg = a.fgeo + e * (D + 1) * nF + f;
      double n[D];
#pragma unroll
      for (int d = 0; d < D; d++) n[d] = fg[d * nF];
      double Uo[V], Up[V];
#pragma unroll
      for (int v = 0; v < V; v++) Uo[v] = line_extrap<N>(lE, u + v * nS, base, str);
      if constexpr (FC == 1) {
        // Roe: the convective flux first, so that its many temporaries are dead before the
        // gradient's lines are gathered (formed in the Rusanov order it needs ~350 registers)
        {
          long ep; int sl, fp;
          unpack_nbr(a.nbr[e * nF + f], ep, sl, fp);
          const long ip = self ? (e * V * nF + f) : (ep * V * nF + fp);
#pragma unroll
          for (int v = 0; v < V; v++) Up[v] = a.Utr[ip + v * nF];
        }
        double Fo[V], G[V];
        if (owner == 0) conv_common_flux<D, FC>(Uo, Up, n, a, Fo);
        else conv_common_flux<D, FC>(Up, Uo, n, a, Fo);
        asm volatile("" ::: "memory");
        visc_normal<D>(Uo, [&](int j, double *dUj) {
#pragma unroll
          for (int v = 0; v < V; v++) {
            dUj[v] = line_extrap<N>(lE, g_l + (j * V + v) * nS, base, str);
            asm volatile("" ::: "memory");
          }
        }, n, a, G);
        if (owner == 0) ldg_common_add<D, BM>(Uo, Up, G, G, a, Fo);
        else ldg_common_add<D, BM>(Up, Uo, G, G, a, Fo);
#pragma unroll
        for (int v = 0; v < V; v++) a.Ftr[(e * V + v) * nF + f] = Fo[v];
        continue;
      }
      double G[V];
      // one variable's line at a time (a compiler barrier after each): 128 VGPRs instead
      // of 160, i.e. one more wave per SIMD, worth more than the loads' overlap
      visc_normal<D>(Uo, [&](int j, double *dUj) {
#pragma unroll
        for (int v = 0; v < V; v++) {
          dUj[v] = line_extrap<N>(lE, g_l + (j * V + v) * nS, base, str);
          asm volatile("" ::: "memory");
        }
      }, n, a, G);
      asm volatile("" ::: "memory");  // load the partner's state only now: fewer live registers
      {
        long ep; int sl, fp;
        unpack_nbr(a.nbr[e * nF + f], ep, sl, fp);
        const long ip = self ? (e * V * nF + f) : (ep * V * nF + fp);  // cached since phase 1
#pragma unroll
        for (int v = 0; v < V; v++) Up[v] = a.Utr[ip + v * nF];
      }
      double Fo[V];
      if (owner == 0) common_flux<D, true, BM, FC>(Uo, Up, G, G, n, 1.0, a, Fo);
      else common_flux<D, true, BM, FC>(Up, Uo, G, G, n, 1.0, a, Fo);
#pragma unroll
      for (int v = 0; v < V; v++) a.Ftr[(e * V + v) * nF + f] = Fo[v];
    }
  }
  __syncthreads();  // the gradient field is consumed
  // pointwise flux contracted with the adjugate rows, over the own spt column of lG
  if (valid) {
    double ij[D][D], detJ;
    load_metric<D>(a, e, t, nS, ij, detJ, false);
    double Us[V], g[V][D];
#pragma unroll
    for (int v = 0; v < V; v++) Us[v] = u[v * nS + t];
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < V; v++) g[v][d] = g_l[(d * V + v) * nS + t];
    const Prim<D> p = primitives<D>(Us, a.gamma);
    const Visc<D> vs = viscous<D>(Us, p, g, a);
    double tF[D][V];
#pragma unroll
    for (int d1 = 0; d1 < D; d1++) {
      conv_proj<D>(Us, p, ij[d1], tF[d1]);
      visc_proj_sub<D>(vs, ij[d1], tF[d1]);
    }
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = tF[d][v];
  }
  __syncthreads();
  if (!valid) return;
  // volume part of the divergence: oppDiv tF (src/elements.cpp:1449-1474)
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
  for (int v = 0; v < V; v++) divF[(e * V + v) * nS + t] = acc[v];
}

template <int D, int N, int BM, int FC, bool SELF>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(OneSidedWaves<D, N, FC>::value)))
void k_visc_onesided(const double *__restrict__ U, double *__restrict__ divF, KArgs a) {
  visc_onesided_block<D, N, BM, FC, SELF>(U, divF, a, blockIdx.x);
}

// The same kernel for workgroups of one wave (64 threads: one hex p=3 element per
// workgroup, the tile the autotuner picks there): the compiler knows the workgroup is a
// single wave, so its barriers need no cross-wave synchronisation.
template <int D, int N, int BM, int FC, bool SELF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OneSidedWaves<D, N, FC>::value)))
void k_visc_onesided_w64(const double *__restrict__ U, double *__restrict__ divF, KArgs a) {
  visc_onesided_block<D, N, BM, FC, SELF>(U, divF, a, blockIdx.x);
}

template <int D, int N, int BM, int FC>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(OneSidedWaves<D, N, FC>::value)))
void k_visc_onesided_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  visc_onesided_block<D, N, BM, FC, true>(cs[c].U, cs[c].out, cs[c].a, (long)blockIdx.x - starts[c]);
}


// divF += oppDiv_fpts Fcomm (src/elements.cpp:1476-1501) with Fcomm = Ftr(owner) * signed
// dA of this side; the owner's Ftr is this element's own or its partner's.
template <int D, int N, int BM>
__device__ __forceinline__ void face_correction_block(double *__restrict__ divF, const KArgs &a, long bid) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  constexpr int owner = (BM == 1) ? 0 : 1;
  // per-element LDS block: the common fluxes (V * nF), then the new state (V * nS) when the
  // RK update writes the next trace
  constexpr int W = V * (nF > nS ? nF : nS);
  extern __shared__ double lds[];
  const int EB = a.EB;
  double *lops = lds;
  double *lF = lops + S::NOPS_PAD;
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int i = tid; i < S::NOPS; i += NT) lops[i] = a.ops[i];
  const long e0 = bid * EB;
  const long neb = min((long)EB, (long)a.E - e0);
  const int el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  double *c = lF + el * W;
  if (valid) {
#pragma unroll
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      long ep; int sl, fp;
      unpack_nbr(a.nbr[e * nF + f], ep, sl, fp);
      const double dAs = a.fgeo[(e * (D + 1) + D) * nF + f];
      const double *src = (sl == owner) ? a.Ftr + e * V * nF + f : a.Ftr + ep * V * nF + fp;
#pragma unroll
      for (int v = 0; v < V; v++) c[v * nF + f] = src[v * nF] * dAs;
    }
  }
  __syncthreads();
  double unew[V];
  if (valid) {
    const int ia = t % N, ib = (t / N) % N, ic = (D == 3) ? t / (N * N) : 0;
    double acc[V];
#pragma unroll
    for (int v = 0; v < V; v++) acc[v] = divF[(e * V + v) * nS + t];
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
    rk_epilogue<D>(a, e, t, nS, acc, unew);
  }
  // the new state's partner-owned trace for the next stage
  if (a.rk.mode == 1 && a.rk.trace) rk_trace<D, N>(a, lops, c, e, t, valid, unew, owner);
}

template <int D, int N, int BM>
__global__ __launch_bounds__(kMaxThreads) void k_face_correction(double *__restrict__ divF, KArgs a) {
  face_correction_block<D, N, BM>(divF, a, blockIdx.x);
}

template <int D, int N, int BM>
__global__ __launch_bounds__(kMaxThreads) void k_face_correction_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  face_correction_block<D, N, BM>(cs[c].out, cs[c].a, (long)blockIdx.x - starts[c]);
}

// ---------------------------------------------------------------- reverse mode
#include "adjoint.hpp"

// ---------------------------------------------------------------- RK + norm
// src/solver.cpp:2897-2941 (stage), 2997-3062 (final), element-major layout.  src, when
// not null, is the p-multigrid source term added to each divF_s (RK_update_source,
// src/solver_kernels.cu:253-303, with the CPU path's arithmetic, src/solver.cpp:2946-2962,
// 3061-3073).
template <int D>
__global__ __launch_bounds__(kMaxThreads) void k_rk_stage(double *__restrict__ U, const double *__restrict__ Uini,
                                                          const double *__restrict__ divF,
                                                          const double *__restrict__ src,
                                                          const double *__restrict__ metric, long total, int V,
                                                          int nS, double alpha, double dt) {
  constexpr int nM = D * D + 1;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long e = i / (V * nS);
    const int s = (int)(i % nS);
    const double detJ = metric[(e * nM + D * D) * nS + s];
    const double r = src ? divF[i] + src[i] : divF[i];
    U[i] = Uini[i] - alpha * dt / detJ * r;
  }
}

template <int D>
__global__ __launch_bounds__(kMaxThreads) void k_rk_final(double *U, const double *Uini,  // may alias (Euler)
                                                          const double *__restrict__ divF,
                                                          const double *__restrict__ src,
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
    for (int st = 0; st < nStages; st++) {
      const double r = src ? divF[st * total + i] + src[i] : divF[st * total + i];
      u -= b[st] * dt / detJ * r;
    }
    U[i] = u;
  }
}

// ---------------------------------------------------------------- p-multigrid
// Transfer between orders (restrict_pmg, prolong_err, prolong_U; src/multigrid.cpp:335-472).
// The reference's operator [nSpts_out][nSpts_in] is a tensor product of one 1D table A
// [n_out][n_in] with itself (src/quads.cpp:363-465, src/hexas.cpp:586-704), so per element
// and variable ("line") out = alpha (A x A [x A]) in + beta out is applied sum-factorised:
// along x, then y, then z, with points x fastest.  A workgroup takes L consecutive lines
// (one contiguous chunk of in and of out), staged in LDS and ping-ponged between passes.
constexpr int kPmgBuf = 2048;  // doubles per LDS buffer: L * max(n_in, n_out)^D <= kPmgBuf
constexpr int kPmgLds = 64 + 2 * kPmgBuf;

template <int D>
__global__ __launch_bounds__(kMaxThreads) void k_pmg_transfer(const double *__restrict__ in, double *__restrict__ out,
                                                              const double *__restrict__ A, long lines, int ni,
                                                              int no, int L, double alpha, double beta) {
  extern __shared__ double lds[];  // kPmgLds doubles
  double *lA = lds, *buf = lds + 64;
  const long l0 = (long)blockIdx.x * L;
  const int nl = (int)min((long)L, lines - l0);
  int szi = 1, szo = 1;
  for (int d = 0; d < D; d++) {
    szi *= ni;
    szo *= no;
  }
  for (int k = threadIdx.x; k < no * ni; k += blockDim.x) lA[k] = A[k];
  const double *gin = in + l0 * szi;
  for (int k = threadIdx.x; k < nl * szi; k += blockDim.x) buf[k] = gin[k];
  __syncthreads();
  double *src = buf, *dst = buf + kPmgBuf;
  int inner = 1, outer = szi / ni;
  for (int d = 0; d < D; d++) {
    // src [nl][outer][ni][inner] -> dst [nl][outer][no][inner]
    const int n = nl * outer * no * inner;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const int q = k % inner, o = (k / inner) % no, lo = k / (inner * no);
      const double *s = src + lo * ni * inner + q;
      double acc = 0.0;
      for (int i = 0; i < ni; i++) acc += lA[o * ni + i] * s[i * inner];
      if (d == D - 1) {  // outer == 1: k runs over [nl][no^D] in output order
        double *g = out + l0 * szo + k;
        *g = (beta == 0.0) ? alpha * acc : alpha * acc + beta * *g;
      } else {
        dst[k] = acc;
      }
    }
    if (d < D - 1) {
      __syncthreads();
      double *t = src;
      src = dst;
      dst = t;
      inner *= no;
      outer /= ni;
    }
  }
}

// z = a x + b y elementwise (device_copy / device_add / device_subtract,
// src/aux_kernels.cu:143-205, as one kernel).
__global__ __launch_bounds__(kMaxThreads) void k_axpby(double *z, double a, const double *x, double b,
                                                       const double *y, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    z[i] = a * x[i] + b * y[i];
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

// Stream copy used as the device's measured bandwidth reference: each thread moves
// kCopyUnroll 16-byte vectors, all loads issued before the stores.
constexpr int kCopyUnroll = 4;
__global__ __launch_bounds__(kMaxThreads) void k_copy(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                                      long n2) {
  const long base = (long)blockIdx.x * blockDim.x * kCopyUnroll + threadIdx.x;
  double2 v[kCopyUnroll];
#pragma unroll
  for (int u = 0; u < kCopyUnroll; u++) {
    const long i = base + (long)u * blockDim.x;
    if (i < n2) v[u] = src[i];
  }
#pragma unroll
  for (int u = 0; u < kCopyUnroll; u++) {
    const long i = base + (long)u * blockDim.x;
    if (i < n2) dst[i] = v[u];
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
  if (d->eles_per_block <= 0) {
    // small meshes (coarse multigrid levels, small cases): keep >= 1024 workgroups, four
    // per CU, so that every SIMD holds a wave; the kernels are latency-bound there, and
    // partly filled waves cost less than idle SIMDs
    const long per_cu4 = ((long)d->nEles + 1023) / 1024;
    if (per_cu4 < eb) eb = per_cu4 < 1 ? 1 : (int)per_cu4;
  }
  while (eb > 1 && (eb * d->nSpts > kMaxThreads || sizeof(double) * (fixed + eb * per_ele) > kMaxLds)) eb--;
  return eb < 1 ? 1 : eb;
}

inline KArgs make_args(const zfr_ele_desc_t *d, int EB) {
  KArgs a;
  a.E = d->nEles; a.EB = EB; a.viscous = d->viscous; a.fix_vis = d->fix_vis;
  a.gamma = d->gamma; a.prandtl = d->prandtl; a.mu = d->mu; a.rt = d->rt; a.c_sth = d->c_sth;
  a.rus_k = d->rus_k; a.ldg_b = d->ldg_b; a.ldg_tau = d->ldg_tau;
  a.ops = d->ops1d; a.metric = d->metric; a.fgeo = d->fgeo; a.nbr = d->nbr; a.Utr = d->Utr; a.Gtr = d->Gtr;
  a.Ftr = d->Ftr;
  a.trace_sel = 0;
  a.rk = KArgs::Rk{};
  return a;
}

enum class Op { Extrap, ExtrapPartner, Grad, Div, OneSided, OneSidedFromTrace, FaceCorr, AdjPrep, AdjMain, AdjGather };

// k_divergence for one convective flux (FC: 0 Rusanov, 1 Roe) and the LDG beta branch.
// Euler: the standalone residual extrapolates the partner traces itself (SELF); with a
// fused RK update it must read them from Utr, since its epilogue overwrites U in place
// while other workgroups may still read it.
template <int D, int N, int FC>
void launch_div(const zfr_ele_desc_t *d, dim3 grid, dim3 block, size_t lds, hipStream_t st, const double *U,
                double *out, const KArgs &a) {
  if (!d->viscous && a.rk.mode != 0)
    hipLaunchKernelGGL((k_divergence<D, N, false, 0, FC, false>), grid, block, lds, st, U, out, a);
  else if (!d->viscous) hipLaunchKernelGGL((k_divergence<D, N, false, 0, FC, true>), grid, block, lds, st, U, out, a);
  else if (d->ldg_b == 0.5) hipLaunchKernelGGL((k_divergence<D, N, true, 1, FC>), grid, block, lds, st, U, out, a);
  else if (d->ldg_b == -0.5) hipLaunchKernelGGL((k_divergence<D, N, true, 2, FC>), grid, block, lds, st, U, out, a);
  else hipLaunchKernelGGL((k_divergence<D, N, true, 0, FC>), grid, block, lds, st, U, out, a);
}

// Extra operands of the reverse-mode launches.
struct AdjPtrs {
  const double *w = nullptr;  // cotangent of divF
  double *Wtr = nullptr;      // [E][V][nF] face weights of w
  double *Xtr = nullptr;      // [E][V][nF] partner trace cotangent at owned points
  const KArgs::Rk *rk = nullptr;  // fused RK update for Div / FaceCorr
};

// k_adj_main for one convective flux (FC: 0 Rusanov, 1 Roe) and the LDG beta branch.
template <int D, int N, int FC>
void launch_adj_main(const zfr_ele_desc_t *d, dim3 grid, dim3 block, size_t lds, hipStream_t st, const double *U,
                     const AdjPtrs &j, double *out, const KArgs &a) {
  if (!d->viscous)
    hipLaunchKernelGGL((k_adj_main<D, N, false, 0, FC>), grid, block, lds, st, U, j.w, j.Wtr, j.Xtr, out, a);
  else if (block.x <= 64) {
    if (d->ldg_b == 0.5)
      hipLaunchKernelGGL((k_adj_main_w64<D, N, true, 1, FC>), grid, block, lds, st, U, j.w, j.Wtr, j.Xtr, out, a);
    else
      hipLaunchKernelGGL((k_adj_main_w64<D, N, true, 2, FC>), grid, block, lds, st, U, j.w, j.Wtr, j.Xtr, out, a);
  } else if (d->ldg_b == 0.5)
    hipLaunchKernelGGL((k_adj_main<D, N, true, 1, FC>), grid, block, lds, st, U, j.w, j.Wtr, j.Xtr, out, a);
  else
    hipLaunchKernelGGL((k_adj_main<D, N, true, 2, FC>), grid, block, lds, st, U, j.w, j.Wtr, j.Xtr, out, a);
}

// LDS doubles per element of each kernel (the owned-point list of the one-sided kernel is
// nF + 1 ints).
template <int D, int N>
size_t lds_per_element(Op op, bool viscous) {
  using S = Shape<D, N>;
  constexpr size_t kUC = (size_t)S::V * S::nS + S::V * S::nF, kG = (size_t)D * S::V * S::nS;
  size_t per_ele = 0;
  switch (op) {
    case Op::Extrap:
    case Op::ExtrapPartner: per_ele = (size_t)S::V * S::nS; break;
    case Op::Grad: per_ele = kUC > kG ? kUC : kG; break;
    case Op::Div: per_ele = kUC + kG; break;
    case Op::OneSided:
    case Op::OneSidedFromTrace:
      per_ele = (size_t)S::V * S::nS + (S::V * S::nF > (int)kG ? S::V * S::nF : kG) + 1;
      break;
    case Op::FaceCorr: per_ele = (size_t)S::V * (S::nF > S::nS ? S::nF : S::nS); break;
    case Op::AdjPrep: per_ele = (size_t)S::V * S::nS; break;
    case Op::AdjMain:
      per_ele = viscous ? (size_t)AdjLds<D, N, true>::per_ele : (size_t)AdjLds<D, N, false>::per_ele;
      break;
    case Op::AdjGather: per_ele = (size_t)S::V * S::nF; break;
  }
  return per_ele;
}

template <int D, int N>
int launch(Op op, const zfr_ele_desc_t *d, const double *U, double *out, hipStream_t st, const AdjPtrs &j = AdjPtrs()) {
  using S = Shape<D, N>;
  if (d->nSpts != S::nS || d->nFpts != S::nF || d->nVars != S::V)
    return zfr::fail("kernel launch: element descriptor does not match (nDims, order)");
  if (d->nEles <= 0) return 0;
  const size_t per_ele = lds_per_element<D, N>(op, d->viscous != 0);
  const bool adj = op == Op::AdjPrep || op == Op::AdjMain || op == Op::AdjGather;
  if (adj && d->viscous && d->ldg_b != 0.5 && d->ldg_b != -0.5)
    return zfr::fail("the residual VJP supports LDG beta = +-0.5 (and Euler)");
  if (d->fconv_type != 0 && d->fconv_type != 2) return zfr::fail("fconv_type must be 0 (Rusanov) or 2 (Roe)");
  if ((op == Op::OneSided || op == Op::OneSidedFromTrace || op == Op::FaceCorr) && d->ldg_b != 0.5 &&
      d->ldg_b != -0.5)
    return zfr::fail("one-sided LDG kernels need ldg_b = +-0.5");
  // the viscous adjoint runs one element per workgroup: its LDS then allows ~7
  // workgroups per CU, more than any larger tile
  const int EB = (op == Op::AdjMain && d->viscous) ? 1 : eles_per_block(d, S::defaultEB(), S::NOPS_PAD, per_ele);
  if (sizeof(double) * (S::NOPS_PAD + EB * per_ele) > kMaxLds)
    return zfr::fail("kernel launch: one element needs more LDS than a workgroup may use (order too high)");
  KArgs a = make_args(d, EB);
  if (j.rk) {
    if (op != Op::Div && op != Op::FaceCorr) return zfr::fail("kernel launch: RK fusion is for the last kernel only");
    a.rk = *j.rk;
  }
  const dim3 grid((unsigned)((d->nEles + EB - 1) / EB)), block((unsigned)(EB * S::nS));
  const size_t lds = sizeof(double) * (S::NOPS_PAD + EB * per_ele);
  switch (op) {
    case Op::Extrap:
    case Op::ExtrapPartner:
      a.trace_sel = (op == Op::ExtrapPartner) ? 1 : 0;
      if (block.x <= 64) hipLaunchKernelGGL((k_extrapolate_w64<D, N>), grid, block, lds, st, U, a);
      else hipLaunchKernelGGL((k_extrapolate<D, N>), grid, block, lds, st, U, a);
      break;
    case Op::Grad:
      hipLaunchKernelGGL((k_gradient<D, N>), grid, block, lds, st, U, out, a);
      break;
    case Op::Div:
      if (d->fconv_type == 2) launch_div<D, N, 1>(d, grid, block, lds, st, U, out, a);
      else launch_div<D, N, 0>(d, grid, block, lds, st, U, out, a);
      break;
    case Op::OneSided:
    case Op::OneSidedFromTrace: {
#define ZK_OS(K_, B_, F_, S_) hipLaunchKernelGGL((K_<D, N, B_, F_, S_>), grid, block, lds, st, U, out, a)
#define ZK_OS2(K_, B_, F_) do { if (op == Op::OneSided) ZK_OS(K_, B_, F_, true); else ZK_OS(K_, B_, F_, false); } while (0)
      if (block.x <= 64) {
        if (d->fconv_type == 2) { if (d->ldg_b == 0.5) ZK_OS2(k_visc_onesided_w64, 1, 1); else ZK_OS2(k_visc_onesided_w64, 2, 1); }
        else { if (d->ldg_b == 0.5) ZK_OS2(k_visc_onesided_w64, 1, 0); else ZK_OS2(k_visc_onesided_w64, 2, 0); }
      } else {
        if (d->fconv_type == 2) { if (d->ldg_b == 0.5) ZK_OS2(k_visc_onesided, 1, 1); else ZK_OS2(k_visc_onesided, 2, 1); }
        else { if (d->ldg_b == 0.5) ZK_OS2(k_visc_onesided, 1, 0); else ZK_OS2(k_visc_onesided, 2, 0); }
      }
#undef ZK_OS2
#undef ZK_OS
      break;
    }
    case Op::FaceCorr:
      if (d->ldg_b == 0.5) hipLaunchKernelGGL((k_face_correction<D, N, 1>), grid, block, lds, st, out, a);
      else hipLaunchKernelGGL((k_face_correction<D, N, 2>), grid, block, lds, st, out, a);
      break;
    case Op::AdjPrep:
      hipLaunchKernelGGL((k_adj_prep<D, N>), grid, block, lds, st, j.w, j.Wtr, a);
      break;
    case Op::AdjMain:
      if (d->fconv_type == 2) launch_adj_main<D, N, 1>(d, grid, block, lds, st, U, j, out, a);
      else launch_adj_main<D, N, 0>(d, grid, block, lds, st, U, j, out, a);
      break;
    case Op::AdjGather:
      if (d->viscous && d->ldg_b == -0.5) hipLaunchKernelGGL((k_adj_gather<D, N, 2>), grid, block, lds, st, j.Xtr, out, a);
      else hipLaunchKernelGGL((k_adj_gather<D, N, 0>), grid, block, lds, st, j.Xtr, out, a);
      break;
  }
  return zfr::check_hip(hipGetLastError(), "kernel launch");
}

int dispatch(Op op, const zfr_ele_desc_t *d, const double *U, double *out, hipStream_t st, const AdjPtrs &j = AdjPtrs()) {
  const int N = d->order + 1;
  if (d->nDims == 2) {
    switch (N) {
      case 1: return launch<2, 1>(op, d, U, out, st, j);
      case 2: return launch<2, 2>(op, d, U, out, st, j);
      case 3: return launch<2, 3>(op, d, U, out, st, j);
      case 4: return launch<2, 4>(op, d, U, out, st, j);
      case 5: return launch<2, 5>(op, d, U, out, st, j);
      case 6: return launch<2, 6>(op, d, U, out, st, j);
    }
  } else if (d->nDims == 3) {
    switch (N) {
      case 1: return launch<3, 1>(op, d, U, out, st, j);
      case 2: return launch<3, 2>(op, d, U, out, st, j);
      case 3: return launch<3, 3>(op, d, U, out, st, j);
      case 4: return launch<3, 4>(op, d, U, out, st, j);
      case 5: return launch<3, 5>(op, d, U, out, st, j);
      case 6: return launch<3, 6>(op, d, U, out, st, j);
    }
  }
  return zfr::fail("kernel launch: supported shapes are quads and hexes with order 0..5");
}

// One launch of a multi-case plan: every case of one kernel, shape and flux branch.
struct MultiGroup {
  Op op = Op::Extrap;
  int D = 0, N = 0, visc = 0, bm = 0, fc = 0;  // bm: LDG branch (0 general, 1 beta = 1/2, 2 beta = -1/2)
  int EB = 1, nc = 0;
  unsigned nblk = 0;
  size_t lds = 0;
  MultiCase *cases = nullptr;  // device [nc]
  long *starts = nullptr;      // device [nc]
};

template <int D, int N>
void launch_multi(const MultiGroup &g, hipStream_t st) {
  const dim3 grid(g.nblk), block((unsigned)(g.EB * Shape<D, N>::nS));
  const MultiCase *c = g.cases;
  const long *s0 = g.starts;
  switch (g.op) {
    case Op::Extrap:
    case Op::ExtrapPartner:
      hipLaunchKernelGGL((k_extrapolate_multi<D, N>), grid, block, g.lds, st, c, s0, g.nc);
      break;
    case Op::Grad:
      hipLaunchKernelGGL((k_gradient_multi<D, N>), grid, block, g.lds, st, c, s0, g.nc);
      break;
    case Op::Div:
#define ZK_DIV(V_, B_, F_) hipLaunchKernelGGL((k_divergence_multi<D, N, V_, B_, F_>), grid, block, g.lds, st, c, s0, g.nc)
      if (!g.visc) { if (g.fc) ZK_DIV(false, 0, 1); else ZK_DIV(false, 0, 0); }
      else if (g.bm == 1) { if (g.fc) ZK_DIV(true, 1, 1); else ZK_DIV(true, 1, 0); }
      else if (g.bm == 2) { if (g.fc) ZK_DIV(true, 2, 1); else ZK_DIV(true, 2, 0); }
      else { if (g.fc) ZK_DIV(true, 0, 1); else ZK_DIV(true, 0, 0); }
#undef ZK_DIV
      break;
    case Op::OneSided:
#define ZK_OSM(B_, F_) hipLaunchKernelGGL((k_visc_onesided_multi<D, N, B_, F_>), grid, block, g.lds, st, c, s0, g.nc)
      if (g.bm == 1) { if (g.fc) ZK_OSM(1, 1); else ZK_OSM(1, 0); }
      else { if (g.fc) ZK_OSM(2, 1); else ZK_OSM(2, 0); }
#undef ZK_OSM
      break;
    case Op::FaceCorr:
      if (g.bm == 1) hipLaunchKernelGGL((k_face_correction_multi<D, N, 1>), grid, block, g.lds, st, c, s0, g.nc);
      else hipLaunchKernelGGL((k_face_correction_multi<D, N, 2>), grid, block, g.lds, st, c, s0, g.nc);
      break;
    default:
      break;
  }
}

template <int D>
void dispatch_multi_n(const MultiGroup &g, hipStream_t st) {
  switch (g.N) {
    case 1: launch_multi<D, 1>(g, st); break;
    case 2: launch_multi<D, 2>(g, st); break;
    case 3: launch_multi<D, 3>(g, st); break;
    case 4: launch_multi<D, 4>(g, st); break;
    case 5: launch_multi<D, 5>(g, st); break;
    case 6: launch_multi<D, 6>(g, st); break;
  }
}

template <int D>
size_t lds_per_element_n(int N, Op op, bool viscous) {
  switch (N) {
    case 1: return lds_per_element<D, 1>(op, viscous);
    case 2: return lds_per_element<D, 2>(op, viscous);
    case 3: return lds_per_element<D, 3>(op, viscous);
    case 4: return lds_per_element<D, 4>(op, viscous);
    case 5: return lds_per_element<D, 5>(op, viscous);
    case 6: return lds_per_element<D, 6>(op, viscous);
  }
  return 0;
}

inline unsigned grid_for(long total) {
  long g = (total + kMaxThreads - 1) / kMaxThreads;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace zk

extern "C" {

namespace {
// Buffers a launch reads or writes through the descriptor: a null one is an error here,
// not a fault on the device.
int need(const void *p, const char *what) {
  return p ? 0 : zfr::fail(std::string("kernel launch: the descriptor's ") + what + " is NULL");
}
int need_geometry(const zfr_ele_desc_t *d) {
  if (!d) return zfr::fail("kernel launch: NULL descriptor");
  int rc = need(d->ops1d, "ops1d");
  if (!rc) rc = need(d->metric, "metric");
  if (!rc) rc = need(d->fgeo, "fgeo");
  if (!rc) rc = need(d->nbr, "nbr");
  return rc;
}
}  // namespace

int zfr_extrapolate(const zfr_ele_desc_t *d, const double *U, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need(d->Utr, "Utr");
  if (rc) return rc;
  return zk::dispatch(zk::Op::Extrap, d, U, nullptr, (hipStream_t)stream);
}

int zfr_extrapolate_partner(const zfr_ele_desc_t *d, const double *U, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need(d->Utr, "Utr");
  if (rc) return rc;
  if (!d->viscous || (d->ldg_b != 0.5 && d->ldg_b != -0.5))
    return zfr::fail("zfr_extrapolate_partner: the descriptor is not on the one-sided LDG path (beta = +-1/2)");
  return zk::dispatch(zk::Op::ExtrapPartner, d, U, nullptr, (hipStream_t)stream);
}

int zfr_gradient(const zfr_ele_desc_t *d, const double *U, double *dU_out, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need(d->Utr, "Utr");
  if (!rc) rc = need(d->Gtr, "Gtr");
  if (rc) return rc;
  if (!d->viscous) return zfr::fail("zfr_gradient: descriptor is inviscid");
  return zk::dispatch(zk::Op::Grad, d, U, dU_out, (hipStream_t)stream);
}

int zfr_divergence(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream) {
  int rc = need_geometry(d);
  if (!rc && d->viscous) rc = need(d->Utr, "Utr");
  if (!rc && d->viscous && d->ldg_b != 0.5 && d->ldg_b != -0.5) rc = need(d->Gtr, "Gtr");
  if (rc) return rc;
  return zk::dispatch(zk::Op::Div, d, U, divF, (hipStream_t)stream);
}

int zfr_gradient_flux(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need(d->Utr, "Utr");
  if (!rc) rc = need(d->Ftr, "Ftr");
  if (rc) return rc;
  if (!d->viscous) return zfr::fail("zfr_gradient_flux: descriptor is inviscid");
  return zk::dispatch(zk::Op::OneSided, d, U, divF, (hipStream_t)stream);
}

int zfr_gradient_flux_traced(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need(d->Utr, "Utr");
  if (!rc) rc = need(d->Ftr, "Ftr");
  if (rc) return rc;
  if (!d->viscous) return zfr::fail("zfr_gradient_flux_traced: descriptor is inviscid");
  return zk::dispatch(zk::Op::OneSidedFromTrace, d, U, divF, (hipStream_t)stream);
}

int zfr_face_correction(const zfr_ele_desc_t *d, double *divF, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need(d->Ftr, "Ftr");
  if (rc) return rc;
  if (!d->viscous) return zfr::fail("zfr_face_correction: descriptor is inviscid");
  return zk::dispatch(zk::Op::FaceCorr, d, nullptr, divF, (hipStream_t)stream);
}

namespace {
int rk_fused(zk::Op op, const zfr_ele_desc_t *d, const double *U, double *divF, const zfr_rk_args_t *r,
             hipStream_t st) {
  if (!r || !r->U || !r->U_ini || (r->mode == 2 && r->nStages > 1 && !r->divF0))
    return zfr::fail("RK fusion: missing buffer");
  int rc = need_geometry(d);
  if (!rc && op == zk::Op::Div) rc = need(d->Utr, "Utr");  // the fused divergence reads the trace
  if (!rc && op == zk::Op::Div && d->viscous && d->ldg_b != 0.5 && d->ldg_b != -0.5) rc = need(d->Gtr, "Gtr");
  if (!rc && op == zk::Op::FaceCorr) rc = need(d->Ftr, "Ftr");
  if (rc) return rc;
  if (r->mode != 1 && r->mode != 2) return zfr::fail("RK fusion: mode must be 1 (stage) or 2 (final)");
  if (r->nStages < 1 || r->nStages > 4 || r->stage < 0 || r->stage >= r->nStages)
    return zfr::fail("RK fusion: stage out of range");
  zk::KArgs::Rk rk{};
  rk.U = r->U;
  rk.Uini = r->U_ini;
  rk.divF0 = r->divF0;
  rk.src = r->source;
  rk.total = (long)d->nEles * d->nVars * d->nSpts;
  rk.alpha = r->alpha;
  rk.dt = r->dt;
  for (int s = 0; s < 4; s++) rk.beta[s] = r->beta[s];
  rk.nStages = r->nStages;
  rk.stage = r->stage;
  rk.mode = r->mode;
  if (r->trace_out && r->mode != 1) return zfr::fail("RK fusion: trace_out is for stage updates (mode 1)");
  if (r->trace_out && op == zk::Op::Div && r->trace_out == d->Utr)
    return zfr::fail("RK fusion: trace_out must not be the trace the divergence kernel reads (d->Utr)");
  rk.trace = r->trace_out;
  zk::AdjPtrs j;
  j.rk = &rk;
  return zk::dispatch(op, d, U, divF, st, j);
}
}  // namespace

int zfr_face_correction_rk(const zfr_ele_desc_t *d, double *divF, const zfr_rk_args_t *rk, void *stream) {
  if (!d->viscous) return zfr::fail("zfr_face_correction_rk: descriptor is inviscid");
  return rk_fused(zk::Op::FaceCorr, d, nullptr, divF, rk, (hipStream_t)stream);
}

int zfr_divergence_rk(const zfr_ele_desc_t *d, const double *U, double *divF, const zfr_rk_args_t *rk,
                      void *stream) {
  return rk_fused(zk::Op::Div, d, U, divF, rk, (hipStream_t)stream);
}

int zfr_rk_stage_source(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF,
                        const double *source, double alpha, double dt, void *stream) {
  const long total = (long)d->nEles * d->nVars * d->nSpts;
  if (total <= 0) return 0;
  if (d->nDims == 2)
    hipLaunchKernelGGL(zk::k_rk_stage<2>, dim3(zk::grid_for(total)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream,
                       U, U_ini, divF, source, d->metric, total, d->nVars, d->nSpts, alpha, dt);
  else
    hipLaunchKernelGGL(zk::k_rk_stage<3>, dim3(zk::grid_for(total)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream,
                       U, U_ini, divF, source, d->metric, total, d->nVars, d->nSpts, alpha, dt);
  return zfr::check_hip(hipGetLastError(), "zfr_rk_stage");
}

int zfr_rk_stage(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF, double alpha,
                 double dt, void *stream) {
  return zfr_rk_stage_source(d, U, U_ini, divF, nullptr, alpha, dt, stream);
}

int zfr_rk_final_source(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF,
                        const double *source, const double *beta, int nStages, double dt, void *stream) {
  if (nStages < 1 || nStages > 4) return zfr::fail("zfr_rk_final: nStages must be 1..4");
  const long total = (long)d->nEles * d->nVars * d->nSpts;
  if (total <= 0) return 0;
  double b[4] = {0, 0, 0, 0};
  for (int i = 0; i < nStages; i++) b[i] = beta[i];
  if (d->nDims == 2)
    hipLaunchKernelGGL(zk::k_rk_final<2>, dim3(zk::grid_for(total)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream,
                       U, U_ini, divF, source, d->metric, total, d->nVars, d->nSpts, b[0], b[1], b[2], b[3],
                       nStages, dt);
  else
    hipLaunchKernelGGL(zk::k_rk_final<3>, dim3(zk::grid_for(total)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream,
                       U, U_ini, divF, source, d->metric, total, d->nVars, d->nSpts, b[0], b[1], b[2], b[3],
                       nStages, dt);
  return zfr::check_hip(hipGetLastError(), "zfr_rk_final");
}

int zfr_rk_final(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF, const double *beta,
                 int nStages, double dt, void *stream) {
  return zfr_rk_final_source(d, U, U_ini, divF, nullptr, beta, nStages, dt, stream);
}

int zfr_pmg_transfer(int nDims, int nVars, long nEles, int n_in, int n_out, const double *A, double alpha,
                     const double *in, double beta, double *out, void *stream) {
  if (nDims != 2 && nDims != 3) return zfr::fail("zfr_pmg_transfer: nDims must be 2 or 3");
  if (n_in < 1 || n_in > 6 || n_out < 1 || n_out > 6)
    return zfr::fail("zfr_pmg_transfer: points per direction must be in 1..6 (orders 0..5)");
  if (nVars < 1 || nEles < 0) return zfr::fail("zfr_pmg_transfer: bad sizes");
  if (!A || !in || !out) return zfr::fail("zfr_pmg_transfer: missing buffer");
  const long lines = nEles * nVars;
  if (lines == 0) return 0;
  const int M = n_in > n_out ? n_in : n_out;
  int cube = 1;
  for (int d = 0; d < nDims; d++) cube *= M;
  int L = zk::kPmgBuf / cube;
  if (L > zk::kMaxThreads) L = zk::kMaxThreads;
  const long blocks = (lines + L - 1) / L;
  if (blocks > 0x7fffffffL) return zfr::fail("zfr_pmg_transfer: too many elements for one launch");
  if (nDims == 2)
    hipLaunchKernelGGL(zk::k_pmg_transfer<2>, dim3((unsigned)blocks), dim3(zk::kMaxThreads),
                       zk::kPmgLds * sizeof(double), (hipStream_t)stream,
                       in, out, A, lines, n_in, n_out, L, alpha, beta);
  else
    hipLaunchKernelGGL(zk::k_pmg_transfer<3>, dim3((unsigned)blocks), dim3(zk::kMaxThreads),
                       zk::kPmgLds * sizeof(double), (hipStream_t)stream,
                       in, out, A, lines, n_in, n_out, L, alpha, beta);
  return zfr::check_hip(hipGetLastError(), "zfr_pmg_transfer");
}

int zfr_axpby(long n, double a, const double *x, double b, const double *y, double *z, void *stream) {
  if (n <= 0) return 0;
  if (!x || !y || !z) return zfr::fail("zfr_axpby: missing buffer");
  hipLaunchKernelGGL(zk::k_axpby, dim3(zk::grid_for(n)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream, z, a, x, b,
                     y, n);
  return zfr::check_hip(hipGetLastError(), "zfr_axpby");
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

int zfr_residual_vjp(const zfr_ele_desc_t *d, const double *U, const double *w, double *v, double *Wtr,
                     double *Xtr, void *stream) {
  if (int rc = need_geometry(d)) return rc;
  if (!d->Utr || !Wtr || !Xtr || !w || !v) return zfr::fail("zfr_residual_vjp: missing buffer");
  zk::AdjPtrs j;
  j.w = w;
  j.Wtr = Wtr;
  j.Xtr = Xtr;
  const hipStream_t st = (hipStream_t)stream;
  int rc = zk::dispatch(zk::Op::AdjPrep, d, U, nullptr, st, j);
  if (!rc) rc = zk::dispatch(zk::Op::AdjMain, d, U, v, st, j);
  if (!rc) rc = zk::dispatch(zk::Op::AdjGather, d, U, v, st, j);
  return rc;
}

// Multi-case residual plan: cases grouped by (kernel, shape, flux branch), one launch per
// group per phase; the case tables live on the device for the plan's lifetime.
struct zfr_multi_s {
  std::vector<zk::MultiGroup> groups;  // launch order: every phase-0 group, then phase 1, phase 2
  std::vector<void *> allocs;
  ~zfr_multi_s() {
    for (void *q : allocs) (void)hipFree(q);
  }
};

int zfr_multi_create(const zfr_ele_desc_t *descs, const double *const *U, double *const *divF, int n,
                     zfr_multi_t **plan) {
  using zk::Op;
  if (!plan) return zfr::fail("zfr_multi_create: null plan pointer");
  *plan = nullptr;
  if (!descs || !U || !divF || n < 1) return zfr::fail("zfr_multi_create: need at least one case");
  // the kernels of each case's residual, by phase (the same choice as the per-solver path)
  struct Key {
    Op op; int D, N, visc, bm, fc;
    bool operator<(const Key &o) const {
      return std::tie(op, D, N, visc, bm, fc) < std::tie(o.op, o.D, o.N, o.visc, o.bm, o.fc);
    }
  };
  std::map<Key, std::vector<int>> phase[3];
  for (int i = 0; i < n; i++) {
    const zfr_ele_desc_t *d = &descs[i];
    const int D = d->nDims, N = d->order + 1;
    if ((D != 2 && D != 3) || N < 1 || N > 6)
      return zfr::fail("zfr_multi_create: supported shapes are quads and hexes with order 0..5");
    long nS = 1, nFF = 1;
    for (int k = 0; k < D; k++) nS *= N;
    for (int k = 0; k < D - 1; k++) nFF *= N;
    if (d->nSpts != nS || d->nFpts != 2 * D * nFF || d->nVars != D + 2)
      return zfr::fail("zfr_multi_create: element descriptor " + std::to_string(i) + " does not match (nDims, order)");
    if (d->fconv_type != 0 && d->fconv_type != 2) return zfr::fail("fconv_type must be 0 (Rusanov) or 2 (Roe)");
    if (!U[i] || !divF[i] || !d->Utr) return zfr::fail("zfr_multi_create: missing buffer of case " + std::to_string(i));
    const int fc = d->fconv_type == 2, visc = d->viscous != 0;
    const int bm = !visc ? 0 : (d->ldg_b == 0.5 ? 1 : (d->ldg_b == -0.5 ? 2 : 0));
    if (d->nEles <= 0) continue;
    if (visc && bm != 0) {
      if (!d->Ftr) return zfr::fail("zfr_multi_create: the one-sided path needs Ftr");
      phase[1][{Op::OneSided, D, N, 1, bm, fc}].push_back(i);  // extrapolates the partner traces itself
      phase[2][{Op::FaceCorr, D, N, 1, bm, 0}].push_back(i);
    } else {
      if (visc && !d->Gtr) return zfr::fail("zfr_multi_create: the general LDG path needs Gtr");
      if (visc) phase[0][{Op::Extrap, D, N, 0, 0, 0}].push_back(i);  // Euler: extrapolated in the divergence
      if (visc) phase[1][{Op::Grad, D, N, 1, 0, 0}].push_back(i);
      phase[2][{Op::Div, D, N, visc, bm, fc}].push_back(i);
    }
  }
  auto m = new zfr_multi_s();
  for (int ph = 0; ph < 3; ph++) {
    for (auto &kv : phase[ph]) {
      const Key &k = kv.first;
      const std::vector<int> &ids = kv.second;
      zk::MultiGroup g;
      g.op = k.op; g.D = k.D; g.N = k.N; g.visc = k.visc; g.bm = k.bm; g.fc = k.fc;
      g.nc = (int)ids.size();
      const size_t per_ele = (k.D == 2) ? zk::lds_per_element_n<2>(k.N, k.op, k.visc != 0)
                                        : zk::lds_per_element_n<3>(k.N, k.op, k.visc != 0);
      const int N = k.N;
      const int nS = (k.D == 2) ? N * N : N * N * N;
      const int nops = 2 * N + N * N + 2 * N, nops_pad = (nops + 7) & ~7;
      // one tile for the group, chosen for its total element count as for one solver
      zfr_ele_desc_t agg = descs[ids[0]];
      agg.nEles = 0;
      for (int i : ids) agg.nEles += descs[i].nEles;
      agg.eles_per_block = 0;
      const int defEB = (zk::kMaxThreads / nS) > 0 ? zk::kMaxThreads / nS : 1;
      g.EB = zk::eles_per_block(&agg, defEB, nops_pad, per_ele);
      g.lds = sizeof(double) * (nops_pad + g.EB * per_ele);
      if (g.lds > zk::kMaxLds) {
        delete m;
        return zfr::fail("zfr_multi_create: one element needs more LDS than a workgroup may use");
      }
      std::vector<zk::MultiCase> hc(ids.size());
      std::vector<long> hs(ids.size());
      long blocks = 0;
      for (size_t q = 0; q < ids.size(); q++) {
        const int i = ids[q];
        hc[q].a = zk::make_args(&descs[i], g.EB);
        hc[q].a.trace_sel = (k.op == Op::ExtrapPartner) ? 1 : 0;
        hc[q].U = U[i];
        hc[q].out = (k.op == Op::Div || k.op == Op::OneSided || k.op == Op::FaceCorr) ? divF[i] : nullptr;
        hs[q] = blocks;
        blocks += (descs[i].nEles + g.EB - 1) / g.EB;
      }
      if (blocks > 0x7fffffffL) {
        delete m;
        return zfr::fail("zfr_multi_create: too many workgroups in one group");
      }
      g.nblk = (unsigned)blocks;
      void *dc = nullptr, *ds = nullptr;
      hipError_t e = hipMalloc(&dc, hc.size() * sizeof(zk::MultiCase));
      if (e == hipSuccess) m->allocs.push_back(dc);
      if (e == hipSuccess) e = hipMalloc(&ds, hs.size() * sizeof(long));
      if (e == hipSuccess) m->allocs.push_back(ds);
      if (e == hipSuccess) e = hipMemcpy(dc, hc.data(), hc.size() * sizeof(zk::MultiCase), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(ds, hs.data(), hs.size() * sizeof(long), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        delete m;
        return zfr::check_hip(e, "zfr_multi_create");
      }
      g.cases = static_cast<zk::MultiCase *>(dc);
      g.starts = static_cast<long *>(ds);
      m->groups.push_back(g);
    }
  }
  *plan = m;
  return 0;
}

int zfr_multi_residual(zfr_multi_t *plan, void *stream) {
  if (!plan) return zfr::fail("zfr_multi_residual: null plan");
  const hipStream_t st = (hipStream_t)stream;
  for (const zk::MultiGroup &g : plan->groups) {
    if (g.nblk == 0) continue;
    if (g.D == 2) zk::dispatch_multi_n<2>(g, st);
    else zk::dispatch_multi_n<3>(g, st);
    const int rc = zfr::check_hip(hipGetLastError(), "zfr_multi_residual");
    if (rc) return rc;
  }
  return 0;
}

int zfr_multi_launches(const zfr_multi_t *plan) { return plan ? (int)plan->groups.size() : -1; }

void zfr_multi_destroy(zfr_multi_t *plan) { delete plan; }

int zfr_stream_copy(double *dst, const double *src, long n, void *stream) {
  if (n <= 0) return 0;
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return zfr::fail("zfr_stream_copy: pointers must be 16-byte aligned");
  const long n2 = n / 2, per_block = (long)zk::kMaxThreads * zk::kCopyUnroll;
  if (n2 > 0)
    hipLaunchKernelGGL(zk::k_copy, dim3((unsigned)((n2 + per_block - 1) / per_block)), dim3(zk::kMaxThreads), 0,
                       (hipStream_t)stream, reinterpret_cast<double2 *>(dst), reinterpret_cast<const double2 *>(src),
                       n2);
  if (n & 1) {
    hipError_t e = hipMemcpyAsync(dst + n - 1, src + n - 1, sizeof(double), hipMemcpyDeviceToDevice, (hipStream_t)stream);
    if (e != hipSuccess) return zfr::check_hip(e, "zfr_stream_copy");
  }
  return zfr::check_hip(hipGetLastError(), "zfr_stream_copy");
}

}  // extern "C"
