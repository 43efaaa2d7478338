// kernels.hpp -- device code of the DFR residual hot path (gfx950), shared by the
// per-shape translation units (kernels_shape.hip, one object per (nDims, order)) and the
// shape-independent kernels and C entry points (kernels.hip).
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
#pragma once

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
constexpr int kMaxRkStages = 4;  // RK44; the stage tables of the RK update hold this many

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
  // workgroup size bound of the per-shape kernels: 256 threads, or one element's points
  // rounded up to whole waves where an element has more (hexes of order 6 to 8)
  static constexpr int kThreads = (nS > kMaxThreads) ? ((nS + 63) / 64) * 64 : kMaxThreads;
};

// The hardware dispatches workgroup b of a grid to XCD b % 8 (each XCD has its own L2).
// xcd_block renumbers the grid so that every XCD walks one contiguous range of element
// blocks in increasing order: the partner lines an element reads from its next neighbours
// along x and y (one and one row of blocks later) then come from that XCD's L2 when the
// neighbour's own workgroup reads them.  A bijection of [0, n) for any n.
constexpr long kXcds = 8;
__device__ __forceinline__ long xcd_block(long b, long n) {
  const long q = n / kXcds, r = n % kXcds, x = b % kXcds, i = b / kXcds;
  return x * q + (x < r ? x : r) + i;
}

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
  // Epilogue of the last kernel of a residual: mode 0 none; 1 RK stage update; 2 final RK
  // combination; 3 srcw -= divF (p-multigrid source term: srcw holds the restricted
  // residual); 4 divF += src (a residual plus the source term)
  struct Rk {
    double *U;
    const double *Uini, *divF0;  // divF0: stage 0 of the [nStages][E][V][nS] residuals
    const double *src;           // p-multigrid source term added to every divF_s (modes 1, 2, 4), or null
    long total;
    double alpha, dt, beta[kMaxRkStages];
    int nStages, stage, mode;
    double *trace;  // mode 1: also write the new state's trace here (the next stage's Utr), or null
    double *Uini_out;         // modes 1, 2: a copy of Uini (read before U is written), or null
    double *corr;             // mode 2: corr = new U - corr_base, or null
    const double *corr_base;
    double *srcw;             // mode 3
    // p-multigrid transfers of each element fused after its residual or update
    // (PMGrid::restrict_pmg, prolong_err): restriction (modes 0, 4) of U and of the
    // residual with the table pmgR [pmgRn][N] into pmgUc, pmgRc ([E][V][pmgRn^D]), the U
    // read from pmgU when the kernel holds no copy of it; prolongation (mode 2 with corr)
    // pmgUf += pmgAlpha * P corr with pmgP [pmgPn][N]
    const double *pmgR, *pmgU;
    double *pmgUc, *pmgRc;
    int pmgRn, pmgPn;
    const double *pmgP;
    double pmgAlpha;
    double *pmgUf;
  } rk;
};

// One case of a multi-case launch: its kernel arguments and arrays.  A multi-case launch
// runs the same kernel over several independent solvers of one shape in one grid:
// workgroup b does block b - starts[c] of case c, where starts[c] <= b < starts[c + 1].
struct MultiCase {
  KArgs a;
  const double *U;
  double *out;
  int N;  // points per direction of the case's elements (the order of a mixed-order launch)
  int D;  // nDims of the case (a launch over quads and hexes together)
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

// Partner map entries: ele << (B + 1) | slot << B | fpt, with B = 8 flux-point bits for
// elements of at most 256 flux points and 9 above (hexes of order 6 to 8).
template <int D, int N>
struct NbrBits {
  static constexpr int B = (Shape<D, N>::nF <= 256) ? 8 : 9;
};

template <int D, int N>
__device__ __forceinline__ void unpack_nbr(unsigned int p, long &ep, int &slot, int &fp) {
  constexpr int B = NbrBits<D, N>::B;
  ep = (long)(p >> (B + 1));
  slot = (int)((p >> B) & 1u);
  fp = (int)(p & ((1u << B) - 1u));
}

template <int D, int N>
__device__ __forceinline__ int nbr_slot(unsigned int p) {
  return (int)((p >> NbrBits<D, N>::B) & 1u);
}

// Which sides' gradients enter the LDG flux (src/faces.cpp:1520-1545).
__device__ __forceinline__ bool need_left_grad(double beta) { return beta != -0.5; }
__device__ __forceinline__ bool need_right_grad(double beta) { return beta != 0.5; }

// Stage EB consecutive elements' U blocks and the 1D tables into LDS.
// stage_U in two halves, so that a kernel can issue other loads between them: the first
// batch of loads into registers, then the LDS stores and the rest.
template <int D, int N>
struct StageRegs {
  double ro[(Shape<D, N>::NOPS + 63) / 64], ru[Shape<D, N>::V];
};

template <int D, int N>
__device__ __forceinline__ void stage_U_load(const double *__restrict__ U, const KArgs &a, long e0, long neb,
                                             StageRegs<D, N> &r) {
  using S = Shape<D, N>;
  constexpr int VS = S::V * S::nS, OB = (S::NOPS + 63) / 64, B = S::V;
  const int tid = threadIdx.x, NT = blockDim.x;
  const double *Ub = U + e0 * VS;
  const long n = neb * VS;
#pragma unroll
  for (int k = 0; k < OB; k++) {
    const int i = tid + k * NT;
    r.ro[k] = (i < S::NOPS) ? a.ops[i] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < B; k++) {
    const long i = tid + (long)k * NT;
    r.ru[k] = (i < n) ? Ub[i] : 0.0;
  }
}

template <int D, int N>
__device__ __forceinline__ void stage_U_store(const double *__restrict__ U, const KArgs &a, double *lops,
                                              double *lU, long e0, long neb, int Ustride, const StageRegs<D, N> &r) {
  using S = Shape<D, N>;
  constexpr int VS = S::V * S::nS, OB = (S::NOPS + 63) / 64, B = S::V;
  const int tid = threadIdx.x, NT = blockDim.x;
  const double *Ub = U + e0 * VS;
  const long n = neb * VS;
#pragma unroll
  for (int k = 0; k < OB; k++) {
    const int i = tid + k * NT;
    if (i < S::NOPS) lops[i] = r.ro[k];
  }
#pragma unroll
  for (int k = 0; k < B; k++) {
    const long i = tid + (long)k * NT;
    if (i < n) {
      const long el = i / VS;
      lU[el * Ustride + (i - el * VS)] = r.ru[k];
    }
  }
  for (int i = tid + OB * NT; i < S::NOPS; i += NT) lops[i] = a.ops[i];
  for (long i = tid + (long)B * NT; i < n; i += NT) {
    const long el = i / VS;
    lU[el * Ustride + (i - el * VS)] = Ub[i];
  }
}

template <int D, int N>
__device__ __forceinline__ void stage_U(const double *__restrict__ U, const KArgs &a, double *lops, double *lU,
                                        long e0, long neb, int Ustride) {
  using S = Shape<D, N>;
  constexpr int VS = S::V * S::nS;
  // every load of the first batch (OB table values and V state values per thread: the
  // whole block when it has >= 64 threads and EB * nS of them) is issued before the
  // first LDS store, so the staging costs one memory round trip, not one per value
  constexpr int OB = (S::NOPS + 63) / 64, B = S::V;
  const int tid = threadIdx.x, NT = blockDim.x;
  const double *Ub = U + e0 * VS;
  const long n = neb * VS;
  double ro[OB], ru[B];
#pragma unroll
  for (int k = 0; k < OB; k++) {
    const int i = tid + k * NT;
    ro[k] = (i < S::NOPS) ? a.ops[i] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < B; k++) {
    const long i = tid + (long)k * NT;
    ru[k] = (i < n) ? Ub[i] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < OB; k++) {
    const int i = tid + k * NT;
    if (i < S::NOPS) lops[i] = ro[k];
  }
#pragma unroll
  for (int k = 0; k < B; k++) {
    const long i = tid + (long)k * NT;
    if (i < n) {
      const long el = i / VS;
      lU[el * Ustride + (i - el * VS)] = ru[k];
    }
  }
  // the rest, for blocks smaller than that
  for (int i = tid + OB * NT; i < S::NOPS; i += NT) lops[i] = a.ops[i];
  for (long i = tid + (long)B * NT; i < n; i += NT) {
    const long el = i / VS;
    lU[el * Ustride + (i - el * VS)] = Ub[i];
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

// Whether a fused epilogue writes the array the residual reads its state from.
inline bool rk_writes_input(const KArgs::Rk &r, const double *U) {
  if (r.mode != 1 && r.mode != 2) return false;
  return r.U == U || r.Uini_out == U || r.corr == U;
}

// partner_trace in two halves: the partner's N-point solution line through fp (V N values
// into registers, with the table row), then the extrapolation, the same sums.
template <int D, int N>
struct PartnerLine {
  double x[Shape<D, N>::V][N], l[N];
};

template <int D, int N>
__device__ __forceinline__ void partner_line_load(const double *__restrict__ U, const double *ops, long ep, int fp,
                                                  PartnerLine<D, N> &p) {
  using S = Shape<D, N>;
  int dim, side, base, str;
  fpt_info<D, N>(fp, dim, side, base, str);
  const double *lE = tab_lE<N>(ops, side);
  const double *Up = U + ep * S::V * S::nS + base;
#pragma unroll
  for (int m = 0; m < N; m++) p.l[m] = lE[m];
#pragma unroll
  for (int v = 0; v < S::V; v++)
#pragma unroll
    for (int m = 0; m < N; m++) p.x[v][m] = Up[v * S::nS + m * str];
}

template <int D, int N>
__device__ __forceinline__ void partner_line_reduce(const PartnerLine<D, N> &p, double *out) {
#pragma unroll
  for (int v = 0; v < Shape<D, N>::V; v++) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < N; m++) s += p.l[m] * p.x[v][m];
    out[v] = s;
  }
}

// RK update of one spt straight after its residual, in the last kernel of the residual
// (the stage update of src/solver.cpp:2897-2941 and the final combination of
// :2997-3062, with the same arithmetic as k_rk_stage / k_rk_final).  acc holds this
// stage's divF at the spt; the other stages' come from memory.  With a source term
// (src/solver.cpp:2946-2962, 3061-3073) every divF_s is replaced by divF_s + src.
// Side outputs: a copy of the starting state (Uini_out), and after a final combination
// the change against a stored state (corr: p-multigrid's coarse correction,
// src/multigrid.cpp:236-256).  Mode 3 subtracts the residual from a p-multigrid source
// term in place (src/multigrid.cpp:474-518).  pre: the spt's starting state, source term
// and detJ loaded ahead by rk_prefetch (null: loaded here).
template <int D>
struct RkPre {
  double u0[D + 2], src[D + 2], detJ;
};

template <int D>
__device__ __forceinline__ void rk_prefetch(const KArgs &a, long e, int t, int nS, RkPre<D> &p) {
  constexpr int V = D + 2, nM = D * D + 1;
  const bool upd = a.rk.mode == 1 || a.rk.mode == 2;
  const bool src = a.rk.src && (upd || a.rk.mode == 4);
  p.detJ = upd ? a.metric[(e * nM + D * D) * nS + t] : 1.0;
#pragma unroll
  for (int v = 0; v < V; v++) {
    const long i = (e * V + v) * nS + t;
    p.u0[v] = upd ? a.rk.Uini[i] : 0.0;
    p.src[v] = src ? a.rk.src[i] : 0.0;
  }
}

// BATCH: the final combination loads every other stage's residual before it consumes
// the first (one memory round trip instead of one per stage and variable; the short
// launches of the coarse p-multigrid levels are latency-bound); it holds (nStages - 1) V
// more values, which the kernels that carry it without a prefetched state (face
// correction, viscous divergence) would pay in occupancy, so they keep the plain loop.
template <int D, bool BATCH = false>
__device__ __forceinline__ void rk_epilogue(const KArgs &a, long e, int t, int nS, const double *acc,
                                            double *unew = nullptr, const RkPre<D> *pre = nullptr) {
  constexpr int V = D + 2, nM = D * D + 1;
  if (a.rk.mode == 0 || a.rk.mode == 4) return;
  if (a.rk.mode == 3) {
#pragma unroll
    for (int v = 0; v < V; v++) a.rk.srcw[(e * V + v) * nS + t] -= acc[v];
    return;
  }
  const double detJ = pre ? pre->detJ : a.metric[(e * nM + D * D) * nS + t];
  if (BATCH && a.rk.mode == 2) {
    const long i0 = e * V * nS + t;
    double u[V], sv[V], rs[kMaxRkStages][V];
#pragma unroll
    for (int v = 0; v < V; v++) {
      u[v] = pre ? pre->u0[v] : a.rk.Uini[i0 + v * nS];
      sv[v] = a.rk.src ? (pre ? pre->src[v] : a.rk.src[i0 + v * nS]) : 0.0;
    }
#pragma unroll
    for (int s = 0; s < kMaxRkStages; s++) {
      if (s < a.rk.nStages && s != a.rk.stage) {
#pragma unroll
        for (int v = 0; v < V; v++) rs[s][v] = a.rk.divF0[s * a.rk.total + i0 + v * nS];
      }
    }
    double cb[V];
    if (a.rk.corr) {
#pragma unroll
      for (int v = 0; v < V; v++) cb[v] = a.rk.corr_base[i0 + v * nS];
    }
    if (a.rk.Uini_out) {
#pragma unroll
      for (int v = 0; v < V; v++) a.rk.Uini_out[i0 + v * nS] = u[v];
    }
    // each variable's sum over the stages in order, as k_rk_final's
#pragma unroll
    for (int s = 0; s < kMaxRkStages; s++) {
      if (s < a.rk.nStages) {
#pragma unroll
        for (int v = 0; v < V; v++) {
          double r = (s == a.rk.stage) ? acc[v] : rs[s][v];
          if (a.rk.src) r += sv[v];
          u[v] -= a.rk.beta[s] * a.rk.dt / detJ * r;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < V; v++) {
      const long i = i0 + v * nS;
      a.rk.U[i] = u[v];
      if (unew) unew[v] = u[v];
      if (a.rk.corr) a.rk.corr[i] = u[v] - cb[v];
    }
    return;
  }
#pragma unroll
  for (int v = 0; v < V; v++) {
    const long i = (e * V + v) * nS + t;
    const double u0 = pre ? pre->u0[v] : a.rk.Uini[i];
    double sv = a.rk.src ? (pre ? pre->src[v] : a.rk.src[i]) : 0.0;
    if (a.rk.Uini_out) a.rk.Uini_out[i] = u0;
    if (a.rk.mode == 1 && a.rk.srcw) {  // form the source term first: srcw = src - divF
      sv = sv - acc[v];
      a.rk.srcw[i] = sv;
    }
    if (a.rk.mode == 1) {
      const double r = a.rk.src ? acc[v] + sv : acc[v];
      const double u = u0 - a.rk.alpha * a.rk.dt / detJ * r;
      a.rk.U[i] = u;
      if (unew) unew[v] = u;
    } else {
      double u = u0;
      for (int s = 0; s < a.rk.nStages; s++) {
        double r = (s == a.rk.stage) ? acc[v] : a.rk.divF0[s * a.rk.total + i];
        if (a.rk.src) r += sv;
        u -= a.rk.beta[s] * a.rk.dt / detJ * r;
      }
      a.rk.U[i] = u;
      if (unew) unew[v] = u;
      if (a.rk.corr) a.rk.corr[i] = u - a.rk.corr_base[i];
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
    if (owner >= 0 && nbr_slot<D, N>(a.nbr[e * nF + f]) == owner) continue;
    int dim, side, base, str;
    fpt_info<D, N>(f, dim, side, base, str);
    const double *lE = tab_lE<N>(lops, side);
#pragma unroll
    for (int v = 0; v < V; v++) a.rk.trace[(e * V + v) * nF + f] = line_extrap<N>(lE, w + v * nS, base, str);
  }
}

// Output o (x fastest) of a tensor-product table A [no][N] (one 1D table per direction)
// applied to one [N^D] block, with k_pmg_transfer's roundings: the sums along x, then y,
// then z, each from 0.0 over increasing input index.  Every output recomputes the partial
// sums of its lines, which are the same numbers the separate passes form.  blk and A are
// in LDS.
template <int D, int N>
__device__ __forceinline__ double tp_out(const double *blk, const double *A, int no, int o) {
  const unsigned uo = o, uno = no;  // unsigned: cheaper division sequences
  const int o1 = uo % uno, o2 = (uo / uno) % uno, o3 = (D == 3) ? uo / (uno * uno) : 0;
  const double *A1 = A + o1 * N, *A2 = A + o2 * N, *A3 = A + o3 * N;
  double a1[N];
#pragma unroll
  for (int i1 = 0; i1 < N; i1++) a1[i1] = A1[i1];
  double s3 = 0.0;
#pragma unroll 1
  for (int i3 = 0; i3 < ((D == 3) ? N : 1); i3++) {
    const double *b = blk + i3 * N * N;
    double s2 = 0.0;
    auto line = [&](int i2) {
      double s1 = 0.0;
#pragma unroll
      for (int i1 = 0; i1 < N; i1++) s1 = __builtin_fma(a1[i1], b[i2 * N + i1], s1);
      s2 = __builtin_fma(A2[i2], s1, s2);
    };
    if constexpr (D == 3) {
      // kept rolled: the hex kernels that carry this epilogue hold no registers to spare
#pragma unroll 1
      for (int i2 = 0; i2 < N; i2++) line(i2);
    } else {
#pragma unroll
      for (int i2 = 0; i2 < N; i2++) line(i2);
    }
    s3 = (D == 3) ? __builtin_fma(A3[i3], s2, s3) : s2;
  }
  return s3;
}

// One element's transfer out [V][no^D] = A blk (add: out += alpha A blk) by the element's nS
// threads; each (variable, output point) pair is one work item.  A thread takes K items at
// a time and, when it adds to the output, loads all K old values before it forms the first
// (one memory round trip per K items, not one per item).  NO > 0: no == NO, a constant.
template <int D, int N, int NO>
__device__ __forceinline__ void pmg_apply(const double *blk, const double *A, int no_rt, double *out, long e, int t,
                                          double alpha, bool add) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, V = S::V, K = 4;
  const int no = NO > 0 ? NO : no_rt;
  const int noD = (D == 3) ? no * no * no : no * no;
  const int total = V * noD;
  double *oe = out + e * total;  // [V][no^D] of this element
#pragma unroll 1
  for (int k0 = t; k0 < total; k0 += K * nS) {
    double old[K];
    if (add) {
#pragma unroll
      for (int j = 0; j < K; j++) {
        const int k = k0 + j * nS;
        old[j] = (k < total) ? oe[k] : 0.0;
      }
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
      const int k = k0 + j * nS;
      if (k < total) {
        const int v = (unsigned)k / (unsigned)noD, o = k - v * noD;
        const double x = tp_out<D, N>(blk + v * nS, A, no, o);
        oe[k] = add ? __builtin_fma(alpha, x, 1.0 * old[j]) : x;
      }
    }
  }
}

// The p-multigrid transfers fused at the end of a residual kernel (every thread of the
// workgroup calls this; it holds barriers).  res: this thread's residual (after the source
// term); unew: its new state (mode 2); ublk: the element's U block in LDS, or null (then U
// is read from a.rk.pmgU); w: a free per-element LDS block of at least V nS doubles; lops:
// the block's operator tables, no longer needed, which take the transfer table (the host
// fuses a transfer only where its table fits their N^2 + 4 N doubles, padded).
template <int D, int N>
__device__ __forceinline__ void pmg_epilogue(const KArgs &a, double *lops, const double *ublk, double *w, long e,
                                             int t, bool valid, const double *res, const double *unew) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, V = S::V;
  const KArgs::Rk &r = a.rk;
  auto apply = [&](const double *blk, int no, double *out, double alpha, bool add) {
    if (!valid) return;
    // the usual transfers (one order down, one up) with their output count known at compile
    // time: the work item's index splits by multiplications instead of integer divisions
    if (no == N - 1) pmg_apply<D, N, (N > 1 ? N - 1 : 0)>(blk, lops, no, out, e, t, alpha, add);
    else if (no == N + 1) pmg_apply<D, N, N + 1>(blk, lops, no, out, e, t, alpha, add);
    else pmg_apply<D, N, 0>(blk, lops, no, out, e, t, alpha, add);
  };
  const bool restr = r.pmgRc && (r.mode == 0 || r.mode == 4);
  const bool prol = r.pmgUf && r.mode == 2 && r.corr;
  if (!restr && !prol) return;
  __syncthreads();  // lops and w are consumed
  const double *A = restr ? r.pmgR : r.pmgP;
  const int nA = (restr ? r.pmgRn : r.pmgPn) * N;
  for (int i = threadIdx.x; i < nA; i += blockDim.x) lops[i] = A[i];
  if (valid) {
#pragma unroll
    for (int v = 0; v < V; v++)
      w[v * nS + t] = restr ? res[v] : unew[v] - r.corr_base[(e * V + v) * nS + t];
  }
  __syncthreads();
  if (prol) {
    apply(w, r.pmgPn, r.pmgUf, r.pmgAlpha, true);
    return;
  }
  apply(w, r.pmgRn, r.pmgRc, 1.0, false);
  if (ublk) {
    apply(ublk, r.pmgRn, r.pmgUc, 1.0, false);
  } else {
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int v = 0; v < V; v++) w[v * nS + t] = r.pmgU[(e * V + v) * nS + t];
    }
    __syncthreads();
    apply(w, r.pmgRn, r.pmgUc, 1.0, false);
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
    slot[k] = (a.trace_sel && el < neb && f < nF) ? nbr_slot<D, N>(a.nbr[e * nF + f]) : -1;
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
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_extrapolate(const double *__restrict__ U, KArgs a) {
  extrapolate_block<D, N>(U, a, xcd_block(blockIdx.x, gridDim.x));
}

// The same kernel for workgroups of one wave (see k_visc_onesided_w64).
template <int D, int N>
__global__ __launch_bounds__(64) void k_extrapolate_w64(const double *__restrict__ U, KArgs a) {
  extrapolate_block<D, N>(U, a, xcd_block(blockIdx.x, gridDim.x));
}

template <int D, int N>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_extrapolate_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
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
    unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
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
    const int sl = nbr_slot<D, N>(a.nbr[e * nF + f]);
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
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_gradient(const double *__restrict__ U, double *__restrict__ dU_out,
                                                          KArgs a) {
  gradient_block<D, N>(U, dU_out, a, xcd_block(blockIdx.x, gridDim.x));
}

template <int D, int N>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_gradient_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  gradient_block<D, N>(cs[c].U, nullptr, cs[c].a, (long)blockIdx.x - starts[c]);
}

// ---------------------------------------------------------------- kernel 3: divergence
// compute_F + compute_divF_spts + compute_common_F + compute_divF_fpts.
// Phases (viscous): U -> LDS | Ucomm -> lC | gradient -> lG | common flux -> lC |
// transformed flux -> lG | divergence.  Inviscid: U -> LDS | flux + common flux | divergence.
// SELF (Euler only): the partner traces are extrapolated here from the partner's U, so an
// Euler residual is this one launch; otherwise they are read from Utr.
// U is not __restrict__: with a fused RK update (a.rk.mode != 0) the epilogue writes a.rk.U,
// which is the same array; U is then read only while staging, before the first barrier.
template <int D, int N, bool VISC, int BM, int FC, bool SELF>
__device__ __forceinline__ void divergence_block(const double *U, double *__restrict__ divF, const KArgs &a,
                                                 long bid) {
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
  const int tid = threadIdx.x, el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  // Euler: everything this thread reads from global memory besides the staged block (the
  // partner states and face geometry of its flux points, its metric, the RK epilogue's
  // inputs) is loaded here, before the staging, so that the loads overlap one another
  // instead of each costing a memory round trip between the phases
  constexpr int PF = VISC ? 1 : S::FPT;
  double pUs[PF][V], pn[PF][D], pdA[PF], pij[D][D], pdet = 1.0;
  int psl[PF];
  RkPre<D> pre;
  if (!VISC && valid) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      const int f = t + k * nS;
      psl[k] = 0;
      if (f >= nF) continue;
      long ep; int sl, fp;
      unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
      psl[k] = sl;
      if constexpr (SELF) {
        partner_trace<D, N>(U, a.ops, ep, fp, pUs[k]);
      } else {
#pragma unroll
        for (int v = 0; v < V; v++) pUs[k][v] = a.Utr[(ep * V + v) * nF + fp];
      }
      const double *fg = a.fgeo + e * (D + 1) * nF + f;
#pragma unroll
      for (int d = 0; d < D; d++) pn[k][d] = fg[d * nF];
      pdA[k] = fg[D * nF];
    }
    load_metric<D>(a, e, t, nS, pij, pdet, false);
    rk_prefetch<D>(a, e, t, nS, pre);
  }
  stage_U<D, N>(U, a, lops, lU, e0, neb, V * nS);
  __syncthreads();
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
  if (!VISC && valid) {
#pragma unroll
    for (int k = 0; k < PF; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
      double Uo[V], UL[V], UR[V], G1[1], G2[1], Fo[V];
#pragma unroll
      for (int v = 0; v < V; v++) Uo[v] = line_extrap<N>(lE, u + v * nS, base, str);
#pragma unroll
      for (int v = 0; v < V; v++) {
        UL[v] = psl[k] ? pUs[k][v] : Uo[v];
        UR[v] = psl[k] ? Uo[v] : pUs[k][v];
      }
      common_flux<D, false, BM, FC>(UL, UR, G1, G2, pn[k], pdA[k], a, Fo);
#pragma unroll
      for (int v = 0; v < V; v++) c[v * nF + f] = Fo[v];
    }
  }
  if (VISC && valid) {
#pragma unroll 1
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
      long ep; int sl, fp;
      unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
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
              for (int v = 0; v < V; v++) {
                dUj[v] = line_extrap<N>(lE, g_l + (j * V + v) * nS, base, str);
                asm volatile("" : "+v"(dUj[v]) : : "memory");  // formed here, before the next line is loaded
              }
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
    if (VISC) {
      load_metric<D>(a, e, t, nS, ij, detJ, false);
    } else {
#pragma unroll
      for (int d1 = 0; d1 < D; d1++)
#pragma unroll
        for (int d2 = 0; d2 < D; d2++) ij[d1][d2] = pij[d1][d2];
    }
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
  double unew[V], res[V];
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
    if (a.rk.mode == 4) {
#pragma unroll
      for (int v = 0; v < V; v++) acc[v] += VISC ? a.rk.src[(e * V + v) * nS + t] : pre.src[v];
    }
#pragma unroll
    for (int v = 0; v < V; v++) divF[(e * V + v) * nS + t] = acc[v];
#pragma unroll
    for (int v = 0; v < V; v++) res[v] = acc[v];
    rk_epilogue<D, !VISC>(a, e, t, nS, acc, unew, VISC ? nullptr : &pre);
  }
  // the new state's full trace for the next stage, into a buffer other than the Utr this
  // launch reads; the element's U block is no longer read
  if (a.rk.mode == 1 && a.rk.trace) rk_trace<D, N>(a, lops, lU + el * V * nS, e, t, valid, unew, -1);
  // p-multigrid transfers (the U block is still the staged state; lG is free)
  if (a.rk.pmgRc || a.rk.pmgUf) pmg_epilogue<D, N>(a, lops, lU + el * V * nS, g_l, e, t, valid, res, unew);
}

// Occupancy target of the divergence kernel: the general-beta hex p=3 Rusanov kernel fits
// four waves per SIMD (it took 130 VGPRs, three waves, once it carried the p-multigrid
// transfer epilogue); other shapes keep the default allocation.
template <int D, int N, bool VISC, int FC>
struct DivWaves {
  static constexpr int value = (D == 3 && N == 4 && VISC && FC == 0) ? 4 : 1;
};

template <int D, int N, bool VISC, int BM, int FC, bool SELF = !VISC>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) __attribute__((amdgpu_waves_per_eu(DivWaves<D, N, VISC, FC>::value)))
void k_divergence(const double *U, double *__restrict__ divF, KArgs a) {
  static_assert(!(VISC && SELF), "the viscous divergence reads the gradient traces, hence the traces too");
  divergence_block<D, N, VISC, BM, FC, SELF>(U, divF, a, xcd_block(blockIdx.x, gridDim.x));
}

// The same kernel for workgroups of one wave (the small meshes of coarse p-multigrid
// levels and small cases): its barriers need no cross-wave synchronisation.
template <int D, int N, bool VISC, int BM, int FC, bool SELF = !VISC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(DivWaves<D, N, VISC, FC>::value)))
void k_divergence_w64(const double *U, double *__restrict__ divF, KArgs a) {
  divergence_block<D, N, VISC, BM, FC, SELF>(U, divF, a, xcd_block(blockIdx.x, gridDim.x));
}

template <int D, int N, bool VISC, int BM, int FC>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_divergence_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
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
// LDS allows per CU, and both face phases fit it (Rusanov 110 VGPRs; Roe at the cap) once
// each gradient line's sum is pinned before the next line is loaded (the value-tied asm
// fences in the face phase: without them the scheduler issued all the face phase's LDS
// loads before their FMAs, 168 VGPRs for Roe, 196 for quads at p=4).  Quads at p=4 with
// Rusanov fit six waves (78-80 VGPRs, no spills; the default allocation took 90 VGPRs,
// five waves); seven spill.  Other shapes keep the default allocation (at six waves the
// quad kernels at p = 1..3 spill).
template <int D, int N, int FC>
struct OneSidedWaves {
  static constexpr int value = (D == 3 && N == 4) ? 4 : ((D == 2 && N == 5 && FC == 0) ? 6 : 1);
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
  // The staging loads and all neighbour entries are issued first, then every partner load,
  // and only then are values consumed: the prologue costs two dependent memory round trips
  // rather than two per flux point plus the staging's.
  int slot[S::FPT];
  double cp[S::FPT][V];
  constexpr bool self = SELF;
  StageRegs<D, N> sr;
  stage_U_load<D, N>(U, a, e0, neb, sr);
  unsigned nb[S::FPT];
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    nb[k] = (valid && f < nF) ? a.nbr[e * nF + f] : 0u;
  }
  PartnerLine<D, N> pl[self ? S::FPT : 1];
  bool own[S::FPT];
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    slot[k] = -1;
    own[k] = false;
    if (valid && f < nF) {
      long ep; int sl, fp;
      unpack_nbr<D, N>(nb[k], ep, sl, fp);
      slot[k] = sl;
      own[k] = sl == owner;
      if (own[k]) {
        if constexpr (self) {
          partner_line_load<D, N>(U, a.ops, ep, fp, pl[k]);
        } else {
#pragma unroll
          for (int v = 0; v < V; v++) cp[k][v] = a.Utr[(ep * V + v) * nF + fp];
        }
      }
    }
  }
  if constexpr (self) {
#pragma unroll
    for (int k = 0; k < S::FPT; k++)
      if (own[k]) partner_line_reduce<D, N>(pl[k], cp[k]);
  }
  stage_U_store<D, N>(U, a, lops, lU, e0, neb, V * nS, sr);
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
          unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
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
            asm volatile("" : "+v"(dUj[v]) : : "memory");  // formed here, before the next line is loaded
          }
        }, n, a, G);
        if (owner == 0) ldg_common_add<D, BM>(Uo, Up, G, G, a, Fo);
        else ldg_common_add<D, BM>(Up, Uo, G, G, a, Fo);
#pragma unroll
        for (int v = 0; v < V; v++) __builtin_nontemporal_store(Fo[v], &a.Ftr[(e * V + v) * nF + f]);
        continue;
      }
      double G[V];
      // one variable's line at a time (a compiler barrier after each): 128 VGPRs instead
      // of 160, i.e. one more wave per SIMD, worth more than the loads' overlap
      visc_normal<D>(Uo, [&](int j, double *dUj) {
#pragma unroll
        for (int v = 0; v < V; v++) {
          dUj[v] = line_extrap<N>(lE, g_l + (j * V + v) * nS, base, str);
          asm volatile("" : "+v"(dUj[v]) : : "memory");  // formed here, before the next line is loaded
        }
      }, n, a, G);
      asm volatile("" ::: "memory");  // load the partner's state only now: fewer live registers
      {
        long ep; int sl, fp;
        unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
        const long ip = self ? (e * V * nF + f) : (ep * V * nF + fp);  // cached since phase 1
#pragma unroll
        for (int v = 0; v < V; v++) Up[v] = a.Utr[ip + v * nF];
      }
      double Fo[V];
      if (owner == 0) common_flux<D, true, BM, FC>(Uo, Up, G, G, n, 1.0, a, Fo);
      else common_flux<D, true, BM, FC>(Up, Uo, G, G, n, 1.0, a, Fo);
#pragma unroll
      for (int v = 0; v < V; v++) __builtin_nontemporal_store(Fo[v], &a.Ftr[(e * V + v) * nF + f]);
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
  // streaming stores for the outputs (divF, Ftr): the next kernel reads them, this one
  // does not, and they should not displace the partner lines other workgroups are about to
  // read from L2 (-0.8% per residual, measured A/B)
#pragma unroll
  for (int v = 0; v < V; v++) __builtin_nontemporal_store(acc[v], &divF[(e * V + v) * nS + t]);
}

template <int D, int N, int BM, int FC, bool SELF>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) __attribute__((amdgpu_waves_per_eu(OneSidedWaves<D, N, FC>::value)))
void k_visc_onesided(const double *__restrict__ U, double *__restrict__ divF, KArgs a) {
  visc_onesided_block<D, N, BM, FC, SELF>(U, divF, a, xcd_block(blockIdx.x, gridDim.x));
}

// The same kernel for workgroups of one wave (64 threads: one hex p=3 element per
// workgroup, the tile the autotuner picks there): the compiler knows the workgroup is a
// single wave, so its barriers need no cross-wave synchronisation.
template <int D, int N, int BM, int FC, bool SELF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OneSidedWaves<D, N, FC>::value)))
void k_visc_onesided_w64(const double *__restrict__ U, double *__restrict__ divF, KArgs a) {
  visc_onesided_block<D, N, BM, FC, SELF>(U, divF, a, xcd_block(blockIdx.x, gridDim.x));
}

template <int D, int N, int BM, int FC>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) __attribute__((amdgpu_waves_per_eu(OneSidedWaves<D, N, FC>::value)))
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
      unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
      const double dAs = a.fgeo[(e * (D + 1) + D) * nF + f];
      const double *src = (sl == owner) ? a.Ftr + e * V * nF + f : a.Ftr + ep * V * nF + fp;
#pragma unroll
      for (int v = 0; v < V; v++) c[v * nF + f] = src[v * nF] * dAs;
    }
  }
  __syncthreads();
  double unew[V], res[V];
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
    if (a.rk.mode == 4) {
#pragma unroll
      for (int v = 0; v < V; v++) acc[v] += a.rk.src[(e * V + v) * nS + t];
    }
#pragma unroll
    for (int v = 0; v < V; v++) __builtin_nontemporal_store(acc[v], &divF[(e * V + v) * nS + t]);
#pragma unroll
    for (int v = 0; v < V; v++) res[v] = acc[v];
    rk_epilogue<D>(a, e, t, nS, acc, unew);
  }
  // the new state's partner-owned trace for the next stage
  if (a.rk.mode == 1 && a.rk.trace) rk_trace<D, N>(a, lops, c, e, t, valid, unew, owner);
  // p-multigrid transfers (the common fluxes in c are consumed; U comes from pmgU)
  if (a.rk.pmgRc || a.rk.pmgUf) pmg_epilogue<D, N>(a, lops, nullptr, c, e, t, valid, res, unew);
}

template <int D, int N, int BM>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_face_correction(double *__restrict__ divF, KArgs a) {
  face_correction_block<D, N, BM>(divF, a, xcd_block(blockIdx.x, gridDim.x));
}

template <int D, int N, int BM>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_face_correction_multi(const MultiCase *__restrict__ cs, const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  face_correction_block<D, N, BM>(cs[c].out, cs[c].a, (long)blockIdx.x - starts[c]);
}

// ---------------------------------------------------------------- reverse mode
#include "adjoint.hpp"

// ---------------------------------------------------------------- dispatch
// LDS budget per workgroup; kept at 64 KiB so several workgroups share a CU.  Shapes whose
// element alone needs more than 256 threads (hexes of order 6 to 8) may take up to the
// 160 KiB a gfx950 workgroup can allocate (one element per workgroup).
constexpr size_t kMaxLds = 64 * 1024;
constexpr size_t kMaxLdsBig = 160 * 1024;
template <int D, int N>
constexpr size_t max_lds() { return Shape<D, N>::nS > kMaxThreads ? kMaxLdsBig : kMaxLds; }
inline size_t max_lds_n(int D, int N) {
  int nS = N * N;
  if (D == 3) nS *= N;
  return nS > kMaxThreads ? kMaxLdsBig : kMaxLds;
}

// Elements per workgroup: the requested tile, limited by the thread count and the LDS
// the kernel needs per element (per_ele doubles).
// one_wave: the launch's workgroup is EB * nS threads, so that a one-wave tile runs as a
// one-wave workgroup; false for launches of kMaxThreads-thread workgroups (multi-case
// launches, persistent plans), where the threads beyond EB * nS would idle
inline int eles_per_block(const zfr_ele_desc_t *d, int defEB, size_t fixed, size_t per_ele, bool one_wave = true) {
  int eb = d->eles_per_block > 0 ? d->eles_per_block : defEB;
  if (d->eles_per_block <= 0) {
    // small meshes (coarse multigrid levels, small cases): keep >= 1024 workgroups, four
    // per CU, so that every SIMD holds a wave; the kernels are latency-bound there, and
    // partly filled waves cost less than idle SIMDs
    const long per_cu4 = ((long)d->nEles + 1023) / 1024;
    if (per_cu4 < eb) {
      eb = per_cu4 < 1 ? 1 : (int)per_cu4;
      // one-wave workgroups (barriers without cross-wave synchronisation) where they
      // still leave >= 1024 of them
      const int eb1 = (64 / d->nSpts) > 0 ? 64 / d->nSpts : 1;
      if (one_wave && eb1 < eb && ((long)d->nEles + eb1 - 1) / eb1 >= 1024) eb = eb1;
    }
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

enum class Op {
  Extrap, ExtrapPartner, Grad, Div, OneSided, OneSidedFromTrace, FaceCorr,
  AdjPrep, AdjMain, AdjGather,  // reverse mode (Euler, LDG beta = +-1/2; AdjPrep and AdjGather also general beta)
  AdjFace, AdjMainGen           // reverse mode, general LDG beta
};

// Extra operands of the reverse-mode launches.
struct AdjPtrs {
  const double *w = nullptr;  // cotangent of divF
  double *Wtr = nullptr;      // [E][V][nF] face weights of w
  double *Xtr = nullptr;      // [E][V][nF] partner trace cotangent at owned points
  const KArgs::Rk *rk = nullptr;  // fused RK update for Div / FaceCorr
};

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
    case Op::AdjFace: per_ele = 0; break;
    case Op::AdjMainGen: per_ele = (size_t)AdjGenLds<D, N>::per_ele; break;
  }
  return per_ele;
}

// One launch of a multi-case plan: every case of one kernel, shape and flux branch.
struct MultiGroup {
  Op op = Op::Extrap;
  int D = 0, N = 0, visc = 0, bm = 0, fc = 0;  // bm: LDG branch (0 general, 1 beta = 1/2, 2 beta = -1/2)
                                               // N = 0: cases of several orders (kernels_mixed.hip);
                                               // with D = 0 also of both nDims
  int EB = 1, nc = 0;
  unsigned nblk = 0;
  size_t lds = 0;
  MultiCase *cases = nullptr;  // device [nc]
  long *starts = nullptr;      // device [nc]
};

template <int D>
size_t lds_per_element_n(int N, Op op, bool viscous) {
  switch (N) {
    case 1: return lds_per_element<D, 1>(op, viscous);
    case 2: return lds_per_element<D, 2>(op, viscous);
    case 3: return lds_per_element<D, 3>(op, viscous);
    case 4: return lds_per_element<D, 4>(op, viscous);
    case 5: return lds_per_element<D, 5>(op, viscous);
    case 6: return lds_per_element<D, 6>(op, viscous);
    case 7: return lds_per_element<D, 7>(op, viscous);
    case 8: return lds_per_element<D, 8>(op, viscous);
    case 9: return lds_per_element<D, 9>(op, viscous);
  }
  if constexpr (D == 2) {
    if (N == 10) return lds_per_element<2, 10>(op, viscous);
  }
  return 0;
}

// Defined and explicitly instantiated per shape in kernels_shape.hip.
template <int D, int N>
int launch(Op op, const zfr_ele_desc_t *d, const double *U, double *out, hipStream_t st, const AdjPtrs &j);
template <int D, int N>
void launch_multi(const MultiGroup &g, hipStream_t st);
// Defined and explicitly instantiated in kernels_mixed.hip for D = 2, 3 and 0 (both).
template <int D>
void launch_mixed(const MultiGroup &g, hipStream_t st);
// The ops a mixed-order launch covers (the Euler divergence and the one-sided LDG pair).
inline bool mixed_op(Op op, int visc) { return (op == Op::Div && !visc) || op == Op::OneSided || op == Op::FaceCorr; }
// Orders a mixed-order launch compiles in (points per direction 1..6); quads of higher
// order get launches of their own shape.
constexpr int kMixedMaxN = 6;

// ---------------------------------------------------------------- persistent plans
// One phase of a persistent plan: an Euler divergence launch (a residual, or a fused RK
// stage / source-term / restriction epilogue around one), recorded instead of launched.
// k_persist runs every phase of a plan in one cooperative launch: its resident workgroups
// walk the phase's element blocks (block b as the single launch's workgroup b) and meet
// at a grid-wide barrier before the next phase.
struct PersistPhase {
  KArgs a;
  const double *U;
  double *out;
  long nblk;     // element blocks of the phase (the single launch's grid)
  int N, D, fc, self;  // points per direction, nDims, Roe, partner traces extrapolated from U
};
// While set on a thread, the Euler divergence launches of that thread are appended here
// and every other launch entry point fails (a plan holds nothing else).
struct PersistRecorder {
  std::vector<PersistPhase> phases;
  size_t lds = 0;  // the largest LDS of the phases
  int D = 0;
};
extern thread_local PersistRecorder *g_persist;
// 0, or a failure status naming the launch that cannot be recorded.
int refuse_while_recording(const char *what);
// Defined and instantiated in kernels_mixed.hip (D = 2, 3).
template <int D>
hipError_t launch_persist(const PersistPhase *phases, int nph, unsigned grid, size_t lds, hipStream_t st);
template <int D>
hipError_t persist_blocks_per_cu(size_t lds, int *n);

}  // namespace zk
