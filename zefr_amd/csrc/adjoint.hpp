// adjoint.hpp -- reverse-mode kernels of the residual (included by kernels.hip, inside
// namespace zk, after the forward kernels and their helpers).
//
// Vector-Jacobian product of one residual: v = (d divF / dU)^T w at the state U, for the
// one-sided LDG path (beta = +-1/2) and for Euler (new; no reference counterpart).
// The forward path, read backwards, is
//   divF = Dss^T-lines(tF) + signed Dsf (Fcomm),  Fcomm = Ftr(owner) dA_side,
//   Ftr  = RusanovLDG(U_L, U_R, G_owner . n),   G = viscous flux of (Utr, E grad),
//   tF   = F(U, grad) . adj,  grad = adj^T (Dss U + Dsf Ucomm) / detJ,
//   Ucomm = partner trace at owned points, own trace elsewhere, Utr = E U.
// The owner side of each face (slot 0 for beta = +1/2 and Euler, slot 1 for -1/2) does all
// the face work, as in the forward path: it combines both sides' weights of its common
// flux, and hands the partner's share of the trace cotangent back through Xtr.
//   k_adj_prep   : Wtr = dA_side * (signed Dsf^T w) at every flux point;
//   k_adj_main   : forward recompute, volume and face adjoints, gradient adjoint; writes
//                  v (all of this element's own contributions) and Xtr at owned points;
//   k_adj_gather : v += E^T (partner Xtr) at the points the partner owns.
#pragma once

// k_adj_gather's BM for general LDG beta: the partner's cotangent at every flux point.
constexpr int kAdjAllPoints = 3;

// Stage EB consecutive [V][nS] blocks of x into LDS (stride V*nS).
template <int D, int N>
__device__ __forceinline__ void stage_block(const double *__restrict__ x, double *lx, long e0, long neb) {
  using S = Shape<D, N>;
  const double *xb = x + e0 * S::V * S::nS;
  for (long i = threadIdx.x; i < neb * S::V * S::nS; i += blockDim.x) lx[i] = xb[i];
}

// Transposed line operator: sum_i Dss[i][m] x[base + i*str].
template <int N>
__device__ __forceinline__ double line_T_Dss(const double *lops, int m, const double *blk, int base, int str) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) s += tab_Dss<N>(lops, i)[m] * blk[base + i * str];
  return s;
}

// sum over the spts t of a line of Dsf[m_t][side] x(t)
template <int N>
__device__ __forceinline__ double line_Dsf(const double *lops, int side, const double *blk, int base, int str) {
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < N; m++) s += tab_Dsf<N>(lops, m)[side] * blk[base + m * str];
  return s;
}

template <int D, int N>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_adj_prep(const double *__restrict__ w, double *__restrict__ Wtr,
                                                          KArgs a) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  extern __shared__ double lds[];
  double *lops = lds;
  double *lW = lops + S::NOPS_PAD;
  const long e0 = xcd_block(blockIdx.x, gridDim.x) * a.EB;
  const long neb = min((long)a.EB, (long)a.E - e0);
  for (int i = threadIdx.x; i < S::NOPS; i += blockDim.x) lops[i] = a.ops[i];
  stage_block<D, N>(w, lW, e0, neb);
  __syncthreads();
  const int tid = threadIdx.x, el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  if (el >= neb) return;
  const double *wl = lW + el * V * nS;
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    if (f >= nF) continue;
    int dim, side, base, str;
    fpt_info<D, N>(f, dim, side, base, str);
    const double dAs = a.fgeo[(e * (D + 1) + D) * nF + f];
    const double sg = side ? dAs : -dAs;
#pragma unroll
    for (int v = 0; v < V; v++) Wtr[(e * V + v) * nF + f] = sg * line_Dsf<N>(lops, side, wl + v * nS, base, str);
  }
}

// LDS per element: U [V][nS]; one region that holds the common solution, then the forward
// gradient [D][V][nS], then (viscous, once the face adjoint has read the gradient) the own
// trace cotangent [V][nF]; the gradient cotangent [D][V][nS] (viscous); the own trace
// cotangent in a region of its own (Euler); the owned-face mask.  w is read from global
// memory (each value by the few threads of its lines, from cache).  The partner's share
// of the trace cotangent goes straight to Xtr; __syncthreads orders this block's global
// accesses too.
template <int D, int N, bool VISC>
struct AdjLds {
  using S = Shape<D, N>;
  static constexpr int R0 = (S::V * S::nF > D * S::V * S::nS) ? S::V * S::nF : D * S::V * S::nS;
  static constexpr int R = VISC ? R0 : 0;
  static constexpr int GB = VISC ? D * S::V * S::nS : 0;
  static constexpr int TO = VISC ? 0 : S::V * S::nF;
  static constexpr int per_ele = S::V * S::nS + R + GB + TO + 1;  // doubles
};

template <int D, int N, bool VISC, int BM, int FC>
__device__ __forceinline__ void adj_main_block(const double *__restrict__ U, const double *__restrict__ w,
                                               double *__restrict__ Wtr, double *__restrict__ Xtr,
                                               double *__restrict__ vout, const KArgs &a, long bid) {
  using S = Shape<D, N>;
  using L = AdjLds<D, N, VISC>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  constexpr int owner = (BM == 2) ? 1 : 0;
  extern __shared__ double lds[];
  const int EB = a.EB;
  double *lops = lds;
  double *lU = lops + S::NOPS_PAD;
  double *lR = lU + EB * V * nS;
  double *lGB = lR + EB * L::R;
  double *lTo = lGB + EB * L::GB;
  unsigned *lmask = reinterpret_cast<unsigned *>(lTo + EB * L::TO);  // owned faces per element
  const long e0 = bid * EB;
  const long neb = min((long)EB, (long)a.E - e0);
  const int tid = threadIdx.x, el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  // the staging loads, this thread's neighbour entries and (viscous) the partner's solution
  // line at each point this element owns are all issued before any is consumed: two
  // dependent memory round trips instead of one per flux point after the staging's
  StageRegs<D, N> sr;
  stage_U_load<D, N>(U, a, e0, neb, sr);
  unsigned nb[S::FPT];
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    nb[k] = (valid && f < nF) ? a.nbr[e * nF + f] : 0u;
  }
  PartnerLine<D, N> pl[VISC ? S::FPT : 1];
  double cp[VISC ? S::FPT : 1][V];
  bool own[S::FPT];
#pragma unroll
  for (int k = 0; k < S::FPT; k++) {
    const int f = t + k * nS;
    own[k] = false;
    if (valid && f < nF) {
      long ep; int sl, fp;
      unpack_nbr<D, N>(nb[k], ep, sl, fp);
      own[k] = sl == owner;
      if (VISC && own[k]) partner_line_load<D, N>(U, a.ops, ep, fp, pl[VISC ? k : 0]);
    }
  }
  if (VISC) {
#pragma unroll
    for (int k = 0; k < S::FPT; k++)
      if (own[k]) partner_line_reduce<D, N>(pl[VISC ? k : 0], cp[VISC ? k : 0]);
  }
  stage_U_store<D, N>(U, a, lops, lU, e0, neb, V * nS, sr);
  if (threadIdx.x < EB) lmask[threadIdx.x] = 0u;
  __syncthreads();
  const double *u = lU + el * V * nS;
  const double *we = w + e * V * nS;
  double *c = lR + el * L::R, *g_l = lR + el * L::R;
  // viscous: the trace cotangent takes the gradient's region after the face adjoint
  double *gb = lGB + el * L::GB, *to = VISC ? lR + el * L::R : lTo + el * L::TO;
  const int ia = t % N, ib = (t / N) % N, ic = (D == 3) ? t / (N * N) : 0;
  // 1. owned list, zeroed own-trace cotangent, common solution (viscous)
  if (valid) {
#pragma unroll
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      if (own[k] && f % S::nFF == 0) atomicOr(&lmask[el], 1u << (f / S::nFF));
      if (VISC) {
        if (own[k]) {
          // the partner's trace, from its U; parked in Utr at this element's own point
#pragma unroll
          for (int v = 0; v < V; v++) {
            c[v * nF + f] = cp[VISC ? k : 0][v];
            a.Utr[(e * V + v) * nF + f] = cp[VISC ? k : 0][v];
          }
        } else {
          int dim, side, base, str;
          fpt_info<D, N>(f, dim, side, base, str);
          const double *lE = tab_lE<N>(lops, side);
#pragma unroll
          for (int v = 0; v < V; v++) c[v * nF + f] = line_extrap<N>(lE, u + v * nS, base, str);
        }
      }
      if (!VISC) {
#pragma unroll
        for (int v = 0; v < V; v++) to[v * nF + f] = 0.0;
      }
    }
  }
  __syncthreads();
  double ij[D][D], detJ = 1.0;
  if (valid) load_metric<D>(a, e, t, nS, ij, detJ, VISC);
  // 2. forward physical gradient at every spt (viscous), into the region of c
  if (VISC) {
    double g[V][D];
    if (valid) spt_gradient<D, N>(lops, u, c, t, ij, detJ, g);
    __syncthreads();  // the common solution is consumed
    if (valid) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = g[v][d];
    }
    __syncthreads();
  }
  // 3. volume adjoint: tFb[d1] = Dss^T w along the d1 lines; F(U, grad) reversed
  double Ub[V];
#pragma unroll
  for (int v = 0; v < V; v++) Ub[v] = 0.0;
  if (valid) {
    double Us[V];
#pragma unroll
    for (int v = 0; v < V; v++) Us[v] = u[v * nS + t];
    const Prim<D> p = primitives<D>(Us, a.gamma);
    double tFb[D][V];
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
      const int str = (d == 0) ? 1 : ((d == 1) ? N : N * N);
      const int base = t - id * str;
#pragma unroll
      for (int v = 0; v < V; v++) tFb[d][v] = line_T_Dss<N>(lops, id, we + v * nS, base, str);
      conv_proj_vjp<D>(Us, p, ij[d], tFb[d], a.gamma, Ub);
    }
    if (VISC) {
      // tF[d1] -= T . adj[d1]: cotangent on T[v][j] is -sum_d1 tFb[d1][v] adj[d1][j]
      viscous_vjp<D>(
          Us, p,
          [&](int j, double *dUj) {
#pragma unroll
            for (int v = 0; v < V; v++) dUj[v] = g_l[(j * V + v) * nS + t];
          },
          [&](int v, int j) {
            double s = 0.0;
#pragma unroll
            for (int d1 = 0; d1 < D; d1++) s += tFb[d1][v] * ij[d1][j];
            return -s;
          },
          [&](int j, const double *gbj) {
#pragma unroll
            for (int v = 0; v < V; v++) gb[(j * V + v) * nS + t] = gbj[v];
          },
          a, Ub);
    }
  }
  __syncthreads();
  // 4. face adjoint at owned points; the partner's trace cotangent goes to Xtr, the
  //    gradient cotangent spreads face by face (lines of one face never meet, so the
  //    accumulation order is fixed)
  const unsigned mask = valid ? lmask[el] : 0u;
  const int cnt = __popc(mask) * S::nFF;
  // viscous: the owned points' trace cotangents go to LDS once the gradient is read; until
  // then each waits in its own point's Wtr entry, which this thread has just read and no
  // other element reads (the partner reads Wtr only at the points it owns)
  int fown[S::FPT];
#pragma unroll
  for (int kk = 0; kk < S::FPT; kk++) fown[kk] = -1;
#pragma unroll
  for (int kk = 0; kk < S::FPT; kk++) {
    // viscous runs use one element per workgroup, so cnt is the same for every thread and
    // a pass with no owned point left (and its barriers) is skipped by all of them
    if (VISC && kk > 0 && kk * nS >= cnt) break;
    const int k = t + kk * nS;
    const bool has = valid && k < cnt;
    int f = 0, dim = 0, side = 0, base = 0, str = 0;
    double dGb[V][D];
    if (has) {
      f = owned_fpt<D, N>(mask, k);
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
      const double *fg = a.fgeo + e * (D + 1) * nF + f;
      double n[D];
#pragma unroll
      for (int d = 0; d < D; d++) n[d] = fg[d * nF];
      double Uo[V], Uob[V], Gb[V];
      {
        long ep; int sl, fp;
        unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
        double Up[V], Fb[V], Upb[V];
        if (VISC) {  // parked in phase 1
#pragma unroll
          for (int v = 0; v < V; v++) Up[v] = a.Utr[(e * V + v) * nF + f];
        } else {
          partner_trace<D, N>(U, a.ops, ep, fp, Up);
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
          Uo[v] = line_extrap<N>(lE, u + v * nS, base, str);
          Fb[v] = Wtr[(e * V + v) * nF + f] + Wtr[(ep * V + v) * nF + fp];
          Uob[v] = Upb[v] = Gb[v] = 0.0;
        }
        if (owner == 0) common_flux_vjp<D, VISC, BM, FC>(Uo, Up, n, 1.0, a, Fb, Uob, Upb, Gb, nullptr);
        else common_flux_vjp<D, VISC, BM, FC>(Up, Uo, n, 1.0, a, Fb, Upb, Uob, Gb, nullptr);
#pragma unroll
        for (int v = 0; v < V; v++) Xtr[(e * V + v) * nF + f] = Upb[v];
      }
      asm volatile("" ::: "memory");
      if (VISC) {
        // G = -T(Uo, grad_o) . n
        const Prim<D> po = primitives<D>(Uo, a.gamma);
        viscous_vjp<D>(
            Uo, po,
            [&](int j, double *dUj) {
#pragma unroll
              for (int v = 0; v < V; v++) {
                dUj[v] = line_extrap<N>(lE, g_l + (j * V + v) * nS, base, str);
                asm volatile("" : "+v"(dUj[v]) : : "memory");  // formed here, before the next line is loaded
              }
            },
            [&](int v, int j) { return -Gb[v] * n[j]; },
            [&](int j, const double *gbj) {
#pragma unroll
              for (int v = 0; v < V; v++) dGb[v][j] = gbj[v];
            },
            a, Uob);
      }
      if (VISC) {
        fown[kk] = f;
#pragma unroll
        for (int v = 0; v < V; v++) Wtr[(e * V + v) * nF + f] = Uob[v];
      } else {
#pragma unroll
        for (int v = 0; v < V; v++) to[v * nF + f] = Uob[v];
      }
    }
    if (VISC) {
      const int face = has ? f / S::nFF : -1;
      const double *lE = tab_lE<N>(lops, side);
#pragma unroll 1
      for (int q = 0; q < 2 * D; q++) {
        __syncthreads();
        if (face == q) {
#pragma unroll
          for (int m = 0; m < N; m++)
#pragma unroll
            for (int j = 0; j < D; j++)
#pragma unroll
              for (int v = 0; v < V; v++) gb[(j * V + v) * nS + base + m * str] += lE[m] * dGb[v][j];
        }
      }
    }
  }
  __syncthreads();
  if (VISC) {
    // the forward gradient is consumed: its region takes the trace cotangent, owned points
    // from the face adjoint, the others in phase 6
#pragma unroll
    for (int kk = 0; kk < S::FPT; kk++) {
      if (fown[kk] < 0) continue;
#pragma unroll
      for (int v = 0; v < V; v++) to[v * nF + fown[kk]] = Wtr[(e * V + v) * nF + fown[kk]];
    }
    // 5. reference-space gradient cotangent: grb[d2][v] = sum_d1 adj[d2][d1] gb[v][d1] / detJ
    //    (the metric read again here: held across the face adjoint it takes registers the
    //    Roe reverse sweep needs)
    if (valid) {
      asm volatile("" ::: "memory");
      load_metric<D>(a, e, t, nS, ij, detJ, true);
      double gv[V][D];
#pragma unroll
      for (int j = 0; j < D; j++)
#pragma unroll
        for (int v = 0; v < V; v++) gv[v][j] = gb[(j * V + v) * nS + t];
      const double inv_det = 1.0 / detJ;
#pragma unroll
      for (int d2 = 0; d2 < D; d2++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          double s = 0.0;
#pragma unroll
          for (int d1 = 0; d1 < D; d1++) s += ij[d2][d1] * gv[v][d1];
          gb[(d2 * V + v) * nS + t] = s * inv_det;
        }
    }
    __syncthreads();
    // 6. Dss^T into U, Dsf^T into the common solution's cotangent: the partner's trace at
    //    owned points (Xtr), this element's trace elsewhere
    if (valid) {
#pragma unroll
      for (int d = 0; d < D; d++) {
        const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
        const int str = (d == 0) ? 1 : ((d == 1) ? N : N * N);
        const int base = t - id * str;
#pragma unroll
        for (int v = 0; v < V; v++) Ub[v] += line_T_Dss<N>(lops, id, gb + (d * V + v) * nS, base, str);
      }
#pragma unroll
      for (int k = 0; k < S::FPT; k++) {
        const int f = t + k * nS;
        if (f >= nF) continue;
        int dim, side, base, str;
        fpt_info<D, N>(f, dim, side, base, str);
        long ep; int sl, fp;
        unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
#pragma unroll
        for (int v = 0; v < V; v++) {
          const double cb = line_Dsf<N>(lops, side, gb + (dim * V + v) * nS, base, str);
          if (sl == owner) Xtr[(e * V + v) * nF + f] += cb;
          else to[v * nF + f] = cb;  // this point's only contribution (the face adjoint wrote owned ones)
        }
      }
    }
    __syncthreads();
  }
  if (!valid) return;
  // 7. own trace cotangent E^T into v
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
    const double l0 = tab_lE<N>(lops, 0)[id], l1 = tab_lE<N>(lops, 1)[id];
    const int f0 = fpt_end<D, N>(ia, ib, ic, d, 0), f1 = fpt_end<D, N>(ia, ib, ic, d, 1);
#pragma unroll
    for (int v = 0; v < V; v++) Ub[v] += l0 * to[v * nF + f0] + l1 * to[v * nF + f1];
  }
#pragma unroll
  for (int v = 0; v < V; v++) vout[(e * V + v) * nS + t] = Ub[v];
}

template <int D, int N, bool VISC, int BM, int FC>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) __attribute__((amdgpu_waves_per_eu(2)))
void k_adj_main(const double *__restrict__ U, const double *__restrict__ w, double *__restrict__ Wtr,
                double *__restrict__ Xtr, double *__restrict__ vout, KArgs a) {
  adj_main_block<D, N, VISC, BM, FC>(U, w, Wtr, Xtr, vout, a, xcd_block(blockIdx.x, gridDim.x));
}

// The same kernel for workgroups of one wave (the viscous adjoint runs one element per
// workgroup; one wave for hex p=3): its barriers need no cross-wave synchronisation.
template <int D, int N, bool VISC, int BM, int FC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_adj_main_w64(const double *__restrict__ U, const double *__restrict__ w, double *__restrict__ Wtr,
                    double *__restrict__ Xtr, double *__restrict__ vout, KArgs a) {
  adj_main_block<D, N, VISC, BM, FC>(U, w, Wtr, Xtr, vout, a, xcd_block(blockIdx.x, gridDim.x));
}

// v += E^T (the owner's cotangent of this element's trace) at the points the partner owns.
template <int D, int N, int BM>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) void k_adj_gather(const double *__restrict__ Xtr, double *__restrict__ vout,
                                                            KArgs a) {
  using S = Shape<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  constexpr int owner = (BM == 2) ? 1 : 0;
  extern __shared__ double lds[];
  const int EB = a.EB;
  double *lops = lds;
  double *lT = lops + S::NOPS_PAD;
  const int tid = threadIdx.x, NT = blockDim.x;
  for (int i = tid; i < S::NOPS; i += NT) lops[i] = a.ops[i];
  const long e0 = xcd_block(blockIdx.x, gridDim.x) * EB;
  const long neb = min((long)EB, (long)a.E - e0);
  const int el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  double *T = lT + el * V * nF;
  if (valid) {
#pragma unroll
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      long ep; int sl, fp;
      unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
#pragma unroll
      for (int v = 0; v < V; v++)
        T[v * nF + f] = (BM != kAdjAllPoints && sl == owner) ? 0.0 : Xtr[(ep * V + v) * nF + fp];
    }
  }
  __syncthreads();
  if (!valid) return;
  const int ia = t % N, ib = (t / N) % N, ic = (D == 3) ? t / (N * N) : 0;
  double acc[V];
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = vout[(e * V + v) * nS + t];
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
    const double l0 = tab_lE<N>(lops, 0)[id], l1 = tab_lE<N>(lops, 1)[id];
    const int f0 = fpt_end<D, N>(ia, ib, ic, d, 0), f1 = fpt_end<D, N>(ia, ib, ic, d, 1);
#pragma unroll
    for (int v = 0; v < V; v++) acc[v] += l0 * T[v * nF + f0] + l1 * T[v * nF + f1];
  }
#pragma unroll
  for (int v = 0; v < V; v++) vout[(e * V + v) * nS + t] = acc[v];
}

// ---------------------------------------------------------------- general LDG beta
// With beta not +-1/2 both sides' gradients enter every common flux (src/faces.cpp:
// 1476-1614) and the common solution mixes both traces, Ucomm = (1/2 - beta) UL +
// (1/2 + beta) UR (src/faces.cpp:1366-1474).  The reverse runs after the forward trace
// and gradient-trace launches (Utr, Gtr at every flux point) and k_adj_prep:
//   k_adj_face     : per flux-point pair, by its slot-0 side: the common flux reversed for
//                    both sides' states and gradient traces; the cotangents overwrite the
//                    traces in place (Utr -> d/dUtr, Gtr -> d/dGtr);
//   k_adj_main_gen : per element: volume adjoint, the gradient-trace cotangents spread
//                    back along the normal lines, the gradient reversed into U and into the
//                    common solution, whose cotangent splits (1/2 -+ beta) between this
//                    element's trace (with the face part, E^T into v) and the partner's
//                    (Xtr, at every point);
//   k_adj_gather   : v += E^T Xtr(partner) at every point (BM = kAdjAllPoints).

template <int D, int N, int FC>
__global__ __launch_bounds__(kMaxThreads) void k_adj_face(const double *__restrict__ Wtr, KArgs a) {
  using S = Shape<D, N>;
  constexpr int nF = S::nF, V = S::V;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)a.E * nF) return;
  const long e = i / nF;
  const int f = (int)(i - e * nF);
  long ep; int sl, fp;
  unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
  if (sl != 0) return;  // each pair is reversed once, by its slot-0 (left) side
  const double *fg = a.fgeo + e * (D + 1) * nF + f;
  double n[D];
#pragma unroll
  for (int d = 0; d < D; d++) n[d] = fg[d * nF];
  double UL[V], UR[V], Fb[V], ULb[V], URb[V], G1b[V], G2b[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    UL[v] = a.Utr[(e * V + v) * nF + f];
    UR[v] = a.Utr[(ep * V + v) * nF + fp];
    // both sides' copies of the flux are the same function of (UL, UR, GL, GR)
    Fb[v] = Wtr[(e * V + v) * nF + f] + Wtr[(ep * V + v) * nF + fp];
    ULb[v] = URb[v] = G1b[v] = G2b[v] = 0.0;
  }
  common_flux_vjp<D, true, 0, FC>(UL, UR, n, 1.0, a, Fb, ULb, URb, G1b, G2b);
  // each side's viscous normal flux G = -T(U, grad) . n, reversed (gradient from Gtr)
  double GLb[V][D], GRb[V][D];
  {
    const Prim<D> pL = primitives<D>(UL, a.gamma);
    viscous_vjp<D>(
        UL, pL,
        [&](int j, double *dUj) {
#pragma unroll
          for (int v = 0; v < V; v++) dUj[v] = a.Gtr[((e * D + j) * V + v) * nF + f];
        },
        [&](int v, int j) { return -G1b[v] * n[j]; },
        [&](int j, const double *gbj) {
#pragma unroll
          for (int v = 0; v < V; v++) GLb[v][j] = gbj[v];
        },
        a, ULb);
  }
  {
    const Prim<D> pR = primitives<D>(UR, a.gamma);
    viscous_vjp<D>(
        UR, pR,
        [&](int j, double *dUj) {
#pragma unroll
          for (int v = 0; v < V; v++) dUj[v] = a.Gtr[((ep * D + j) * V + v) * nF + fp];
        },
        [&](int v, int j) { return -G2b[v] * n[j]; },
        [&](int j, const double *gbj) {
#pragma unroll
          for (int v = 0; v < V; v++) GRb[v][j] = gbj[v];
        },
        a, URb);
  }
  // in place: this pair's two entries are read and written by this thread only
#pragma unroll
  for (int v = 0; v < V; v++) {
    a.Utr[(e * V + v) * nF + f] = ULb[v];
    a.Utr[(ep * V + v) * nF + fp] = URb[v];
  }
#pragma unroll
  for (int j = 0; j < D; j++)
#pragma unroll
    for (int v = 0; v < V; v++) {
      a.Gtr[((e * D + j) * V + v) * nF + f] = GLb[v][j];
      a.Gtr[((ep * D + j) * V + v) * nF + fp] = GRb[v][j];
    }
}

// LDS per element: U [V][nS]; one region for the common solution, then the forward
// gradient [D][V][nS], then the trace cotangent [V][nF]; the gradient cotangent [D][V][nS].
template <int D, int N>
struct AdjGenLds {
  using S = Shape<D, N>;
  static constexpr int R = (S::V * S::nF > D * S::V * S::nS) ? S::V * S::nF : D * S::V * S::nS;
  static constexpr int per_ele = S::V * S::nS + R + D * S::V * S::nS;  // doubles
};

template <int D, int N>
__device__ __forceinline__ void adj_main_general_block(const double *__restrict__ U, const double *__restrict__ w,
                                                       double *__restrict__ Xtr, double *__restrict__ vout,
                                                       const KArgs &a, long bid) {
  using S = Shape<D, N>;
  using L = AdjGenLds<D, N>;
  constexpr int nS = S::nS, nF = S::nF, V = S::V;
  extern __shared__ double lds[];
  const int EB = a.EB;
  double *lops = lds;
  double *lU = lops + S::NOPS_PAD;
  double *lR = lU + EB * V * nS;
  double *lGB = lR + EB * L::R;
  const long e0 = bid * EB;
  const long neb = min((long)EB, (long)a.E - e0);
  stage_U<D, N>(U, a, lops, lU, e0, neb, V * nS);
  __syncthreads();
  const int tid = threadIdx.x, el = tid / nS, t = tid - el * nS;
  const long e = e0 + el;
  const bool valid = el < neb;
  const double *u = lU + el * V * nS;
  const double *we = w + e * V * nS;
  double *c = lR + el * L::R, *g_l = c, *to = c;
  double *gb = lGB + el * D * V * nS;
  const int ia = t % N, ib = (t / N) % N, ic = (D == 3) ? t / (N * N) : 0;
  const double beta = a.ldg_b;
  // 1. common solution, the partner's trace extrapolated from its U (Utr now holds cotangents)
  if (valid) {
#pragma unroll
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
      long ep; int sl, fp;
      unpack_nbr<D, N>(a.nbr[e * nF + f], ep, sl, fp);
      double pt[V];
      partner_trace<D, N>(U, a.ops, ep, fp, pt);
#pragma unroll
      for (int v = 0; v < V; v++) {
        const double own = line_extrap<N>(lE, u + v * nS, base, str);
        const double l = sl ? pt[v] : own, r = sl ? own : pt[v];
        c[v * nF + f] = 0.5 * (l + r) - beta * (l - r);
      }
    }
  }
  __syncthreads();
  double ij[D][D], detJ = 1.0;
  if (valid) load_metric<D>(a, e, t, nS, ij, detJ, true);
  // 2. forward physical gradient, into the common solution's region
  {
    double g[V][D];
    if (valid) spt_gradient<D, N>(lops, u, c, t, ij, detJ, g);
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) g_l[(d * V + v) * nS + t] = g[v][d];
    }
    __syncthreads();
  }
  // 3. volume adjoint: tFb[d1] = Dss^T w along the d1 lines; F(U, grad) reversed
  double Ub[V];
#pragma unroll
  for (int v = 0; v < V; v++) Ub[v] = 0.0;
  if (valid) {
    double Us[V];
#pragma unroll
    for (int v = 0; v < V; v++) Us[v] = u[v * nS + t];
    const Prim<D> p = primitives<D>(Us, a.gamma);
    double tFb[D][V];
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
      const int str = (d == 0) ? 1 : ((d == 1) ? N : N * N);
      const int base = t - id * str;
#pragma unroll
      for (int v = 0; v < V; v++) tFb[d][v] = line_T_Dss<N>(lops, id, we + v * nS, base, str);
      conv_proj_vjp<D>(Us, p, ij[d], tFb[d], a.gamma, Ub);
    }
    viscous_vjp<D>(
        Us, p,
        [&](int j, double *dUj) {
#pragma unroll
          for (int v = 0; v < V; v++) dUj[v] = g_l[(j * V + v) * nS + t];
        },
        [&](int v, int j) {
          double s = 0.0;
#pragma unroll
          for (int d1 = 0; d1 < D; d1++) s += tFb[d1][v] * ij[d1][j];
          return -s;
        },
        [&](int j, const double *gbj) {
#pragma unroll
          for (int v = 0; v < V; v++) gb[(j * V + v) * nS + t] = gbj[v];
        },
        a, Ub);
  }
  // 4. the gradient-trace cotangents (k_adj_face, in Gtr) spread back along the normal
  //    lines, face by face (lines of one face never meet, so the order is fixed)
#pragma unroll 1
  for (int q = 0; q < 2 * D; q++) {
    __syncthreads();
    if (!valid) continue;
#pragma unroll
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF || f / S::nFF != q) continue;
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const double *lE = tab_lE<N>(lops, side);
#pragma unroll
      for (int j = 0; j < D; j++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          const double gt = a.Gtr[((e * D + j) * V + v) * nF + f];
#pragma unroll
          for (int m = 0; m < N; m++) gb[(j * V + v) * nS + base + m * str] += lE[m] * gt;
        }
    }
  }
  __syncthreads();
  // 5. reference-space gradient cotangent: grb[d2][v] = sum_d1 adj[d2][d1] gb[v][d1] / detJ
  if (valid) {
    double gv[V][D];
#pragma unroll
    for (int j = 0; j < D; j++)
#pragma unroll
      for (int v = 0; v < V; v++) gv[v][j] = gb[(j * V + v) * nS + t];
    const double inv_det = 1.0 / detJ;
#pragma unroll
    for (int d2 = 0; d2 < D; d2++)
#pragma unroll
      for (int v = 0; v < V; v++) {
        double s = 0.0;
#pragma unroll
        for (int d1 = 0; d1 < D; d1++) s += ij[d2][d1] * gv[v][d1];
        gb[(d2 * V + v) * nS + t] = s * inv_det;
      }
  }
  __syncthreads();
  // 6. Dss^T into U; Dsf^T into the common solution's cotangent, split between this
  //    element's trace (added to its face cotangent from Utr) and the partner's (Xtr)
  if (valid) {
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
      const int str = (d == 0) ? 1 : ((d == 1) ? N : N * N);
      const int base = t - id * str;
#pragma unroll
      for (int v = 0; v < V; v++) Ub[v] += line_T_Dss<N>(lops, id, gb + (d * V + v) * nS, base, str);
    }
#pragma unroll
    for (int k = 0; k < S::FPT; k++) {
      const int f = t + k * nS;
      if (f >= nF) continue;
      int dim, side, base, str;
      fpt_info<D, N>(f, dim, side, base, str);
      const int sl = nbr_slot<D, N>(a.nbr[e * nF + f]);
      const double ao = sl ? (0.5 + beta) : (0.5 - beta), ap = sl ? (0.5 - beta) : (0.5 + beta);
#pragma unroll
      for (int v = 0; v < V; v++) {
        const double cb = line_Dsf<N>(lops, side, gb + (dim * V + v) * nS, base, str);
        to[v * nF + f] = a.Utr[(e * V + v) * nF + f] + ao * cb;
        Xtr[(e * V + v) * nF + f] = ap * cb;
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  // 7. own trace cotangent E^T into v
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int id = (d == 0) ? ia : ((d == 1) ? ib : ic);
    const double l0 = tab_lE<N>(lops, 0)[id], l1 = tab_lE<N>(lops, 1)[id];
    const int f0 = fpt_end<D, N>(ia, ib, ic, d, 0), f1 = fpt_end<D, N>(ia, ib, ic, d, 1);
#pragma unroll
    for (int v = 0; v < V; v++) Ub[v] += l0 * to[v * nF + f0] + l1 * to[v * nF + f1];
  }
#pragma unroll
  for (int v = 0; v < V; v++) vout[(e * V + v) * nS + t] = Ub[v];
}

template <int D, int N>
__global__ __launch_bounds__((Shape<D, N>::kThreads)) __attribute__((amdgpu_waves_per_eu(2)))
void k_adj_main_gen(const double *__restrict__ U, const double *__restrict__ w, double *__restrict__ Xtr,
                    double *__restrict__ vout, KArgs a) {
  adj_main_general_block<D, N>(U, w, Xtr, vout, a, xcd_block(blockIdx.x, gridDim.x));
}
