// kernels_shape.hip -- the kernel launches of one element shape: compiled once per
// (nDims, points per direction) with -DZK_D=<2|3> -DZK_N=<1..10 for quads, 1..9 for
// hexes>, so the shapes build in parallel and a change to one shape's code rebuilds one
// object.
#include "kernels.hpp"

#if !defined(ZK_D) || !defined(ZK_N)
#error "kernels_shape.hip is compiled with -DZK_D=<nDims> -DZK_N=<order + 1>"
#endif

namespace zk {

// k_divergence for one convective flux (FC: 0 Rusanov, 1 Roe) and the LDG beta branch.
// Euler: the residual extrapolates the partner traces from U itself (SELF), unless a
// fused epilogue writes U in place (the RK update, or the copy of the starting state):
// other workgroups may then still be reading it, so the traces come from Utr.
template <int D, int N, int FC>
void launch_div(const zfr_ele_desc_t *d, dim3 grid, dim3 block, size_t lds, hipStream_t st, const double *U,
                double *out, const KArgs &a) {
  if (!d->viscous && block.x <= 64) {  // one wave per workgroup
    if (rk_writes_input(a.rk, U))
      hipLaunchKernelGGL((k_divergence_w64<D, N, false, 0, FC, false>), grid, block, lds, st, U, out, a);
    else
      hipLaunchKernelGGL((k_divergence_w64<D, N, false, 0, FC, true>), grid, block, lds, st, U, out, a);
  } else if (!d->viscous && rk_writes_input(a.rk, U))
    hipLaunchKernelGGL((k_divergence<D, N, false, 0, FC, false>), grid, block, lds, st, U, out, a);
  else if (!d->viscous) hipLaunchKernelGGL((k_divergence<D, N, false, 0, FC, true>), grid, block, lds, st, U, out, a);
  else if (d->ldg_b == 0.5) hipLaunchKernelGGL((k_divergence<D, N, true, 1, FC>), grid, block, lds, st, U, out, a);
  else if (d->ldg_b == -0.5) hipLaunchKernelGGL((k_divergence<D, N, true, 2, FC>), grid, block, lds, st, U, out, a);
  else hipLaunchKernelGGL((k_divergence<D, N, true, 0, FC>), grid, block, lds, st, U, out, a);
}

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

// Shapes with the adjoint main kernels (not hexes of order 8: 199 KiB of LDS per element).
template <int D, int N>
constexpr bool kAdjoint = !(D == 3 && N > 8);

template <int D, int N>
int launch(Op op, const zfr_ele_desc_t *d, const double *U, double *out, hipStream_t st, const AdjPtrs &j) {
  using S = Shape<D, N>;
  if (d->nSpts != S::nS || d->nFpts != S::nF || d->nVars != S::V)
    return zfr::fail("kernel launch: element descriptor does not match (nDims, order)");
  if (d->nEles <= 0) return 0;
  const size_t per_ele = lds_per_element<D, N>(op, d->viscous != 0);
  const bool general = d->viscous && d->ldg_b != 0.5 && d->ldg_b != -0.5;
  if (op == Op::AdjMain && general) return zfr::fail("k_adj_main is the one-sided LDG adjoint (beta = +-0.5)");
  if ((op == Op::AdjMain || op == Op::AdjMainGen) && !kAdjoint<D, N>)
    return zfr::fail("the residual VJP covers hexes to order 7: at order 8 one element's adjoint needs more LDS "
                     "than a workgroup can allocate");
  if ((op == Op::AdjFace || op == Op::AdjMainGen) && !general)
    return zfr::fail("the general-beta adjoint launches need a viscous descriptor with beta != +-0.5");
  if ((op == Op::AdjFace || op == Op::AdjMainGen) && (!d->Utr || !d->Gtr))
    return zfr::fail("the general-beta adjoint needs the trace buffers Utr and Gtr");
  if (d->fconv_type != 0 && d->fconv_type != 2) return zfr::fail("fconv_type must be 0 (Rusanov) or 2 (Roe)");
  if ((op == Op::OneSided || op == Op::OneSidedFromTrace || op == Op::FaceCorr) && d->ldg_b != 0.5 &&
      d->ldg_b != -0.5)
    return zfr::fail("one-sided LDG kernels need ldg_b = +-0.5");
  // the viscous adjoint runs one element per workgroup: its LDS then allows ~7
  // workgroups per CU, more than any larger tile
  const int EB = (op == Op::AdjMain && d->viscous) ? 1 : eles_per_block(d, S::defaultEB(), S::NOPS_PAD, per_ele, !g_persist);
  if (sizeof(double) * (S::NOPS_PAD + EB * per_ele) > max_lds<D, N>())
    return zfr::fail("kernel launch: one element needs more LDS than a workgroup may use (order too high)");
  KArgs a = make_args(d, EB);
  if (j.rk) {
    if (op != Op::Div && op != Op::FaceCorr) return zfr::fail("kernel launch: RK fusion is for the last kernel only");
    a.rk = *j.rk;
  }
  const dim3 grid((unsigned)((d->nEles + EB - 1) / EB)), block((unsigned)(EB * S::nS));
  const size_t lds = sizeof(double) * (S::NOPS_PAD + EB * per_ele);
  if (g_persist) {  // recording a persistent plan: append the phase instead of launching
    if (op != Op::Div || d->viscous)
      return zfr::fail("persistent plan: only Euler divergence launches (residuals and fused RK stages of an "
                       "inviscid run) can be recorded");
    if (N > kMixedMaxN) return zfr::fail("persistent plan: the persistent kernel covers orders 0..5");
    if (g_persist->D != 0 && g_persist->D != D) return zfr::fail("persistent plan: one nDims per plan");
    g_persist->phases.push_back({a, U, out, (long)grid.x, N, D, d->fconv_type == 2 ? 1 : 0,
                                 rk_writes_input(a.rk, U) ? 0 : 1});
    g_persist->D = D;
    if (lds > g_persist->lds) g_persist->lds = lds;
    return 0;
  }
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
      if constexpr (kAdjoint<D, N>) {
        if (d->fconv_type == 2) launch_adj_main<D, N, 1>(d, grid, block, lds, st, U, j, out, a);
        else launch_adj_main<D, N, 0>(d, grid, block, lds, st, U, j, out, a);
      }
      break;
    case Op::AdjGather:
      if (general) hipLaunchKernelGGL((k_adj_gather<D, N, kAdjAllPoints>), grid, block, lds, st, j.Xtr, out, a);
      else if (d->viscous && d->ldg_b == -0.5) hipLaunchKernelGGL((k_adj_gather<D, N, 2>), grid, block, lds, st, j.Xtr, out, a);
      else hipLaunchKernelGGL((k_adj_gather<D, N, 0>), grid, block, lds, st, j.Xtr, out, a);
      break;
    case Op::AdjFace: {  // one thread per (element, flux point); slot-0 points work
      const long total = (long)d->nEles * S::nF;
      const dim3 fgrid((unsigned)((total + kMaxThreads - 1) / kMaxThreads)), fblock(kMaxThreads);
      if (d->fconv_type == 2) hipLaunchKernelGGL((k_adj_face<D, N, 1>), fgrid, fblock, 0, st, j.Wtr, a);
      else hipLaunchKernelGGL((k_adj_face<D, N, 0>), fgrid, fblock, 0, st, j.Wtr, a);
      break;
    }
    case Op::AdjMainGen:
      if constexpr (kAdjoint<D, N>) hipLaunchKernelGGL((k_adj_main_gen<D, N>), grid, block, lds, st, U, j.w, j.Xtr, out, a);
      break;
  }
  return zfr::check_hip(hipGetLastError(), "kernel launch");
}

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

template int launch<ZK_D, ZK_N>(Op, const zfr_ele_desc_t *, const double *, double *, hipStream_t, const AdjPtrs &);
template void launch_multi<ZK_D, ZK_N>(const MultiGroup &, hipStream_t);

}  // namespace zk
