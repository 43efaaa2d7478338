// physics_host.cpp -- host builds of the pointwise physics in physics.hpp (the same
// templates the kernels instantiate), exported so the forward fluxes can be compared with
// the reference's formulas and the hand-written reverse-mode derivatives with finite
// differences, on the CPU.

#include <cstring>

#include "common.hpp"
#include "physics.hpp"
#include "zefr_amd.h"

namespace {

using namespace zphys;

// Full flux tensor F[v][d] = Fconv - Fvisc (include/flux.hpp:15-202) from the projected
// forms with n = e_d.
template <int D>
void point_flux(const zfr_phys_t &a, const double *U, const double *dU, double *F) {
  constexpr int V = D + 2;
  const Prim<D> p = primitives<D>(U, a.gamma);
  Visc<D> s{};
  if (dU) s = viscous<D>(U, p, reinterpret_cast<const double(*)[D]>(dU), a);
  for (int d = 0; d < D; d++) {
    double n[D] = {};
    n[d] = 1.0;
    double Fn[V];
    conv_proj<D>(U, p, n, Fn);
    if (dU) visc_proj_sub<D>(s, n, Fn);
    for (int v = 0; v < V; v++) F[v * D + d] = Fn[v];
  }
}

template <int D>
void point_flux_vjp(const zfr_phys_t &a, const double *U, const double *dU, const double *Fb, double *Ub,
                    double *dUb) {
  constexpr int V = D + 2;
  const Prim<D> p = primitives<D>(U, a.gamma);
  for (int v = 0; v < V; v++) Ub[v] = 0.0;
  for (int d = 0; d < D; d++) {
    double n[D] = {}, Fbd[V];
    n[d] = 1.0;
    for (int v = 0; v < V; v++) Fbd[v] = Fb[v * D + d];
    conv_proj_vjp<D>(U, p, n, Fbd, a.gamma, Ub);
  }
  if (!dU) return;
  // F = Fconv - T with T the viscous tensor: cotangent on T is -Fb
  viscous_vjp<D>(
      U, p, [&](int j, double *dUj) { for (int v = 0; v < V; v++) dUj[v] = dU[v * D + j]; },
      [&](int v, int j) { return -Fb[v * D + j]; },
      [&](int j, const double *g) { for (int v = 0; v < V; v++) dUb[v * D + j] = g[v]; }, a, Ub);
}

// One side's viscous normal flux from its gradient, as the kernels form it.
template <int D>
void side_visc(const zfr_phys_t &a, const double *U, const double *dU, const double *n, double *G) {
  visc_normal<D>(U, [&](int j, double *dUj) { for (int v = 0; v < D + 2; v++) dUj[v] = dU[v * D + j]; }, n, a, G);
}

template <int D, int BM, int FC>
void common(const zfr_phys_t &a, int viscous, const double *UL, const double *UR, const double *dUL,
            const double *dUR, const double *n, double dAs, double *F) {
  constexpr int V = D + 2;
  if (!viscous) {
    common_flux<D, false, 0, FC>(UL, UR, nullptr, nullptr, n, dAs, a, F);
    return;
  }
  double GL[V], GR[V];
  if (BM != 2) side_visc<D>(a, UL, dUL, n, GL);
  if (BM != 1) side_visc<D>(a, UR, dUR, n, GR);
  if (BM == 0) common_flux<D, true, 0, FC>(UL, UR, GL, GR, n, dAs, a, F);
  else if (BM == 1) common_flux<D, true, 1, FC>(UL, UR, GL, nullptr, n, dAs, a, F);
  else common_flux<D, true, 2, FC>(UL, UR, GR, nullptr, n, dAs, a, F);
}

template <int D, int BM, int FC>
void common_vjp(const zfr_phys_t &a, int viscous, const double *UL, const double *UR, const double *dUL,
                const double *dUR, const double *n, double dAs, const double *Fb, double *ULb, double *URb,
                double *dULb, double *dURb) {
  constexpr int V = D + 2;
  for (int v = 0; v < V; v++) ULb[v] = URb[v] = 0.0;
  if (!viscous) {
    common_flux_vjp<D, false, 0, FC>(UL, UR, n, dAs, a, Fb, ULb, URb, nullptr, nullptr);
    return;
  }
  double G1b[V] = {}, G2b[V] = {};
  common_flux_vjp<D, true, BM, FC>(UL, UR, n, dAs, a, Fb, ULb, URb, G1b, G2b);
  // G = -T . n for the side whose gradient enters: cotangent on T[v][j] is -Gb[v] n_j
  auto side = [&](const double *U, const double *dU, const double *Gb, double *Ub, double *dUb) {
    const Prim<D> p = primitives<D>(U, a.gamma);
    viscous_vjp<D>(
        U, p, [&](int j, double *dUj) { for (int v = 0; v < V; v++) dUj[v] = dU[v * D + j]; },
        [&](int v, int j) { return -Gb[v] * n[j]; },
        [&](int j, const double *g) { for (int v = 0; v < V; v++) dUb[v * D + j] += g[v]; }, a, Ub);
  };
  for (int i = 0; i < V * D; i++) dULb[i] = dURb[i] = 0.0;
  if (BM == 0) {
    side(UL, dUL, G1b, ULb, dULb);
    side(UR, dUR, G2b, URb, dURb);
  } else if (BM == 1) {
    side(UL, dUL, G1b, ULb, dULb);
  } else {
    side(UR, dUR, G1b, URb, dURb);
  }
}

int bm_of(double beta) { return beta == 0.5 ? 1 : (beta == -0.5 ? 2 : 0); }

}  // namespace

extern "C" {

int zfr_host_point_flux(int nDims, const zfr_phys_t *a, const double *U, const double *dU, double *F) {
  return zfr::guard([&] {
    if (nDims == 2) point_flux<2>(*a, U, dU, F);
    else if (nDims == 3) point_flux<3>(*a, U, dU, F);
    else zfr::raise("nDims must be 2 or 3");
  });
}

int zfr_host_point_flux_vjp(int nDims, const zfr_phys_t *a, const double *U, const double *dU, const double *Fb,
                            double *Ub, double *dUb) {
  return zfr::guard([&] {
    if (nDims == 2) point_flux_vjp<2>(*a, U, dU, Fb, Ub, dUb);
    else if (nDims == 3) point_flux_vjp<3>(*a, U, dU, Fb, Ub, dUb);
    else zfr::raise("nDims must be 2 or 3");
  });
}

int zfr_host_common_flux(int nDims, const zfr_phys_t *a, int viscous, const double *UL, const double *UR,
                         const double *dUL, const double *dUR, const double *n, double dAs, double *F) {
  return zfr::guard([&] {
    const int bm = bm_of(a->ldg_b);
    if (nDims != 2 && nDims != 3) zfr::raise("nDims must be 2 or 3");
    if (a->fconv_type != 0 && a->fconv_type != 2) zfr::raise("fconv_type must be 0 (Rusanov) or 2 (Roe)");
    const bool roe = a->fconv_type == 2;
    using Fn = void (*)(const zfr_phys_t &, int, const double *, const double *, const double *, const double *,
                        const double *, double, double *);
    static const Fn table[2][2][3] = {
        {{common<2, 0, 0>, common<2, 1, 0>, common<2, 2, 0>}, {common<2, 0, 1>, common<2, 1, 1>, common<2, 2, 1>}},
        {{common<3, 0, 0>, common<3, 1, 0>, common<3, 2, 0>}, {common<3, 0, 1>, common<3, 1, 1>, common<3, 2, 1>}}};
    table[nDims - 2][roe ? 1 : 0][bm](*a, viscous, UL, UR, dUL, dUR, n, dAs, F);
  });
}

int zfr_host_common_flux_vjp(int nDims, const zfr_phys_t *a, int viscous, const double *UL, const double *UR,
                             const double *dUL, const double *dUR, const double *n, double dAs, const double *Fb,
                             double *ULb, double *URb, double *dULb, double *dURb) {
  return zfr::guard([&] {
    if (nDims != 2 && nDims != 3) zfr::raise("nDims must be 2 or 3");
    if (a->fconv_type != 0 && a->fconv_type != 2) zfr::raise("fconv_type must be 0 (Rusanov) or 2 (Roe)");
    const int bm = bm_of(a->ldg_b);
    using Fn = void (*)(const zfr_phys_t &, int, const double *, const double *, const double *, const double *,
                        const double *, double, const double *, double *, double *, double *, double *);
    static const Fn table[2][2][3] = {
        {{common_vjp<2, 0, 0>, common_vjp<2, 1, 0>, common_vjp<2, 2, 0>},
         {common_vjp<2, 0, 1>, common_vjp<2, 1, 1>, common_vjp<2, 2, 1>}},
        {{common_vjp<3, 0, 0>, common_vjp<3, 1, 0>, common_vjp<3, 2, 0>},
         {common_vjp<3, 0, 1>, common_vjp<3, 1, 1>, common_vjp<3, 2, 1>}}};
    table[nDims - 2][a->fconv_type == 2 ? 1 : 0][bm](*a, viscous, UL, UR, dUL, dUR, n, dAs, Fb, ULb, URb, dULb,
                                                     dURb);
  });
}

}  // extern "C"
