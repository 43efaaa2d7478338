// kernels.hip -- shape-independent kernels (RK updates, p-multigrid transfers, residual
// norm, stream copy) and the kernel-level C entry points of zefr_amd.h.  The residual
// kernels themselves live in kernels.hpp and are launched from kernels_shape.hip.

#include "kernels.hpp"

#include <algorithm>
#include <memory>
#include <mutex>

namespace zk {

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
constexpr int kPmgTab = 128;  // doubles for the [n_out][n_in] table (n <= 10)
constexpr int kPmgLds = kPmgTab + 2 * kPmgBuf;

template <int D>
__global__ __launch_bounds__(kMaxThreads) void k_pmg_transfer(const double *__restrict__ in, double *__restrict__ out,
                                                              const double *__restrict__ A, long lines, int ni,
                                                              int no, int L, double alpha, double beta) {
  extern __shared__ double lds[];  // kPmgLds doubles
  double *lA = lds, *buf = lds + kPmgTab;
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
      // explicit fmas: the p-multigrid epilogues of the residual kernels (tp_out,
      // pmg_epilogue) form the same numbers
      double acc = 0.0;
      for (int i = 0; i < ni; i++) acc = __builtin_fma(lA[o * ni + i], s[i * inner], acc);
      if (d == D - 1) {  // outer == 1: k runs over [nl][no^D] in output order
        double *g = out + l0 * szo + k;
        *g = (beta == 0.0) ? alpha * acc : __builtin_fma(alpha, acc, beta * *g);
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

thread_local PersistRecorder *g_persist = nullptr;

int refuse_while_recording(const char *what) {
  if (!g_persist) return 0;
  return zfr::fail(std::string("persistent plan: ") + what +
                   " cannot be recorded (a plan holds Euler divergence launches only)");
}

int dispatch(Op op, const zfr_ele_desc_t *d, const double *U, double *out, hipStream_t st, const AdjPtrs &j = AdjPtrs()) {
  if (op != Op::Div)
    if (int rc = refuse_while_recording("a launch other than the Euler divergence")) return rc;
  const int N = d->order + 1;
  if (d->nDims == 2) {
    switch (N) {
      case 1: return launch<2, 1>(op, d, U, out, st, j);
      case 2: return launch<2, 2>(op, d, U, out, st, j);
      case 3: return launch<2, 3>(op, d, U, out, st, j);
      case 4: return launch<2, 4>(op, d, U, out, st, j);
      case 5: return launch<2, 5>(op, d, U, out, st, j);
      case 6: return launch<2, 6>(op, d, U, out, st, j);
      case 7: return launch<2, 7>(op, d, U, out, st, j);
      case 8: return launch<2, 8>(op, d, U, out, st, j);
      case 9: return launch<2, 9>(op, d, U, out, st, j);
      case 10: return launch<2, 10>(op, d, U, out, st, j);
    }
  } else if (d->nDims == 3) {
    switch (N) {
      case 1: return launch<3, 1>(op, d, U, out, st, j);
      case 2: return launch<3, 2>(op, d, U, out, st, j);
      case 3: return launch<3, 3>(op, d, U, out, st, j);
      case 4: return launch<3, 4>(op, d, U, out, st, j);
      case 5: return launch<3, 5>(op, d, U, out, st, j);
      case 6: return launch<3, 6>(op, d, U, out, st, j);
      case 7: return launch<3, 7>(op, d, U, out, st, j);
      case 8: return launch<3, 8>(op, d, U, out, st, j);
      case 9: return launch<3, 9>(op, d, U, out, st, j);
    }
  }
  return zfr::fail("kernel launch: supported shapes are quads of order 0..9 and hexes of order 0..8");
}

template <int D>
void dispatch_multi_n(const MultiGroup &g, hipStream_t st) {
  if (g.N == 0) {
    launch_mixed<D>(g, st);
    return;
  }
  switch (g.N) {
    case 1: launch_multi<D, 1>(g, st); break;
    case 2: launch_multi<D, 2>(g, st); break;
    case 3: launch_multi<D, 3>(g, st); break;
    case 4: launch_multi<D, 4>(g, st); break;
    case 5: launch_multi<D, 5>(g, st); break;
    case 6: launch_multi<D, 6>(g, st); break;
  }
  switch (g.N) {
    case 7: launch_multi<D, 7>(g, st); break;
    case 8: launch_multi<D, 8>(g, st); break;
    case 9: launch_multi<D, 9>(g, st); break;
  }
  if constexpr (D == 2) {
    if (g.N == 10) launch_multi<2, 10>(g, st);
  }
}

inline unsigned grid_for(long total) {
  long g = (total + kMaxThreads - 1) / kMaxThreads;
  if (g > 8192) g = 8192;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace zk

namespace {
// Buffers a launch reads or writes through the descriptor: a null one is an error here,
// not a fault on the device.
int need(const void *p, const char *what) {
  return p ? 0 : zfr::fail(std::string("kernel launch: the descriptor's ") + what + " is NULL");
}
// A pointer argument of the entry point itself.
int need_arg(const void *p, const char *what) {
  return p ? 0 : zfr::fail(std::string("kernel launch: ") + what + " is NULL");
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

extern "C" {

int zfr_extrapolate(const zfr_ele_desc_t *d, const double *U, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need_arg(U, "U");
  if (!rc) rc = need(d->Utr, "Utr");
  if (rc) return rc;
  return zk::dispatch(zk::Op::Extrap, d, U, nullptr, (hipStream_t)stream);
}

int zfr_extrapolate_partner(const zfr_ele_desc_t *d, const double *U, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need_arg(U, "U");
  if (!rc) rc = need(d->Utr, "Utr");
  if (rc) return rc;
  if (!d->viscous || (d->ldg_b != 0.5 && d->ldg_b != -0.5))
    return zfr::fail("zfr_extrapolate_partner: the descriptor is not on the one-sided LDG path (beta = +-1/2)");
  return zk::dispatch(zk::Op::ExtrapPartner, d, U, nullptr, (hipStream_t)stream);
}

int zfr_gradient(const zfr_ele_desc_t *d, const double *U, double *dU_out, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need_arg(U, "U");
  if (!rc) rc = need(d->Utr, "Utr");
  if (!rc) rc = need(d->Gtr, "Gtr");
  if (rc) return rc;
  if (!d->viscous) return zfr::fail("zfr_gradient: descriptor is inviscid");
  return zk::dispatch(zk::Op::Grad, d, U, dU_out, (hipStream_t)stream);
}

int zfr_divergence(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need_arg(U, "U");
  if (!rc) rc = need_arg(divF, "divF");
  if (!rc && d->viscous) rc = need(d->Utr, "Utr");
  if (!rc && d->viscous && d->ldg_b != 0.5 && d->ldg_b != -0.5) rc = need(d->Gtr, "Gtr");
  if (rc) return rc;
  return zk::dispatch(zk::Op::Div, d, U, divF, (hipStream_t)stream);
}

int zfr_gradient_flux(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need_arg(U, "U");
  if (!rc) rc = need_arg(divF, "divF");
  if (!rc) rc = need(d->Utr, "Utr");
  if (!rc) rc = need(d->Ftr, "Ftr");
  if (rc) return rc;
  if (!d->viscous) return zfr::fail("zfr_gradient_flux: descriptor is inviscid");
  return zk::dispatch(zk::Op::OneSided, d, U, divF, (hipStream_t)stream);
}

int zfr_gradient_flux_traced(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need_arg(U, "U");
  if (!rc) rc = need_arg(divF, "divF");
  if (!rc) rc = need(d->Utr, "Utr");
  if (!rc) rc = need(d->Ftr, "Ftr");
  if (rc) return rc;
  if (!d->viscous) return zfr::fail("zfr_gradient_flux_traced: descriptor is inviscid");
  return zk::dispatch(zk::Op::OneSidedFromTrace, d, U, divF, (hipStream_t)stream);
}

int zfr_face_correction(const zfr_ele_desc_t *d, double *divF, void *stream) {
  int rc = need_geometry(d);
  if (!rc) rc = need_arg(divF, "divF");
  if (!rc) rc = need(d->Ftr, "Ftr");
  if (rc) return rc;
  if (!d->viscous) return zfr::fail("zfr_face_correction: descriptor is inviscid");
  return zk::dispatch(zk::Op::FaceCorr, d, nullptr, divF, (hipStream_t)stream);
}

}  // extern "C"

namespace {
int rk_fused(zk::Op op, const zfr_ele_desc_t *d, const double *U, double *divF, const zfr_rk_args_t *r,
             hipStream_t st) {
  if (!r) return zfr::fail("RK fusion: missing buffer");
  const bool restr = r->restrict_U || r->restrict_divF || r->restrict_A;
  const bool prol = r->prolong_out || r->prolong_A;
  if (r->mode < 0 || r->mode > 4 || (r->mode == 0 && !restr))
    return zfr::fail("RK fusion: mode must be 1 (stage), 2 (final), 3 (source -= divF), 4 (divF += source) "
                     "or 0 (residual) with a restriction");
  if (restr) {
    if (r->mode != 0 && r->mode != 4) return zfr::fail("RK fusion: the restriction goes with modes 0 and 4");
    if (!r->restrict_U || !r->restrict_divF || !r->restrict_A)
      return zfr::fail("RK fusion: restriction needs restrict_A, restrict_U and restrict_divF");
    if (r->restrict_n < 1 || r->restrict_n > d->order + 1)
      return zfr::fail("RK fusion: restrict_n must be in 1..order+1");
    if (op == zk::Op::FaceCorr && !r->restrict_src)
      return zfr::fail("RK fusion: the face correction restricts U from restrict_src");
  }
  if (prol) {
    if (r->mode != 2 || !r->corr_out) return zfr::fail("RK fusion: the prolongation goes with mode 2 and corr_out");
    if (!r->prolong_out || !r->prolong_A) return zfr::fail("RK fusion: prolongation needs prolong_A and prolong_out");
    if (!zfr::pmg_prolong_fits(d->order + 1, r->prolong_n))
      return zfr::fail("RK fusion: prolong_n must be in order+1..10 with a [prolong_n][order+1] table that fits "
                       "the kernel's operator-table region (prolong_n (order+1) <= (order+1)^2 + 4 (order+1), "
                       "padded to 8); otherwise prolong with zfr_pmg_transfer");
  }
  const bool update = r->mode == 1 || r->mode == 2;
  if (update && (!r->U || !r->U_ini || (r->mode == 2 && r->nStages > 1 && !r->divF0)))
    return zfr::fail("RK fusion: missing buffer");
  if (r->mode == 3 && !r->source_out) return zfr::fail("RK fusion: mode 3 needs source_out");
  if (r->source_out && r->mode != 3 && !(r->mode == 1 && r->source))
    return zfr::fail("RK fusion: source_out is for mode 3, or for mode 1 with source (the stage forms the source "
                     "term first)");
  if (r->mode == 4 && !r->source) return zfr::fail("RK fusion: mode 4 needs source");
  if (r->U_ini_out && !update) return zfr::fail("RK fusion: U_ini_out is for RK updates (modes 1, 2)");
  if (r->U_ini_out && r->U_ini_out == r->U_ini) return zfr::fail("RK fusion: U_ini_out must not be U_ini");
  if (r->corr_out && (r->mode != 2 || !r->corr_base))
    return zfr::fail("RK fusion: corr_out is for final combinations (mode 2) and needs corr_base");
  if (r->trace_out && r->mode != 1) return zfr::fail("RK fusion: trace_out is for stage updates (mode 1)");
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
  rk.trace = r->trace_out;
  rk.Uini_out = r->U_ini_out;
  rk.corr = r->corr_out;
  rk.corr_base = r->corr_base;
  rk.srcw = r->source_out;
  if (restr) {
    rk.pmgR = r->restrict_A;
    rk.pmgRn = r->restrict_n;
    rk.pmgUc = r->restrict_U;
    rk.pmgRc = r->restrict_divF;
    rk.pmgU = r->restrict_src;
  }
  if (prol) {
    rk.pmgP = r->prolong_A;
    rk.pmgPn = r->prolong_n;
    rk.pmgAlpha = r->prolong_alpha;
    rk.pmgUf = r->prolong_out;
  }
  // the divergence reads the trace when viscous, or when its epilogue writes the state it
  // reads (Euler then reads the partners' traces instead of their U)
  const bool reads_tr = op == zk::Op::Div && (d->viscous || zk::rk_writes_input(rk, U));
  int rc = need_geometry(d);
  if (!rc && reads_tr) rc = need(d->Utr, "Utr");
  if (!rc && op == zk::Op::Div && d->viscous && d->ldg_b != 0.5 && d->ldg_b != -0.5) rc = need(d->Gtr, "Gtr");
  if (!rc && op == zk::Op::FaceCorr) rc = need(d->Ftr, "Ftr");
  if (rc) return rc;
  if (update && (r->nStages < 1 || r->nStages > zk::kMaxRkStages || r->stage < 0 || r->stage >= r->nStages))
    return zfr::fail("RK fusion: stage out of range");
  if (r->trace_out && reads_tr && r->trace_out == d->Utr)
    return zfr::fail("RK fusion: trace_out must not be the trace the divergence kernel reads (d->Utr)");
  zk::AdjPtrs j;
  j.rk = &rk;
  return zk::dispatch(op, d, U, divF, st, j);
}
}  // namespace

extern "C" {

int zfr_face_correction_rk(const zfr_ele_desc_t *d, double *divF, const zfr_rk_args_t *rk, void *stream) {
  if (int rc = need_geometry(d)) return rc;
  if (int rc = need_arg(divF, "divF")) return rc;
  if (!d->viscous) return zfr::fail("zfr_face_correction_rk: descriptor is inviscid");
  return rk_fused(zk::Op::FaceCorr, d, nullptr, divF, rk, (hipStream_t)stream);
}

int zfr_divergence_rk(const zfr_ele_desc_t *d, const double *U, double *divF, const zfr_rk_args_t *rk,
                      void *stream) {
  if (int rc = need_geometry(d)) return rc;
  if (int rc = need_arg(U, "U")) return rc;
  if (int rc = need_arg(divF, "divF")) return rc;
  return rk_fused(zk::Op::Div, d, U, divF, rk, (hipStream_t)stream);
}

int zfr_rk_stage_source(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF,
                        const double *source, double alpha, double dt, void *stream) {
  if (int rc = zk::refuse_while_recording("zfr_rk_stage")) return rc;
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
  if (int rc = zk::refuse_while_recording("zfr_rk_final")) return rc;
  if (nStages < 1 || nStages > zk::kMaxRkStages) return zfr::fail("zfr_rk_final: nStages must be 1..4");
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
  if (int rc = zk::refuse_while_recording("zfr_pmg_transfer")) return rc;
  if (nDims != 2 && nDims != 3) return zfr::fail("zfr_pmg_transfer: nDims must be 2 or 3");
  if (n_in < 1 || n_in > 10 || n_out < 1 || n_out > 10)
    return zfr::fail("zfr_pmg_transfer: points per direction must be in 1..10 (orders 0..9)");
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
  if (int rc = zk::refuse_while_recording("zfr_axpby")) return rc;
  if (n <= 0) return 0;
  if (!x || !y || !z) return zfr::fail("zfr_axpby: missing buffer");
  hipLaunchKernelGGL(zk::k_axpby, dim3(zk::grid_for(n)), dim3(zk::kMaxThreads), 0, (hipStream_t)stream, z, a, x, b,
                     y, n);
  return zfr::check_hip(hipGetLastError(), "zfr_axpby");
}

}  // extern "C"

namespace {
// Block partials of the residual norm: at most kNormMaxBlocks workgroups, V <= kNormMaxVars.
constexpr int kNormMaxBlocks = 1024, kNormMaxVars = 8;

int norm_checks(const zfr_ele_desc_t *d, const double *divF, int res_type, const double *res_out, const char *who) {
  if (!d) return zfr::fail(std::string(who) + ": NULL descriptor");
  if (res_type < 0 || res_type > 2) return zfr::fail(std::string(who) + ": res_type must be 0, 1 or 2");
  if (d->nDims != 2 && d->nDims != 3) return zfr::fail(std::string(who) + ": nDims must be 2 or 3");
  if (d->nVars != d->nDims + 2) return zfr::fail(std::string(who) + ": nVars must be nDims + 2");
  if (d->nEles < 0 || d->nSpts < 1) return zfr::fail(std::string(who) + ": bad sizes");
  if (!divF || !res_out || !d->metric) return zfr::fail(std::string(who) + ": missing buffer (divF, res_out or metric)");
  return 0;
}

// Partials into part[nblk][V] (every entry written before it is read), then the final
// reduction into res_out[V].
int norm_launch(const zfr_ele_desc_t *d, const double *divF, int res_type, double *part, double *res_out,
                hipStream_t st) {
  if (int rc = zk::refuse_while_recording("zfr_residual_norm")) return rc;
  const long npts = (long)d->nEles * d->nSpts;
  unsigned nblk = zk::grid_for(npts);
  if (nblk > (unsigned)kNormMaxBlocks) nblk = kNormMaxBlocks;
  if (d->nDims == 2)
    hipLaunchKernelGGL(zk::k_norm_partial<2>, dim3(nblk), dim3(zk::kMaxThreads), 0, st, divF, d->metric,
                       (long)d->nEles, d->nSpts, res_type, part);
  else
    hipLaunchKernelGGL(zk::k_norm_partial<3>, dim3(nblk), dim3(zk::kMaxThreads), 0, st, divF, d->metric,
                       (long)d->nEles, d->nSpts, res_type, part);
  hipLaunchKernelGGL(zk::k_norm_final, dim3(1), dim3(64), 0, st, part, (int)nblk, d->nVars, res_type, res_out);
  return zfr::check_hip(hipGetLastError(), "zfr_residual_norm");
}
}  // namespace

extern "C" {

int zfr_residual_norm(const zfr_ele_desc_t *d, const double *divF, int res_type, double *work, double *res_out,
                      void *stream) {
  if (int rc = norm_checks(d, divF, res_type, res_out, "zfr_residual_norm")) return rc;
  if (!work) return zfr::fail("zfr_residual_norm: missing buffer (work)");
  return norm_launch(d, divF, res_type, work, res_out, (hipStream_t)stream);
}

}  // extern "C"

namespace {
// The block partials of zfr_residual_norm_static: a fixed array in every device's global
// memory (64 KiB, held until the process ends), so that call allocates nothing.  Its
// host-side name is not a device address: host code uses only the address that
// hipGetSymbolAddress returns for the current device.
constexpr long kScratchDoubles = (long)kNormMaxBlocks * kNormMaxVars;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int scratch_for_current_device(double **out) {
  int dev = -1;
  if (int rc = zfr::check_hip(hipGetDevice(&dev), "zfr_residual_norm_static: hipGetDevice")) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return zfr::fail("zfr_residual_norm_static: device index " + std::to_string(dev) + " beyond " +
                     std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (int rc = zfr::check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                                "zfr_residual_norm_static: hipGetSymbolAddress(scratch)"))
      return rc;
    if (!p) return zfr::fail("zfr_residual_norm_static: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" {

int zfr_residual_norm_static(const zfr_ele_desc_t *d, const double *divF, int res_type, double *res_out,
                             void *stream) {
  if (int rc = norm_checks(d, divF, res_type, res_out, "zfr_residual_norm_static")) return rc;
  if (d->nVars > kNormMaxVars)
    return zfr::fail("zfr_residual_norm_static: nVars " + std::to_string(d->nVars) + " exceeds the scratch limit " +
                     std::to_string(kNormMaxVars));
  const long need = (long)std::min<long>(zk::grid_for((long)d->nEles * d->nSpts), kNormMaxBlocks) * d->nVars;
  if (need > kScratchDoubles)
    return zfr::fail("zfr_residual_norm_static: " + std::to_string(need) + " partials exceed the scratch array of " +
                     std::to_string(kScratchDoubles) + " doubles");
  double *part = nullptr;
  if (int rc = scratch_for_current_device(&part)) return rc;
  return norm_launch(d, divF, res_type, part, res_out, (hipStream_t)stream);
}

}  // extern "C"

extern "C" {

int zfr_residual_vjp(const zfr_ele_desc_t *d, const double *U, const double *w, double *v, double *Wtr,
                     double *Xtr, void *stream) {
  if (int rc = need_geometry(d)) return rc;
  if (!d->Utr || !Wtr || !Xtr || !w || !v) return zfr::fail("zfr_residual_vjp: missing buffer");
  zk::AdjPtrs j;
  j.w = w;
  j.Wtr = Wtr;
  j.Xtr = Xtr;
  const hipStream_t st = (hipStream_t)stream;
  if (d->viscous && d->ldg_b != 0.5 && d->ldg_b != -0.5) {
    // general LDG beta: forward traces and gradient traces, then the two-sided reverse
    if (!d->Gtr) return zfr::fail("zfr_residual_vjp: general LDG beta needs the gradient trace buffer Gtr");
    int rc = zk::dispatch(zk::Op::Extrap, d, U, nullptr, st);
    if (!rc) rc = zk::dispatch(zk::Op::Grad, d, U, nullptr, st);
    if (!rc) rc = zk::dispatch(zk::Op::AdjPrep, d, U, nullptr, st, j);
    if (!rc) rc = zk::dispatch(zk::Op::AdjFace, d, U, nullptr, st, j);
    if (!rc) rc = zk::dispatch(zk::Op::AdjMainGen, d, U, v, st, j);
    if (!rc) rc = zk::dispatch(zk::Op::AdjGather, d, U, v, st, j);
    return rc;
  }
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
    if ((D != 2 && D != 3) || N < 1 || N > (D == 2 ? 10 : 9))
      return zfr::fail("zfr_multi_create: supported shapes are quads of order 0..9 and hexes of order 0..8");
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
      // one launch per phase for every order and element type (N = 0, D = 0: mixed launch;
      // narrowed to one nDims below when the group holds only one)
      // (orders above the mixed launches' range: one launch per shape)
      const bool mixed = N <= zk::kMixedMaxN;
      phase[1][{Op::OneSided, mixed ? 0 : D, mixed ? 0 : N, 1, bm, fc}].push_back(i);  // extrapolates the partner traces itself
      phase[2][{Op::FaceCorr, mixed ? 0 : D, mixed ? 0 : N, 1, bm, 0}].push_back(i);
    } else {
      if (visc && !d->Gtr) return zfr::fail("zfr_multi_create: the general LDG path needs Gtr");
      if (visc) phase[0][{Op::Extrap, D, N, 0, 0, 0}].push_back(i);  // Euler: extrapolated in the divergence
      if (visc) phase[1][{Op::Grad, D, N, 1, 0, 0}].push_back(i);
      const bool own = visc || N > zk::kMixedMaxN;
      phase[2][{Op::Div, own ? D : 0, own ? N : 0, visc, bm, fc}].push_back(i);
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
      if (g.D == 0) {  // a mixed group: its own nDims' kernel when every case has one nDims
        const int d0 = descs[ids[0]].nDims;
        bool one = true;
        for (int i : ids) one = one && descs[i].nDims == d0;
        if (one) g.D = d0;
      }
      // one tile per order, chosen for that order's total element count in the group, as
      // for one solver but never shrunk to one wave (a mixed launch's workgroups have
      // kMaxThreads threads); the launch takes the largest LDS of its orders
      std::map<int, int> eb_of;  // tile per shape, keyed nDims * 16 + N
      g.lds = 0;
      g.EB = 0;
      for (int i : ids) {
        const int N = descs[i].order + 1, cD = descs[i].nDims, shape = cD * 16 + N;
        if (eb_of.count(shape)) continue;
        const size_t per_ele = (cD == 2) ? zk::lds_per_element_n<2>(N, k.op, k.visc != 0)
                                         : zk::lds_per_element_n<3>(N, k.op, k.visc != 0);
        const int nS = (cD == 2) ? N * N : N * N * N;
        const int nops = 2 * N + N * N + 2 * N, nops_pad = (nops + 7) & ~7;
        zfr_ele_desc_t agg = descs[i];
        agg.nEles = 0;
        for (int i2 : ids)
          if (descs[i2].order + 1 == N && descs[i2].nDims == cD) agg.nEles += descs[i2].nEles;
        agg.eles_per_block = 0;
        const int defEB = (zk::kMaxThreads / nS) > 0 ? zk::kMaxThreads / nS : 1;
        const int eb = zk::eles_per_block(&agg, defEB, nops_pad, per_ele, false);
        eb_of[shape] = eb;
        const size_t lds = sizeof(double) * (nops_pad + eb * per_ele);
        if (lds > zk::max_lds_n(cD, N)) {
          delete m;
          return zfr::fail("zfr_multi_create: one element needs more LDS than a workgroup may use");
        }
        g.lds = std::max(g.lds, lds);
        g.EB = std::max(g.EB, eb);
      }
      std::vector<zk::MultiCase> hc(ids.size());
      std::vector<long> hs(ids.size());
      long blocks = 0;
      for (size_t q = 0; q < ids.size(); q++) {
        const int i = ids[q];
        const int N = descs[i].order + 1, eb = eb_of[descs[i].nDims * 16 + N];
        hc[q].a = zk::make_args(&descs[i], eb);
        hc[q].a.trace_sel = (k.op == Op::ExtrapPartner) ? 1 : 0;
        hc[q].U = U[i];
        hc[q].out = (k.op == Op::Div || k.op == Op::OneSided || k.op == Op::FaceCorr) ? divF[i] : nullptr;
        hc[q].N = N;
        hc[q].D = descs[i].nDims;
        hs[q] = blocks;
        blocks += (descs[i].nEles + eb - 1) / eb;
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
  if (int rc = zk::refuse_while_recording("zfr_multi_residual")) return rc;
  if (!plan) return zfr::fail("zfr_multi_residual: null plan");
  const hipStream_t st = (hipStream_t)stream;
  for (const zk::MultiGroup &g : plan->groups) {
    if (g.nblk == 0) continue;
    if (g.D == 0) zk::launch_mixed<0>(g, st);  // quads and hexes in one launch
    else if (g.D == 2) zk::dispatch_multi_n<2>(g, st);
    else zk::dispatch_multi_n<3>(g, st);
    const int rc = zfr::check_hip(hipGetLastError(), "zfr_multi_residual");
    if (rc) return rc;
  }
  return 0;
}

int zfr_multi_launches(const zfr_multi_t *plan) { return plan ? (int)plan->groups.size() : -1; }

void zfr_multi_destroy(zfr_multi_t *plan) { delete plan; }

int zfr_stream_copy(double *dst, const double *src, long n, void *stream) {
  if (int rc = zk::refuse_while_recording("zfr_stream_copy")) return rc;
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

// ---------------------------------------------------------------- persistent plans
struct zfr_persist_s {
  int D = 0, nph = 0;
  unsigned grid = 0, max_grid = 0;
  size_t lds = 0;
  zk::PersistPhase *phases = nullptr;  // device [nph]
  ~zfr_persist_s() {
    if (phases) (void)hipFree(phases);
  }
};

extern "C" {

int zfr_persist_record_begin(void) {
  if (zk::g_persist) return zfr::fail("zfr_persist_record_begin: this thread is already recording");
  zk::g_persist = new zk::PersistRecorder();
  return 0;
}

int zfr_persist_record_end(zfr_persist_t **plan) {
  std::unique_ptr<zk::PersistRecorder> rec(zk::g_persist);
  zk::g_persist = nullptr;
  if (!rec) return zfr::fail("zfr_persist_record_end: this thread is not recording");
  if (!plan) return 0;  // the recording is discarded
  *plan = nullptr;
  if (rec->phases.empty()) return zfr::fail("zfr_persist_record_end: no launch was recorded");
  int dev = 0, cus = 0, coop = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
  if (e == hipSuccess)
    e = (rec->D == 2) ? zk::persist_blocks_per_cu<2>(rec->lds, &per_cu) : zk::persist_blocks_per_cu<3>(rec->lds, &per_cu);
  if (e != hipSuccess) return zfr::check_hip(e, "zfr_persist_record_end");
  if (!coop) return zfr::fail("zfr_persist_record_end: the device has no cooperative launches");
  if (per_cu < 1 || cus < 1) return zfr::fail("zfr_persist_record_end: the persistent kernel does not fit a CU");
  // at most every workgroup that can be resident (the cooperative launch refuses more) and
  // the largest phase's element blocks; by default one per CU: each grid-wide barrier costs
  // about 0.1 us per workgroup here (10 Euler residuals of 128^2 quads at p = 4: 1.03 ms
  // with 768 workgroups, 0.78 with 512, 0.66 with 256, 0.86 with 128)
  long most = 0;
  for (const auto &ph : rec->phases) most = std::max(most, ph.nblk);
  const long max_grid = std::max(1L, std::min((long)per_cu * cus, most));
  auto m = std::make_unique<zfr_persist_s>();
  m->D = rec->D;
  m->nph = (int)rec->phases.size();
  m->max_grid = (unsigned)max_grid;
  m->grid = (unsigned)std::min(max_grid, (long)cus);
  m->lds = rec->lds;
  e = hipMalloc(&m->phases, rec->phases.size() * sizeof(zk::PersistPhase));
  if (e == hipSuccess)
    e = hipMemcpy(m->phases, rec->phases.data(), rec->phases.size() * sizeof(zk::PersistPhase),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) return zfr::check_hip(e, "zfr_persist_record_end");
  *plan = m.release();
  return 0;
}

int zfr_persist_launch(const zfr_persist_t *plan, void *stream) {
  if (!plan) return zfr::fail("zfr_persist_launch: null plan");
  if (int rc = zk::refuse_while_recording("zfr_persist_launch")) return rc;
  const hipError_t e = (plan->D == 2)
                           ? zk::launch_persist<2>(plan->phases, plan->nph, plan->grid, plan->lds, (hipStream_t)stream)
                           : zk::launch_persist<3>(plan->phases, plan->nph, plan->grid, plan->lds, (hipStream_t)stream);
  return zfr::check_hip(e, "zfr_persist_launch");
}

int zfr_persist_info(const zfr_persist_t *plan, int *phases, int *workgroups) {
  if (!plan) return zfr::fail("zfr_persist_info: null plan");
  if (phases) *phases = plan->nph;
  if (workgroups) *workgroups = (int)plan->grid;
  return 0;
}

int zfr_persist_set_workgroups(zfr_persist_t *plan, int n) {
  if (!plan) return zfr::fail("zfr_persist_set_workgroups: null plan");
  if (n < 1 || (unsigned)n > plan->max_grid)
    return zfr::fail("zfr_persist_set_workgroups: n must be in 1.." + std::to_string(plan->max_grid) +
                     " (the workgroups that are resident at once)");
  plan->grid = (unsigned)n;
  return 0;
}

void zfr_persist_destroy(zfr_persist_t *plan) { delete plan; }

}  // extern "C"
