// kernels_mixed.hip -- multi-case launches over cases of different orders, and of different
// element types (quads and hexes): one launch per (kernel, flux branch) for a whole
// multi-case plan.  Compiled with -DZK_D=2 or 3 (cases of that nDims, any order) and with
// -DZK_D=0 (quads and hexes together).  Workgroup b finds its case by a binary search over
// the cases' first workgroups, then switches on that case's shape (uniform across the
// workgroup) into the single-shape block function, with the case's own tile; threads beyond
// its EB * nS idle.  The launch takes the largest LDS any of its cases needs; the register
// budget is the largest of the shapes compiled in, which is why single-nDims groups keep
// their own kernel.
#include <hip/hip_cooperative_groups.h>

#include "kernels.hpp"

#if !defined(ZK_D)
#error "kernels_mixed.hip is compiled with -DZK_D=<nDims, or 0 for quads and hexes together>"
#endif

namespace zk {

#define ZK_SWITCH_N_OF(DD, CALL)                         \
  switch (m.N) {                                         \
    case 1: { constexpr int N = 1, D = DD; CALL; } break; \
    case 2: { constexpr int N = 2, D = DD; CALL; } break; \
    case 3: { constexpr int N = 3, D = DD; CALL; } break; \
    case 4: { constexpr int N = 4, D = DD; CALL; } break; \
    case 5: { constexpr int N = 5, D = DD; CALL; } break; \
    case 6: { constexpr int N = 6, D = DD; CALL; } break; \
  }
// DIMS 2 or 3: that nDims; 0: the case's own nDims (quads and hexes in one launch)
#define ZK_SWITCH_N(CALL)                                \
  if constexpr (DIMS == 2) { ZK_SWITCH_N_OF(2, CALL) }   \
  else if constexpr (DIMS == 3) { ZK_SWITCH_N_OF(3, CALL) } \
  else if (m.D == 2) { ZK_SWITCH_N_OF(2, CALL) }         \
  else { ZK_SWITCH_N_OF(3, CALL) }

template <int DIMS, bool VISC, int BM, int FC>
__global__ __launch_bounds__(kMaxThreads) void k_divergence_mixed(const MultiCase *__restrict__ cs,
                                                                  const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  const MultiCase &m = cs[c];
  const long b = (long)blockIdx.x - starts[c];
  ZK_SWITCH_N((divergence_block<D, N, VISC, BM, FC, !VISC>(m.U, m.out, m.a, b)))
}

template <int DIMS, int BM, int FC>
__global__ __launch_bounds__(kMaxThreads) void k_visc_onesided_mixed(const MultiCase *__restrict__ cs,
                                                                     const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  const MultiCase &m = cs[c];
  const long b = (long)blockIdx.x - starts[c];
  ZK_SWITCH_N((visc_onesided_block<D, N, BM, FC, true>(m.U, m.out, m.a, b)))
}

template <int DIMS, int BM>
__global__ __launch_bounds__(kMaxThreads) void k_face_correction_mixed(const MultiCase *__restrict__ cs,
                                                                       const long *__restrict__ starts, int nc) {
  const int c = find_case(starts, nc, blockIdx.x);
  const MultiCase &m = cs[c];
  const long b = (long)blockIdx.x - starts[c];
  ZK_SWITCH_N((face_correction_block<D, N, BM>(m.out, m.a, b)))
}
#if ZK_D != 0
// A persistent plan (zfr_persist_*): every phase's element blocks walked by the resident
// workgroups, one grid-wide barrier between phases (cooperative launch: all workgroups are
// resident, so every barrier is reached by all of them).  Each block runs the Euler
// divergence of the single launch's workgroup b with the phase's own tile; threads
// beyond its EB * nS idle.
// Each variant is a function of its own (not inlined): the registers of one variant are
// allocated on their own, so the kernel needs the largest variant's, not those of all of
// them inlined into one body (256 VGPRs and 76 AGPRs for the quad kernel).
template <int D, int N, int FC, bool SELF>
__device__ __attribute__((noinline)) void persist_div(const PersistPhase *m, long bid) {
  divergence_block<D, N, false, 0, FC, SELF>(m->U, m->out, m->a, bid);
}

template <int D, int N>
__device__ __forceinline__ void persist_block(const PersistPhase &m, long bid) {
  if (m.fc) {
    if (m.self) persist_div<D, N, 1, true>(&m, bid);
    else persist_div<D, N, 1, false>(&m, bid);
  } else {
    if (m.self) persist_div<D, N, 0, true>(&m, bid);
    else persist_div<D, N, 0, false>(&m, bid);
  }
}

template <int DIMS>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(3)))
void k_persist(const PersistPhase *__restrict__ ph, int nph) {
  for (int i = 0; i < nph; i++) {
    const PersistPhase &m = ph[i];
    for (long b = blockIdx.x; b < m.nblk; b += gridDim.x) {
      const long bid = xcd_block(b, m.nblk);
      ZK_SWITCH_N((persist_block<D, N>(m, bid)))
      __syncthreads();  // the block's LDS is reused by the next one
    }
    if (i + 1 < nph) cooperative_groups::this_grid().sync();
  }
}
#endif
#undef ZK_SWITCH_N
#undef ZK_SWITCH_N_OF

template <int D>
void launch_mixed(const MultiGroup &g, hipStream_t st) {
  const dim3 grid(g.nblk), block(kMaxThreads);
  const MultiCase *c = g.cases;
  const long *s0 = g.starts;
  switch (g.op) {
    case Op::Div:
#define ZK_DIV(V_, B_, F_) hipLaunchKernelGGL((k_divergence_mixed<D, V_, B_, F_>), grid, block, g.lds, st, c, s0, g.nc)
      if (!g.visc) { if (g.fc) ZK_DIV(false, 0, 1); else ZK_DIV(false, 0, 0); }
#undef ZK_DIV
      break;
    case Op::OneSided:
#define ZK_OSM(B_, F_) hipLaunchKernelGGL((k_visc_onesided_mixed<D, B_, F_>), grid, block, g.lds, st, c, s0, g.nc)
      if (g.bm == 1) { if (g.fc) ZK_OSM(1, 1); else ZK_OSM(1, 0); }
      else { if (g.fc) ZK_OSM(2, 1); else ZK_OSM(2, 0); }
#undef ZK_OSM
      break;
    case Op::FaceCorr:
      if (g.bm == 1) hipLaunchKernelGGL((k_face_correction_mixed<D, 1>), grid, block, g.lds, st, c, s0, g.nc);
      else hipLaunchKernelGGL((k_face_correction_mixed<D, 2>), grid, block, g.lds, st, c, s0, g.nc);
      break;
    default:
      break;
  }
}

template void launch_mixed<ZK_D>(const MultiGroup &, hipStream_t);

#if ZK_D != 0
template <int D>
hipError_t launch_persist(const PersistPhase *phases, int nph, unsigned grid, size_t lds, hipStream_t st) {
  void *args[] = {(void *)&phases, (void *)&nph};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_persist<D>), dim3(grid), dim3(kMaxThreads),
                                    args, (unsigned)lds, st);
}

template <int D>
hipError_t persist_blocks_per_cu(size_t lds, int *n) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_persist<D>, kMaxThreads, lds);
}

template hipError_t launch_persist<ZK_D>(const PersistPhase *, int, unsigned, size_t, hipStream_t);
template hipError_t persist_blocks_per_cu<ZK_D>(size_t, int *);
#endif

}  // namespace zk
