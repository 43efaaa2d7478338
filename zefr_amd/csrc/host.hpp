// host.hpp -- host-side classes of the MI355X residual path.
//
// The class names and public methods follow the reference (FRSolver, Elements, Faces,
// PMGrid; include/solver.hpp:34-233, include/elements.hpp:237-308,
// include/faces.hpp:147-167, include/multigrid.hpp:17-44) so callers are unchanged.
// Their internals are a fresh design: the host builds connectivity, operators and
// metrics once, lays them out element-major for the device, and every hot-path call is
// a handful of HIP kernel launches through the kernel C ABI (zfr_* in zefr_amd.h).
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <functional>
#include <cstddef>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "zefr_amd.h"

namespace zfr {

enum EleType { LINE_T = 0, QUAD_T = 2, HEX_T = 4 };

// ---------------------------------------------------------------- input (input.cpp)
void read_input_file(const std::string &file, zefr_input_t &in);
void initialize_inputs(zefr_input_t &in);
void apply_nondim(zefr_input_t &in);
void finish_freestream(zefr_input_t &in);
void set_stages(zefr_input_t &in);

// ---------------------------------------------------------------- basis (basis.cpp)
std::vector<double> gauss_legendre_points(unsigned n);
std::vector<double> gauss_legendre_weights(unsigned n);
double lagrange(const std::vector<double> &grid, unsigned mode, double x);
double lagrange_d1(const std::vector<double> &grid, unsigned mode, double x);

// Reference element of a tensor-product quad/hex of order p: point sets in the
// reference's orderings (src/quads.cpp:45-190, src/hexas.cpp:51-292), the 1D DFR tables
// the kernels contract with, and linear shape functions.
// Highest supported order per nDims: quads 9, hexes 8 (the residual VJP of hexes to 7).
inline int max_order(int nDims) { return nDims == 2 ? 9 : 8; }

struct RefElement {
  int nDims = 0, order = 0, N = 0, nSpts = 0, nFpts = 0, nFptsPerFace = 0, nFaces = 0, nNodes = 0;
  std::vector<double> loc1d, locDFR, wts1d;
  std::vector<double> loc_spts, loc_fpts, tnorm, weights_spts;  // [n][nDims]
  std::vector<int> idx_spts, idx_fpts;                          // [n][nDims]
  std::vector<double> ops1d;  // lE[2][N], Dss[N][N], Dsf[N][2]

  void build(int nDims, int order);
  void shape(const double *loc, double *val, double *dval) const;  // linear, gmsh node order
  std::vector<double> dense_operator(const std::string &name) const;
};

// ---------------------------------------------------------------- mesh (mesh.cpp)
using FaceKey = std::array<unsigned, 4>;  // sorted node ids, UINT_MAX padded

struct Mesh {
  int nDims = 0, etype = 0, nNodes = 0, nEles = 0, nNodesPerEle = 0;
  std::vector<double> coord;              // [nNodes][nDims]
  std::vector<int> ele2nodes;             // [nEles][nNodesPerEle]
  std::vector<FaceKey> per_faces;         // periodic boundary faces (sorted node ids)
  std::vector<int> per_pair;              // index into per_faces of each face's partner
};

Mesh read_gmsh(const std::string &file, const zefr_input_t &in);
void couple_periodic(Mesh &m);

struct Connectivity {
  int nFptsPerFace = 0, nGfpts = 0, nGfpts_int = 0, nFaces = 0;
  std::vector<int> fpt2gfpt;         // [nFpts][nEles], after orientation
  std::vector<signed char> slot;     // [nFpts][nEles]
  std::vector<int> gf_ele, gf_fpt;   // [2][nGfpts]: element / element fpt of each side
};

// fpt_coord(e, f, x) writes the physical coordinates of flux point f of element e.
Connectivity build_connectivity(const Mesh &m, const RefElement &re,
                                const std::function<void(int, int, double *)> &fpt_coord);

// ---------------------------------------------------------------- device buffers
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf();
  void alloc(size_t nbytes);
  void release();
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// ---------------------------------------------------------------- solver classes
class Faces;

// One element-type block (include/elements.hpp:237-308).  Geometry is computed per
// element straight into the element-major device layout; the reference-layout arrays
// (coord_spts, inv_jaco_spts, ...) are produced only on request by export_array().
class Elements {
 public:
  int etype = 0, nDims = 0, order = 0, nEles = 0, nSpts = 0, nFpts = 0, nVars = 0, nNodes = 0;
  RefElement re;
  std::vector<double> shp_s, dshp_s, shp_f, dshp_f;  // shape values [n][nNodes], derivatives [n][nDims][nNodes]
  std::vector<double> X;                             // element nodes [E][nNodes][nDims]
  std::vector<double> metric;                        // [E][D*D+1][nS]: adjugate(d1,d2), detJ
  // device data (element-major)
  DevBuf U_spts_d, U_ini_d, divF_d, dU_spts_d, metric_d, fgeo_d, nbr_d, ops_d, Utr_d, Gtr_d, Ftr_d;
  DevBuf Utr2_d;  // second trace buffer: fused RK stages on the divergence path alternate with Utr_d
  std::vector<double> U_host, dU_host;  // host mirrors for get_u_spts / get_du_spts
  int tile = 0;                         // elements per workgroup (0 = kernel default)

  void setup_geometry(const Mesh &m);
  void point_coord(int e, bool fpt, int p, double *x) const;
  void jacobian(int e, bool fpt, int p, double J[3][3]) const;  // J[r][d] = dx_d / dxi_r
  // adjugate A (A[a][b] in the reference's inv_jaco(a, pt, b) order) and determinant
  static double adjugate(int D, const double J[3][3], double A[3][3]);
  std::vector<double> initial_state(const zefr_input_t &in) const;  // [E][V][nS]
  std::vector<double> export_array(const std::string &name) const;  // reference layouts
  // phase: 0 trace, 1 gradient / one-sided flux, 2 divergence / face correction (the
  // autotuner's per-launch tiles); -1 = the solver-wide tile
  zfr_ele_desc_t desc(const zefr_input_t &in, int phase = -1) const;
  int tile_phase[3] = {0, 0, 0};  // per-launch tiles picked by FRSolver::autotune (0 = tile)
  int tile_for(int phase) const { return (phase >= 0 && phase < 3 && tile_phase[phase] > 0) ? tile_phase[phase] : tile; }
  // the four tile settings a captured launch sequence bakes in (compared element-wise)
  std::array<int, 4> tile_key() const { return {tile, tile_phase[0], tile_phase[1], tile_phase[2]}; }
  size_t state_size() const { return (size_t)nEles * nVars * nSpts; }
};

// Interface data (include/faces.hpp:147-167): per element flux point, the unit normal of
// its global flux point (slot-0 side) and this side's signed area, plus the partner map
// that replaces the reference's per-gfpt pointer views.
class Faces {
 public:
  int nGfpts = 0;
  std::vector<double> fgeo;   // [E][D+1][nF]
  std::vector<unsigned> nbr;  // [E][nF]: partner ele << (B + 1) | own slot << B | partner fpt (B: nbr_fpt_bits)
  void setup(const Elements &e, const Connectivity &c);
  // reference layouts: norm [nDims][nGfpts], dA [2][nGfpts]
  std::vector<double> export_array(const std::string &name, const Elements &e, const Connectivity &c) const;
};

class FRSolver {
 public:
  explicit FRSolver(zefr_input_t *input, int order = -1);
  ~FRSolver();
  void setup_host();
  void setup_device();
  void setup() { setup_host(); setup_device(); }

  // src/solver.cpp:1494-1840
  void compute_residual(unsigned stage, int color = -1);
  void compute_residual_start(unsigned stage, int color = -1);
  void compute_residual_mid(unsigned stage, int color = -1);
  void compute_residual_finish(unsigned stage, int color = -1);
  // src/solver.cpp:2599-3097, 3458-3736
  void update();
  void step_RK();
  void step_RK_stage(int stage);
  void step_RK_stage_start(int stage);
  void step_RK_stage_mid(int stage);
  void step_RK_stage_finish(int stage);
  // src/solver.cpp:6990-7137 (no history file; res[nVars])
  void report_residuals(double *res);
  void compute_gradient_export();  // fills dU_spts_d with the physical gradient of U
  void synchronize();
  int autotune(int reps);
  bool onesided() const;  // viscous with LDG beta = +-1/2: one-sided gradient path
  void time_kernels(int reps, double *ms_out);
  // The residual of an arbitrary device state U [E][V][nS] into divF (same layout) on the
  // caller's stream st (nullptr: the null stream), with the kernels of compute_residual;
  // and its vector-Jacobian product v = (d divF / dU)^T w.  Both use the solver's trace
  // scratch, so they must not overlap other work of this solver.
  void residual_at(const double *U, double *divF, hipStream_t st);
  void residual_vjp(const double *U, const double *w, double *v, hipStream_t st);

  zefr_input_t *input;
  int order;
  Mesh mesh;
  Connectivity conn;
  std::shared_ptr<Elements> eles;
  std::shared_ptr<Faces> faces;
  hipStream_t stream = nullptr;
  double flow_time = 0.0, prev_time = 0.0, res_max = 0.0;
  int current_iter = 0;
  bool device_ready = false;
  DevBuf norm_work_d, norm_out_d;
  DevBuf Wtr_d, Xtr_d;  // reverse-mode scratch [E][V][nF] each, allocated on first use
  // Time-step execution: the RK update fused into each residual's last kernel, and the
  // whole step replayed from a captured HIP graph (recaptured when dt or the tile change).
  bool fuse_rk = true;
  bool use_graph = false;
  void set_step_mode(bool fuse, bool graph);
  // FRSolver::update(sourceBT) (src/solver.cpp:2599-2646): one step with a p-multigrid
  // source term [E][V][nS] added to every stage residual (nullptr: plain update)
  void update(const double *source);
  // The same, also storing the starting state in save and, after the step, writing
  // corr = U - corr_base (each nullptr for none).  With fuse_rk and several stages the
  // copy and the difference are written by the step's own kernels; otherwise by a copy
  // and an axpby around the step.
  void update(const double *source, double *save, double *corr, const double *corr_base);
  // p-multigrid's prolongation of the correction written by a step (prolong_err,
  // src/multigrid.cpp:394-433): out += alpha * (A [n][order+1] per direction) corr
  struct Prolong {
    const double *A = nullptr;
    int n = 0;
    double alpha = 0.0;
    double *out = nullptr;
  };
  // With fuse_rk: the step's final combination also prolongs the correction (corr and
  // corr_base needed); the caller prolongs it otherwise.  form_source (fuse_rk, several
  // stages): source holds the restricted residual, and stage 0 first turns it into the
  // source term (PMGrid::compute_source_term: source -= the stage-0 residual, which is
  // the residual at the state the step starts from) before it uses it, so the source
  // term needs no residual launch of its own.
  void update(const double *source, double *save, double *corr, const double *corr_base, const Prolong &pro,
              bool form_source = false);
  // With fuse_rk: the stage-0 residual (plus source when not null) and, in the same last
  // kernel, U and that residual restricted with R [n_out][order+1] per direction into Uc
  // and Rc (restrict_pmg, src/multigrid.cpp:335-392).
  void residual_restrict(const double *source, const double *R, int n_out, double *Uc, double *Rc);
  // The stage-0 residual into divF(0) with a p-multigrid epilogue: mode 3 subtracts it
  // from source_out (source_out -= divF), mode 4 adds the source term to it
  // (divF += source).  With fuse_rk the last residual kernel does it; otherwise an axpby.
  void residual_source(int mode, double *source_out, const double *source);
  bool own_stream = true;  // false for p-multigrid levels, which run on the fine stream
  // The trace buffer the residual phases read, when not Utr_d (fused RK steps on the
  // divergence path alternate two buffers; see step_RK_eager)
  double *trace_in = nullptr;
  // The state the residual phases read, when not U_spts_d (stages 1.. of a fused step,
  // whose stage states live in U_ini_d; see step_RK_eager)
  const double *state_in = nullptr;
  // PyFR-format solution files (src/solver.cpp:4275-4837): write returns the file name
  std::string write_solution_pyfr(const std::string &prefix);
  void restart_pyfr(const std::string &restart_case, unsigned restart_iter);
  void restart_solution();  // src/solver.cpp:212-222, restart_type 2 only

 private:
  void rk_stage_update(int stage);
  void rk_final_update();
  void finish_with_rk(int stage, double *trace_out, bool split);
  const double *cur_state() const;
  void step_RK_eager();
  void drop_graph();
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double graph_dt = 0.0;
  std::array<int, 4> graph_tile = {-1, -1, -1, -1};
  bool graph_fused = false;
  const double *graph_source = nullptr;
  std::array<const void *, 6> graph_extra = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double graph_pro_alpha = 0.0;
  int graph_pro_n = 0;
  const double *rk_source = nullptr;  // source term of the step in progress
  double *rk_save = nullptr;          // store the step's starting state here
  double *rk_corr = nullptr;          // after the step: rk_corr = U - rk_corr_base
  const double *rk_corr_base = nullptr;
  Prolong rk_prolong;                  // with rk_corr: prolong the correction (fused steps)
  double *rk_form_source = nullptr;   // stage 0 forms the source term here (update's form_source)
};

// p-multigrid (include/multigrid.hpp:17-44, src/multigrid.cpp:18-518) on quads/hexes:
// coarse FRSolvers at the orders of mg_levels on the fine solver's mesh, all on the fine
// solver's stream, with explicit RK smoothing under a source term.  Level 0 is the
// solver passed in; grids[0] is left empty as in the reference.
class PMGrid {
 public:
  void setup(zefr_input_t *input, FRSolver &solver);
  void cycle(FRSolver &solver);
  void v_cycle(FRSolver &solver, int level);
  void restrict_pmg(FRSolver &grid_f, FRSolver &grid_c, double *res_out);
  void prolong_err(FRSolver &grid_c, const double *correction, FRSolver &grid_f);
  void prolong_U(FRSolver &grid_c, FRSolver &grid_f);
  void compute_source_term(FRSolver &grid, double *source);

  zefr_input_t *input = nullptr;
  int order = 0, nLevels = 0;
  std::vector<std::shared_ptr<FRSolver>> grids;
  // per level: source term, stored solution and correction [E][V][nS(level)] (level >= 1;
  // the fine level has no buffers: its correction is prolonged straight into U)
  std::deque<DevBuf> sourcesBT, solutionsBT, correctionsBT;
  // 1D transfer tables per level (row-major [n_out][n_in]): oppRes maps level n to n+1,
  // oppPro maps level n to n-1 (src/quads.cpp:363-465, src/hexas.cpp:586-704)
  std::vector<std::vector<double>> oppRes, oppPro;
  std::deque<DevBuf> oppRes_d, oppPro_d;
  // V-cycles replayed from one captured HIP graph (every level, transfer and update of
  // the cycle); recaptured when dt, rel_fac, a tile or the fusion mode change
  bool use_graph = false;
  // V-cycles replayed as one persistent launch (zfr_persist_*); recorded like the graph
  void set_persistent(bool on);
  bool persistent() const { return use_persistent; }
  ~PMGrid();

 private:
  bool use_persistent = false;
  zfr_persist_t *persist = nullptr;
  std::vector<double> persist_key;
  void check_persistent(FRSolver &solver);
  // run_v between begin() and end() (a stream capture or a plan recording: nothing runs),
  // abort() instead of end() if it throws; each level's iteration counter and time are
  // left as they were, and the updates of each level in one cycle are returned
  std::vector<int> record_cycle(FRSolver &solver, const std::function<void()> &begin,
                                const std::function<void()> &abort, const std::function<void()> &end);
  void advance_counters(FRSolver &solver, const std::vector<int> &iters);
  std::vector<int> persist_iters;
  FRSolver &level(FRSolver &fine, int n) { return n == 0 ? fine : *grids[n]; }
  void run_v(FRSolver &solver);
  void drop_graph();
  std::vector<double> graph_key(const FRSolver &solver) const;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<double> key;
  std::vector<int> iters_per_cycle;  // updates of each level in one cycle (counters on replay)
  void transfer(const FRSolver &from, const FRSolver &to, const DevBuf &A, double alpha, const double *in,
                double beta, double *out);
};

// 1D p-multigrid tables (src/quads.cpp:363-465 in one direction): the chained restriction
// from order p down to order r < p, and the chained prolongation from order p up to q > p,
// each the product of single-order Lagrange interpolations between Gauss-Legendre point sets.
std::vector<double> pmg_restriction_1d(int p, int r);
std::vector<double> pmg_prolongation_1d(int p, int q);

}  // namespace zfr
