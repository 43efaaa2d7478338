// multigrid.cpp -- p-multigrid (PMGrid) on quad / hex meshes.
//
// The cycle is the reference's (src/multigrid.cpp:18-518): coarse FRSolvers at the orders
// of mg_levels on the same mesh; restriction of the solution and of the stage-0 residual;
// a source term that makes each coarse problem reproduce the restricted fine residual at
// the restricted state; explicit RK smoothing under that source; and prolongation of each
// coarse level's change back to the next finer level.  Every level runs on the fine
// solver's stream, so the whole cycle is one ordered sequence of launches.  The operator
// products go through zfr_pmg_transfer (the 1D interpolation table applied per direction,
// where the reference multiplies by the dense tensor operator), the vector updates through
// zfr_axpby, the smoothing through the solvers' own (RK-fused) steps.
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"

namespace zfr {

namespace {

// Interpolation between Gauss-Legendre point sets: A[i][j] = l_j(x_i), with l_j the
// Lagrange polynomials on n_from points and x_i the n_to points.
std::vector<double> interp_1d(int n_from, int n_to) {
  const auto xf = gauss_legendre_points((unsigned)n_from), xt = gauss_legendre_points((unsigned)n_to);
  std::vector<double> A((size_t)n_to * n_from);
  for (int i = 0; i < n_to; i++)
    for (int j = 0; j < n_from; j++) A[(size_t)i * n_from + j] = lagrange(xf, (unsigned)j, xt[i]);
  return A;
}

// C[m][n] = A[m][k] B[k][n], row-major, summed in increasing k.
std::vector<double> matmul(const std::vector<double> &A, const std::vector<double> &B, int m, int k, int n) {
  std::vector<double> C((size_t)m * n, 0.0);
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) {
      double s = 0.0;
      for (int l = 0; l < k; l++) s += A[(size_t)i * k + l] * B[(size_t)l * n + j];
      C[(size_t)i * n + j] = s;
    }
  return C;
}

}  // namespace

// Restriction from order p to order r < p: the single-order steps (order P+1 -> P, the
// order-(P+1) interpolant evaluated at the order-P points) multiplied as
// ((I[r<-r+1] I[r+1<-r+2]) I[r+2<-r+3]) ... I[p-1<-p], the association of
// src/quads.cpp:416-457.  The composition is not the direct interpolation from p to r.
std::vector<double> pmg_restriction_1d(int p, int r) {
  if (r < 0 || r >= p || p > 9) raise("p-multigrid restriction needs 0 <= coarse order < fine order <= 9");
  std::vector<double> R = interp_1d(r + 2, r + 1);
  for (int P = r + 1; P < p; P++) R = matmul(R, interp_1d(P + 2, P + 1), r + 1, P + 1, P + 2);
  return R;  // [r+1][p+1]
}

// Prolongation from order p to order q > p: ((I[q<-q-1] I[q-1<-q-2]) ...) I[p+1<-p]
// (src/quads.cpp:373-414).  Interpolation reproduces polynomials, so this equals the direct
// interpolation from p to q up to rounding.
std::vector<double> pmg_prolongation_1d(int p, int q) {
  if (p < 0 || q <= p || q > 9) raise("p-multigrid prolongation needs 0 <= coarse order < fine order <= 9");
  std::vector<double> Pm = interp_1d(q, q + 1);
  for (int P = q - 1; P > p; P--) Pm = matmul(Pm, interp_1d(P, P + 1), q + 1, P + 1, P);
  return Pm;  // [q+1][p+1]
}

// PMGrid::setup (src/multigrid.cpp:18-103).
void PMGrid::setup(zefr_input_t *in, FRSolver &solver) {
  input = in;
  order = solver.order;
  nLevels = (int)in->nLevels;
  if (nLevels == 0) raise("Need to provide mg_levels and/or mg_steps to run multigrid!");
  if ((int)in->mg_levels[0] != order) raise("Invalid mg_levels provided. First level must equal order!");
  // the reference reads mg_levels[1] unconditionally and builds empty operators for a
  // non-decreasing list; both are rejected here
  if (nLevels < 2) raise("mg_levels needs at least two levels");
  for (int n = 1; n < nLevels; n++)
    if (in->mg_levels[n] >= in->mg_levels[n - 1]) raise("mg_levels must decrease strictly from the fine order");
  grids.clear();
  grids.emplace_back();  // level 0 is the solver itself
  oppRes.assign(nLevels, {});
  oppPro.assign(nLevels, {});
  sourcesBT.clear();
  solutionsBT.clear();
  correctionsBT.clear();
  oppRes_d.clear();
  oppPro_d.clear();
  const bool device = solver.device_ready;
  oppRes[0] = pmg_restriction_1d(order, (int)in->mg_levels[1]);
  for (int n = 1; n < nLevels; n++) {
    const int P = (int)in->mg_levels[n], P_pro = (int)in->mg_levels[n - 1];
    auto g = std::make_shared<FRSolver>(in, P);
    g->own_stream = false;
    g->stream = solver.stream;
    g->setup_host();
    if (device) g->setup_device();
    grids.push_back(g);
    oppPro[n] = pmg_prolongation_1d(P, P_pro);
    if (n + 1 < nLevels) oppRes[n] = pmg_restriction_1d(P, (int)in->mg_levels[n + 1]);
  }
  for (int n = 0; n < nLevels; n++) {
    oppRes_d.emplace_back();
    oppPro_d.emplace_back();
    sourcesBT.emplace_back();
    solutionsBT.emplace_back();
    correctionsBT.emplace_back();
  }
  if (!device) return;
  auto upload = [&](DevBuf &b, const std::vector<double> &v) {
    b.alloc(v.size() * sizeof(double));
    if (!v.empty()) hip_or_throw(hipMemcpy(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
  };
  for (int n = 0; n < nLevels; n++) {
    upload(oppRes_d[n], oppRes[n]);
    upload(oppPro_d[n], oppPro[n]);
    if (n == 0) continue;
    const size_t bytes = grids[n]->eles->state_size() * sizeof(double);
    for (auto *b : {&sourcesBT[n], &solutionsBT[n], &correctionsBT[n]}) {
      b->alloc(bytes);
      hip_or_throw(hipMemsetAsync(b->p, 0, bytes, solver.stream), "hipMemsetAsync");
    }
  }
  hip_or_throw(hipStreamSynchronize(solver.stream), "hipStreamSynchronize");
}

void PMGrid::transfer(const FRSolver &from, const FRSolver &to, const DevBuf &A, double alpha, const double *in,
                      double beta, double *out) {
  const Elements &ef = *from.eles;
  if (!A.p) raise("p-multigrid: transfer table missing for this pair of levels");
  status_or_throw(zfr_pmg_transfer(ef.nDims, ef.nVars, ef.nEles, from.order + 1, to.order + 1, A.as<double>(),
                                   alpha, in, beta, out, from.stream));
}

static int level_of(const std::vector<std::shared_ptr<FRSolver>> &grids, const FRSolver &g) {
  for (size_t n = 1; n < grids.size(); n++)
    if (grids[n].get() == &g) return (int)n;
  return 0;
}

// PMGrid::restrict_pmg (src/multigrid.cpp:335-392): solution and stage-0 residual.  The
// restricted residual goes to res_out (the coarse divF in the reference; the cycle passes
// the coarse level's source buffer, where compute_source_term consumes it).
void PMGrid::restrict_pmg(FRSolver &grid_f, FRSolver &grid_c, double *res_out) {
  const DevBuf &R = oppRes_d[level_of(grids, grid_f)];
  transfer(grid_f, grid_c, R, 1.0, grid_f.eles->U_spts_d.as<double>(), 0.0, grid_c.eles->U_spts_d.as<double>());
  transfer(grid_f, grid_c, R, 1.0, grid_f.eles->divF_d.as<double>(), 0.0, res_out);
}

// PMGrid::prolong_err (src/multigrid.cpp:394-433): U_f += rel_fac * P correction_c.
void PMGrid::prolong_err(FRSolver &grid_c, const double *correction, FRSolver &grid_f) {
  transfer(grid_c, grid_f, oppPro_d[level_of(grids, grid_c)], input->rel_fac, correction, 1.0,
           grid_f.eles->U_spts_d.as<double>());
}

// PMGrid::prolong_U (src/multigrid.cpp:435-472): U_f = P U_c.
void PMGrid::prolong_U(FRSolver &grid_c, FRSolver &grid_f) {
  transfer(grid_c, grid_f, oppPro_d[level_of(grids, grid_c)], 1.0, grid_c.eles->U_spts_d.as<double>(), 0.0,
           grid_f.eles->U_spts_d.as<double>());
}

// PMGrid::compute_source_term (src/multigrid.cpp:474-518): source = restricted residual
// minus the coarse residual at the restricted state.  source holds the restricted
// residual on entry (restrict_pmg writes it there); the residual's last kernel subtracts.
void PMGrid::compute_source_term(FRSolver &grid, double *source) { grid.residual_source(3, source, nullptr); }

// PMGrid::v_cycle (src/multigrid.cpp:105-261) from `level` (0 = the fine solver) down.
// Each coarse level's stored solution is written by its first smoothing step (the state
// it starts from), its correction (U - stored solution, the CPU path's form; the GPU
// path of the reference leaves the difference in U as well, which nothing reads
// afterwards) by its last one.  With fused steps the residual restricted to the next
// level (with the source term added) and the restricted U are written by the residual's
// last kernel, and the correction is prolonged into the next finer level by the last
// step's final combination; otherwise separate launches do the same arithmetic.
void PMGrid::v_cycle(FRSolver &solver, int lvl) {
  auto level_at = [&](int n) -> FRSolver & { return n == lvl ? solver : *grids[n]; };
  // residual (+ source) of level n, restricted with U into level n + 1
  auto restrict_from = [&](int n, const double *src) {
    FRSolver &g = level_at(n), &gc = *grids[n + 1];
    double *Rc = sourcesBT[n + 1].as<double>();
    if (g.fuse_rk) {
      g.residual_restrict(src, oppRes_d[n].as<double>(), gc.order + 1, gc.eles->U_spts_d.as<double>(), Rc);
    } else {
      if (src) g.residual_source(4, nullptr, src);
      else g.compute_residual(0);
      restrict_pmg(g, gc, Rc);
    }
  };
  // the correction of level n is prolonged by its last step's final combination when the
  // step is fused and the table fits the kernel's operator-table region
  auto fused_prolong = [&](int n) {
    return grids[n]->fuse_rk && zfr::pmg_prolong_fits(grids[n]->order + 1, level_at(n - 1).order + 1);
  };
  auto prolong = [&](int n) -> FRSolver::Prolong {
    FRSolver &gf = level_at(n - 1);
    FRSolver::Prolong p;
    p.A = oppPro_d[n].as<double>();
    p.n = gf.order + 1;
    p.alpha = input->rel_fac;
    p.out = gf.eles->U_spts_d.as<double>();
    return p;
  };
  // steps smoothing steps of level n under its source term: the first stores the solution
  // (store), the last writes the correction and, fused, prolongs it (finish)
  // (form: the first step's stage 0 forms the source term from the restricted residual)
  auto smooth = [&](int n, unsigned steps, bool store, bool finish, bool form) {
    FRSolver &g = *grids[n];
    double *src = sourcesBT[n].as<double>(), *sol = solutionsBT[n].as<double>(),
           *corr = correctionsBT[n].as<double>();
    for (unsigned s = 0; s < steps; s++) {
      const bool last = finish && s + 1 == steps;
      g.update(src, (store && s == 0) ? sol : nullptr, last ? corr : nullptr, last ? sol : nullptr,
               (last && fused_prolong(n)) ? prolong(n) : FRSolver::Prolong{}, form && s == 0);
    }
  };
  // the correction of a level without smoothing steps, and the unfused prolongation
  auto finish_level = [&](int n, bool stepped) {
    FRSolver &g = *grids[n];
    double *sol = solutionsBT[n].as<double>(), *corr = correctionsBT[n].as<double>();
    if (!stepped)
      status_or_throw(zfr_axpby((long)g.eles->state_size(), 1.0, g.eles->U_spts_d.as<double>(), -1.0, sol, corr,
                                g.stream));
    if (!stepped || !fused_prolong(n)) prolong_err(g, corr, level_at(n - 1));
  };

  if (lvl != nLevels - 1) restrict_from(lvl, nullptr);
  for (int n = lvl + 1; n < nLevels; n++) {
    FRSolver &g = *grids[n];
    Elements &e = *g.eles;
    const unsigned steps = input->mg_steps[n];
    const bool coarsest = n == nLevels - 1;
    // the source term is the residual at the restricted state subtracted from the
    // restricted residual; a fused first step forms it in its stage 0, whose residual is
    // that residual
    const bool form = g.fuse_rk && steps > 0 && g.input->nStages > 1;
    if (!form) compute_source_term(g, sourcesBT[n].as<double>());
    if (steps == 0)
      hip_or_throw(hipMemcpyAsync(solutionsBT[n].p, e.U_spts_d.p, e.state_size() * sizeof(double),
                                  hipMemcpyDeviceToDevice, g.stream), "hipMemcpyAsync");
    // the coarsest level is not smoothed again on the way up: its last step here ends it
    smooth(n, steps, true, coarsest, form);
    if (coarsest) finish_level(n, steps > 0);
    else restrict_from(n, sourcesBT[n].as<double>());
  }
  for (int n = nLevels - 2; n > lvl; n--) {
    const unsigned steps = input->mg_steps[n];
    smooth(n, steps, false, true, false);
    finish_level(n, steps > 0);
  }
}

PMGrid::~PMGrid() {
  drop_graph();
  if (persist) zfr_persist_destroy(persist);
}

void PMGrid::drop_graph() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  graph_exec = nullptr;
  graph = nullptr;
}

void PMGrid::set_persistent(bool on) {
  use_persistent = on;
  if (persist) zfr_persist_destroy(persist);
  persist = nullptr;
  persist_key.clear();
}

std::vector<double> PMGrid::graph_key(const FRSolver &solver) const {
  std::vector<double> k = {input->dt, input->rel_fac, (double)solver.fuse_rk};
  for (int t : solver.eles->tile_key()) k.push_back((double)t);
  for (int n = 1; n < nLevels; n++) {
    for (int t : grids[n]->eles->tile_key()) k.push_back((double)t);
    k.push_back((double)grids[n]->fuse_rk);
  }
  for (int n = 0; n < nLevels; n++) k.push_back((double)input->mg_steps[n]);
  return k;
}

std::vector<int> PMGrid::record_cycle(FRSolver &solver, const std::function<void()> &begin,
                                      const std::function<void()> &abort, const std::function<void()> &end) {
  std::vector<FRSolver *> lv = {&solver};
  for (int n = 1; n < nLevels; n++) lv.push_back(grids[n].get());
  std::vector<bool> saved;
  std::vector<int> before;
  for (auto *g : lv) {
    saved.push_back(g->use_graph);
    g->use_graph = false;  // the levels' own step graphs are not nested into the recording
    before.push_back(g->current_iter);
  }
  auto restore = [&] {
    for (size_t i = 0; i < lv.size(); i++) lv[i]->use_graph = saved[i];
  };
  begin();
  try {
    run_v(solver);
  } catch (...) {
    abort();
    restore();
    throw;
  }
  try {
    end();
  } catch (...) {
    restore();
    throw;
  }
  std::vector<int> iters;
  for (size_t i = 0; i < lv.size(); i++) {
    lv[i]->use_graph = saved[i];
    iters.push_back(lv[i]->current_iter - before[i]);
    lv[i]->current_iter = before[i];  // the recording ran nothing; each replay advances them
    lv[i]->flow_time -= iters.back() * input->dt;
  }
  return iters;
}

void PMGrid::advance_counters(FRSolver &solver, const std::vector<int> &iters) {
  for (int n = 0; n < nLevels; n++) {
    FRSolver &g = level(solver, n);
    g.prev_time = g.flow_time;
    g.current_iter += iters[n];
    g.flow_time += iters[n] * input->dt;
  }
}

// What a V-cycle must be to run as one persistent launch: only Euler divergence launches
// (inviscid levels, fused steps, no level without smoothing steps, every prolongation
// fused).  Anything else would need a launch the plan cannot hold.
void PMGrid::check_persistent(FRSolver &solver) {
  if (input->viscous)
    raise("persistent V-cycle: inviscid runs only (a viscous residual is two or three kernels of other kinds)");
  if (input->nStages != 4) raise("persistent V-cycle: RK44 steps only");
  for (int n = 0; n < nLevels; n++) {
    FRSolver &g = level(solver, n);
    if (!g.fuse_rk) raise("persistent V-cycle: needs fused steps (set_step_mode fuse_rk)");
    if (n > 0 && input->mg_steps[n] == 0) raise("persistent V-cycle: every coarse level needs a smoothing step");
    if (n > 0 && !zfr::pmg_prolong_fits(g.order + 1, level(solver, n - 1).order + 1))
      raise("persistent V-cycle: the prolongation from order " + std::to_string(g.order) + " to " +
            std::to_string(level(solver, n - 1).order) + " is not fused into the step (its table is too large)");
  }
}

// One V-cycle: mg_steps[0] fine steps, then v_cycle from the fine level.
void PMGrid::run_v(FRSolver &solver) {
  for (unsigned s = 0; s < input->mg_steps[0]; s++) solver.update();
  v_cycle(solver, 0);
}

// PMGrid::cycle (src/multigrid.cpp:263-333).  No history or solution files are written.
// With use_graph a V-cycle is captured once on the fine stream (every level runs on it)
// and replayed; the iteration counters and times the captured updates would have
// advanced are advanced on the host for each replay.
void PMGrid::cycle(FRSolver &solver) {
  if (!solver.device_ready) raise("solver is not set up on the device (call setup_solver)");
  const std::string type = input->mg_cycle;
  if (type == "V" && use_persistent) {
    const auto k = graph_key(solver);
    if (!persist || k != persist_key) {
      check_persistent(solver);
      if (persist) zfr_persist_destroy(persist);
      persist = nullptr;
      zfr_persist_t *p = nullptr;
      persist_iters = record_cycle(
          solver, [&] { status_or_throw(zfr_persist_record_begin()); },
          [&] { (void)zfr_persist_record_end(nullptr); }, [&] { status_or_throw(zfr_persist_record_end(&p)); });
      persist = p;
      persist_key = k;
    }
    status_or_throw(zfr_persist_launch(persist, solver.stream));
    advance_counters(solver, persist_iters);
    return;
  }
  if (type == "V" && use_graph) {
    const auto k = graph_key(solver);
    if (!graph_exec || k != key) {
      drop_graph();
      iters_per_cycle = record_cycle(
          solver,
          [&] {
            hip_or_throw(hipStreamBeginCapture(solver.stream, hipStreamCaptureModeThreadLocal),
                         "hipStreamBeginCapture");
          },
          [&] {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(solver.stream, &g);
            if (g) (void)hipGraphDestroy(g);
          },
          [&] { hip_or_throw(hipStreamEndCapture(solver.stream, &graph), "hipStreamEndCapture"); });
      hip_or_throw(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      key = k;
    }
    hip_or_throw(hipGraphLaunch(graph_exec, solver.stream), "hipGraphLaunch");
    advance_counters(solver, iters_per_cycle);
    return;
  }
  if (type == "V") {
    run_v(solver);
  } else if (type == "FMG") {
    for (int n = nLevels - 1; n > 0; n--) {
      for (unsigned c = 0; c < input->FMG_vcycles; c++) {
        grids[n]->update();
        v_cycle(*grids[n], n);
      }
      grids[n]->update();
      if (n > 1) prolong_U(*grids[n], *grids[n - 1]);
    }
    prolong_U(*grids[1], solver);
    std::memset(input->mg_cycle, 0, sizeof(input->mg_cycle));
    std::strcpy(input->mg_cycle, "V");
    cycle(solver);
  } else {
    raise("Multigrid cycle type unknown!");
  }
}

}  // namespace zfr
