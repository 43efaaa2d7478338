// capi.cpp -- the Zefr object and the extern "C" entry points of include/zefr_amd.h.
//
// Zefr mirrors the reference's library object (include/zefr.hpp:24-209,
// src/zefr.cpp:222-480): read_input -> init_inputs -> setup_solver -> do_step /
// do_rk_stage*.  Every entry point catches exceptions and reports them through
// zefr_last_error() with a nonzero status, which the Python layer raises as
// RuntimeError (the reference has no exception translation at its SWIG boundary).
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"
#include "zefr_amd.h"

namespace zfr {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

}  // namespace zfr

struct zefr_s {
  zefr_input_t input;
  bool have_input = false;
  std::unique_ptr<zfr::FRSolver> solver;
  std::unique_ptr<zfr::PMGrid> pmg;  // with p_multi (src/zefr.cpp:321-328)

  zfr::FRSolver &s() {
    if (!solver) zfr::raise("solver is not set up (call setup_solver first)");
    return *solver;
  }
  zfr::Elements &el() { return *s().eles; }

  void copy_u_to_host() {
    auto &e = el();
    e.U_host.resize(e.state_size());
    zfr::hip_or_throw(hipMemcpyAsync(e.U_host.data(), e.U_spts_d.p, e.U_spts_d.bytes, hipMemcpyDeviceToHost,
                                     s().stream), "D2H U_spts");
    s().synchronize();
  }
  void copy_du_to_host() {
    auto &e = el();
    s().compute_gradient_export();
    e.dU_host.resize((size_t)e.nDims * e.state_size());
    zfr::hip_or_throw(hipMemcpyAsync(e.dU_host.data(), e.dU_spts_d.p, e.dU_spts_d.bytes, hipMemcpyDeviceToHost,
                                     s().stream), "D2H dU_spts");
    s().synchronize();
  }
  void strides(int *ele, int *spt, int *var, int *dim) {
    auto &e = el();
    if (ele) *ele = e.nVars * e.nSpts;
    if (spt) *spt = 1;
    if (var) *var = e.nSpts;
    if (dim) *dim = e.nVars * e.nSpts;  // gradient layout [E][D][V][nS]
  }
};

namespace {
zefr_t *g_zefr = nullptr;

void check_etype(int etype) {
  if (etype != 0) zfr::raise("single element-type meshes only: etype must be 0");
}
}  // namespace

extern "C" {

const char *zefr_last_error(void) { return zfr::last_error(); }

int zefr_read_input_file(const char *file, zefr_input_t *out) {
  return zfr::guard([&] { zfr::read_input_file(file, *out); });
}

int zefr_initialize_inputs(zefr_input_t *inp) {
  return zfr::guard([&] { zfr::initialize_inputs(*inp); });
}

zefr_t *zefr_create(void) {
  return zfr::guard_val<zefr_t *>(nullptr, [&] {
    auto *z = new zefr_s();
    std::memset(&z->input, 0, sizeof(z->input));
    return z;
  });
}

void zefr_destroy(zefr_t *z) {
  if (z == g_zefr) g_zefr = nullptr;
  delete z;
}

int zefr_read_input(zefr_t *z, const char *file) {
  return zfr::guard([&] {
    zfr::read_input_file(file, z->input);
    z->have_input = true;
  });
}

zefr_input_t *zefr_get_input(zefr_t *z) { return z ? &z->input : nullptr; }

int zefr_init_inputs(zefr_t *z) {
  return zfr::guard([&] { zfr::initialize_inputs(z->input); z->have_input = true; });
}

int zefr_setup_host(zefr_t *z) {
  return zfr::guard([&] {
    z->pmg.reset();
    z->solver = std::make_unique<zfr::FRSolver>(&z->input);
    z->solver->setup_host();
    if (z->input.p_multi) {
      z->pmg = std::make_unique<zfr::PMGrid>();
      z->pmg->setup(&z->input, *z->solver);
    }
  });
}

int zefr_setup_solver(zefr_t *z) {
  return zfr::guard([&] {
    int ndev = 0;
    zfr::hip_or_throw(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (ndev < 1) zfr::raise("no HIP device available");
    z->pmg.reset();
    z->solver = std::make_unique<zfr::FRSolver>(&z->input);
    z->solver->setup_host();
    z->solver->setup_device();
    if (z->input.p_multi) {
      z->pmg = std::make_unique<zfr::PMGrid>();
      z->pmg->setup(&z->input, *z->solver);
    }
  });
}

int zefr_do_step(zefr_t *z) {
  return zfr::guard([&] {
    if (z->pmg) z->pmg->cycle(z->s());
    else z->s().update();
    z->input.iter++;
  });
}

int zefr_do_rk_stage(zefr_t *z, int iter, int stage) {
  return zfr::guard([&] {
    z->input.iter = iter;
    z->s().step_RK_stage(stage);
    if (stage == (int)z->input.nStages - 1) z->input.iter++;
  });
}

int zefr_do_rk_stage_start(zefr_t *z, int, int stage) {
  return zfr::guard([&] { z->s().step_RK_stage_start(stage); });
}

int zefr_do_rk_stage_mid(zefr_t *z, int, int stage) {
  return zfr::guard([&] { z->s().step_RK_stage_mid(stage); });
}

int zefr_do_rk_stage_finish(zefr_t *z, int iter, int stage) {
  return zfr::guard([&] {
    z->s().step_RK_stage_finish(stage);
    z->input.iter = iter;
    z->s().current_iter = iter;
  });
}

int zefr_write_residual(zefr_t *z, double *res) {
  return zfr::guard([&] { z->s().report_residuals(res); });
}

int zefr_compute_residual(zefr_t *z, unsigned int stage) {
  return zfr::guard([&] { z->s().compute_residual(stage); });
}
int zefr_compute_residual_start(zefr_t *z, unsigned int stage) {
  return zfr::guard([&] { z->s().compute_residual_start(stage); });
}
int zefr_compute_residual_mid(zefr_t *z, unsigned int stage) {
  return zfr::guard([&] { z->s().compute_residual_mid(stage); });
}
int zefr_compute_residual_finish(zefr_t *z, unsigned int stage) {
  return zfr::guard([&] { z->s().compute_residual_finish(stage); });
}
int zefr_update(zefr_t *z) {
  return zfr::guard([&] { z->s().update(); });
}
int zefr_synchronize(zefr_t *z) {
  return zfr::guard([&] { z->s().synchronize(); });
}

double *zefr_get_u_spts_d(zefr_t *z, int *ele_stride, int *spt_stride, int *var_stride, int etype) {
  return zfr::guard_val<double *>(nullptr, [&] {
    check_etype(etype);
    z->strides(ele_stride, spt_stride, var_stride, nullptr);
    return z->el().U_spts_d.as<double>();
  });
}

double *zefr_get_du_spts_d(zefr_t *z, int *ele_stride, int *spt_stride, int *var_stride, int *dim_stride,
                           int etype) {
  return zfr::guard_val<double *>(nullptr, [&] {
    check_etype(etype);
    z->s().compute_gradient_export();
    z->s().synchronize();
    auto &e = z->el();
    if (ele_stride) *ele_stride = e.nDims * e.nVars * e.nSpts;
    if (spt_stride) *spt_stride = 1;
    if (var_stride) *var_stride = e.nSpts;
    if (dim_stride) *dim_stride = e.nVars * e.nSpts;
    return e.dU_spts_d.as<double>();
  });
}

double *zefr_get_u_spts(zefr_t *z, int *ele_stride, int *spt_stride, int *var_stride, int etype) {
  return zfr::guard_val<double *>(nullptr, [&] {
    check_etype(etype);
    z->copy_u_to_host();
    z->strides(ele_stride, spt_stride, var_stride, nullptr);
    return z->el().U_host.data();
  });
}

double *zefr_get_du_spts(zefr_t *z, int *ele_stride, int *spt_stride, int *var_stride, int *dim_stride,
                         int etype) {
  return zfr::guard_val<double *>(nullptr, [&] {
    check_etype(etype);
    z->copy_du_to_host();
    auto &e = z->el();
    if (ele_stride) *ele_stride = e.nDims * e.nVars * e.nSpts;
    if (spt_stride) *spt_stride = 1;
    if (var_stride) *var_stride = e.nSpts;
    if (dim_stride) *dim_stride = e.nVars * e.nSpts;
    return e.dU_host.data();
  });
}

double zefr_get_u_spt(zefr_t *z, int ele, int spt, int var) {
  return zfr::guard_val<double>(0.0 / 0.0, [&] {
    auto &e = z->el();
    if (ele < 0 || ele >= e.nEles || spt < 0 || spt >= e.nSpts || var < 0 || var >= e.nVars)
      zfr::raise("get_u_spt: index out of range");
    double v;
    zfr::hip_or_throw(hipMemcpy(&v, e.U_spts_d.as<double>() + ((size_t)ele * e.nVars + var) * e.nSpts + spt,
                                sizeof(double), hipMemcpyDeviceToHost), "D2H");
    return v;
  });
}

double zefr_get_grad_spt(zefr_t *z, int ele, int spt, int dim, int var) {
  return zfr::guard_val<double>(0.0 / 0.0, [&] {
    auto &e = z->el();
    if (ele < 0 || ele >= e.nEles || spt < 0 || spt >= e.nSpts || var < 0 || var >= e.nVars || dim < 0 ||
        dim >= e.nDims)
      zfr::raise("get_grad_spt: index out of range");
    z->s().compute_gradient_export();
    z->s().synchronize();
    double v;
    zfr::hip_or_throw(
        hipMemcpy(&v, e.dU_spts_d.as<double>() + (((size_t)ele * e.nDims + dim) * e.nVars + var) * e.nSpts + spt,
                  sizeof(double), hipMemcpyDeviceToHost), "D2H");
    return v;
  });
}

int zefr_get_dims(zefr_t *z, int *out) {
  return zfr::guard([&] {
    auto &s = z->s();
    auto &e = *s.eles;
    const int v[10] = {e.nEles, e.nSpts, e.nFpts, e.nVars, e.nDims, s.conn.nGfpts, s.conn.nGfpts_int,
                       (int)z->input.nStages, e.order, e.etype};
    std::memcpy(out, v, sizeof(v));
  });
}

int zefr_copy_state(zefr_t *z, int which, int stage, double *host, long n, int to_device) {
  return zfr::guard([&] {
    auto &e = z->el();
    const size_t st = e.state_size();
    if ((size_t)n < st) zfr::raise("copy_state: buffer too small");
    double *dev = nullptr;
    if (which == 0) dev = e.U_spts_d.as<double>();
    else if (which == 1) dev = e.U_ini_d.as<double>();
    else if (which == 2) {
      if (stage < 0 || stage >= (int)z->input.nStages) zfr::raise("copy_state: stage out of range");
      dev = e.divF_d.as<double>() + (size_t)stage * st;
    } else zfr::raise("copy_state: which must be 0 (U), 1 (U_ini) or 2 (divF)");
    if (!dev) zfr::raise("copy_state: device arrays are not allocated");
    z->s().synchronize();
    zfr::hip_or_throw(hipMemcpy(to_device ? (void *)dev : (void *)host, to_device ? (void *)host : (void *)dev,
                                st * sizeof(double), to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost),
                      "copy_state");
  });
}

int zefr_pmg_copy_state(zefr_t *z, int level, int which, double *host, long n, int to_device) {
  return zfr::guard([&] {
    if (!z->pmg) zfr::raise("pmg_copy_state: p-multigrid is not set up (p_multi 0)");
    auto &pm = *z->pmg;
    if (level < 1 || level >= pm.nLevels) zfr::raise("pmg_copy_state: level must be in 1..nLevels-1");
    auto &g = *pm.grids[level];
    auto &e = *g.eles;
    const size_t st = e.state_size();
    if ((size_t)n < st) zfr::raise("pmg_copy_state: buffer too small");
    double *dev = nullptr;
    switch (which) {
      case 0: dev = e.U_spts_d.as<double>(); break;
      case 2: dev = e.divF_d.as<double>(); break;
      case 3: dev = pm.sourcesBT[level].as<double>(); break;
      case 4: dev = pm.solutionsBT[level].as<double>(); break;
      case 5: dev = pm.correctionsBT[level].as<double>(); break;
      default: zfr::raise("pmg_copy_state: which must be 0 (U), 2 (divF), 3 (source), 4 (solution) or 5 (correction)");
    }
    if (!dev) zfr::raise("pmg_copy_state: device arrays are not allocated");
    z->s().synchronize();
    zfr::hip_or_throw(hipMemcpy(to_device ? (void *)dev : (void *)host, to_device ? (void *)host : (void *)dev,
                                st * sizeof(double), to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost),
                      "pmg_copy_state");
  });
}

int zefr_write_solution(zefr_t *z) {
  return zfr::guard([&] {
    if (z->input.write_pyfr) z->s().write_solution_pyfr(z->input.output_prefix);
  });
}

int zefr_restart_solution(zefr_t *z) {
  return zfr::guard([&] {
    if (z->input.restart) z->s().restart_solution();
  });
}

int zefr_get_fpt_maps(zefr_t *z, int *fpt2gfpt, signed char *slot) {
  return zfr::guard([&] {
    auto &c = z->s().conn;
    std::memcpy(fpt2gfpt, c.fpt2gfpt.data(), c.fpt2gfpt.size() * sizeof(int));
    std::memcpy(slot, c.slot.data(), c.slot.size());
  });
}

long zefr_get_host_array(zefr_t *z, const char *name, double *out, long n) {
  return zfr::guard_val<long>(-1, [&]() -> long {
    auto &s = z->s();
    auto &e = *s.eles;
    const std::string nm = name;
    std::vector<double> tmp;
    const std::vector<double> *src = nullptr;
    if (nm == "oppE" || nm == "oppD" || nm == "oppD_fpts" || nm == "oppDiv" || nm == "oppDiv_fpts") {
      tmp = e.re.dense_operator(nm);
      src = &tmp;
    } else if (nm == "coord_spts" || nm == "coord_fpts" || nm == "jaco_det_spts" || nm == "inv_jaco_spts" ||
               nm == "inv_jaco_fpts") {
      tmp = e.export_array(nm);
      src = &tmp;
    } else if (nm == "norm" || nm == "dA") {
      tmp = s.faces->export_array(nm, e, s.conn);
      src = &tmp;
    } else if (nm == "weights_spts") src = &e.re.weights_spts;
    else if (nm == "ops1d") src = &e.re.ops1d;
    else if (nm == "U_init") src = &e.U_host;
    else zfr::raise("unknown host array " + nm);
    if (!out) return (long)src->size();
    if ((long)src->size() > n) zfr::raise("get_host_array: buffer too small");
    std::memcpy(out, src->data(), src->size() * sizeof(double));
    return (long)src->size();
  });
}

int zefr_autotune(zefr_t *z, int reps) {
  return zfr::guard_val<int>(-1, [&] { return z->s().autotune(reps < 1 ? 3 : reps); });
}

int zefr_time_residual_kernels(zefr_t *z, int reps, double *ms_out) {
  return zfr::guard([&] { z->s().time_kernels(reps, ms_out); });
}

int zefr_residual_at(zefr_t *z, const double *U, double *divF, void *stream) {
  return zfr::guard([&] { z->s().residual_at(U, divF, (hipStream_t)stream); });
}

int zefr_residual_vjp(zefr_t *z, const double *U, const double *w, double *v, void *stream) {
  return zfr::guard([&] { z->s().residual_vjp(U, w, v, (hipStream_t)stream); });
}

int zefr_set_tile(zefr_t *z, int eles_per_block) {
  return zfr::guard([&] {
    if (eles_per_block < 0) zfr::raise("tile must be >= 0");
    z->el().tile = eles_per_block;
    for (int &t : z->el().tile_phase) t = 0;
  });
}

int zefr_set_tiles(zefr_t *z, const int *tiles) {
  return zfr::guard([&] {
    for (int k = 0; k < 3; k++)
      if (tiles[k] < 0) zfr::raise("tiles must be >= 0");
    for (int k = 0; k < 3; k++) z->el().tile_phase[k] = tiles[k];
  });
}

int zefr_get_tiles(zefr_t *z, int *tiles) {
  return zfr::guard([&] {
    for (int k = 0; k < 3; k++) tiles[k] = z->el().tile_for(k);
  });
}

int zefr_set_step_mode(zefr_t *z, int fuse_rk, int use_graph) {
  return zfr::guard([&] {
    z->s().set_step_mode(fuse_rk != 0, use_graph != 0);
    if (z->pmg) {
      for (int n = 1; n < z->pmg->nLevels; n++) z->pmg->grids[n]->set_step_mode(fuse_rk != 0, false);
      z->pmg->use_graph = use_graph != 0;
    }
  });
}

int zefr_set_pmg_persistent(zefr_t *z, int on) {
  return zfr::guard([&] {
    if (!z->pmg) zfr::raise("set_pmg_persistent: p-multigrid is not set up (p_multi 0)");
    z->pmg->set_persistent(on != 0);
  });
}

// ---------------------------------------------------------------- batches
// Residuals of independent solvers: a multi-case plan over their element blocks (one
// launch per kernel, shape and flux branch; zfr_multi_create), captured once as a HIP
// graph on the batch's own stream, so one graph launch evaluates every member.
struct zefr_batch_s {
  std::vector<zfr::FRSolver *> solvers;
  hipStream_t stream = nullptr;
  zfr_multi_t *plan = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~zefr_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (plan) zfr_multi_destroy(plan);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

zefr_batch_t *zefr_batch_create(zefr_t **zs, int n) {
  zefr_batch_t *b = nullptr;
  const int rc = zfr::guard([&] {
    if (!zs || n < 1) zfr::raise("batch_create: need at least one solver");
    auto batch = std::make_unique<zefr_batch_s>();
    std::vector<zfr_ele_desc_t> descs;
    std::vector<const double *> U;
    std::vector<double *> divF;
    for (int i = 0; i < n; i++) {
      if (!zs[i]) zfr::raise("batch_create: null solver");
      zfr::FRSolver &s = zs[i]->s();
      if (!s.device_ready) zfr::raise("batch_create: every solver must be set up on the device");
      s.synchronize();
      batch->solvers.push_back(&s);
      descs.push_back(s.eles->desc(*s.input));
      U.push_back(s.eles->U_spts_d.as<double>());
      divF.push_back(s.eles->divF_d.as<double>());
    }
    zfr::status_or_throw(zfr_multi_create(descs.data(), U.data(), divF.data(), n, &batch->plan));
    zfr::hip_or_throw(hipStreamCreateWithFlags(&batch->stream, hipStreamNonBlocking), "hipStreamCreate");
    hipStream_t s0 = batch->stream;
    zfr::hip_or_throw(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const int lrc = zfr_multi_residual(batch->plan, s0);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s0, &g);
    if (lrc) {
      if (g) (void)hipGraphDestroy(g);
      zfr::status_or_throw(lrc);
    }
    zfr::hip_or_throw(ec, "hipStreamEndCapture");
    batch->graph = g;
    zfr::hip_or_throw(hipGraphInstantiate(&batch->exec, batch->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    b = batch.release();
  });
  return rc ? nullptr : b;
}

int zefr_batch_launches(zefr_batch_t *b) { return b ? zfr_multi_launches(b->plan) : -1; }

int zefr_batch_residual(zefr_batch_t *b) {
  return zfr::guard([&] {
    if (!b) zfr::raise("batch_residual: null batch");
    zfr::hip_or_throw(hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
  });
}

int zefr_batch_synchronize(zefr_batch_t *b) {
  return zfr::guard([&] {
    if (!b) zfr::raise("batch_synchronize: null batch");
    zfr::hip_or_throw(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
  });
}

void zefr_batch_destroy(zefr_batch_t *b) { delete b; }

// ---------------------------------------------------------------- singleton
int zefr_initialize(const char *input_file) {
  return zfr::guard([&] {
    if (!g_zefr) g_zefr = new zefr_s();
    zfr::read_input_file(input_file, g_zefr->input);
    g_zefr->have_input = true;
  });
}

zefr_t *zefr_get_zefr_object(void) { return g_zefr; }

void zefr_set_zefr_object(zefr_t *z) {
  if (g_zefr && g_zefr != z) delete g_zefr;
  g_zefr = z;
}

int zefr_finalize(void) {
  delete g_zefr;
  g_zefr = nullptr;
  return 0;
}

int zefr_use_gpus(void) { return 1; }

static zefr_t *singleton() {
  if (!g_zefr) zfr::set_error("zefr is not initialized (call zefr_initialize)");
  return g_zefr;
}

double *zefr_get_q_spts(int *ele_stride, int *spt_stride, int *var_stride, int etype) {
  zefr_t *z = singleton();
  return z ? zefr_get_u_spts(z, ele_stride, spt_stride, var_stride, etype) : nullptr;
}
double *zefr_get_dq_spts(int *ele_stride, int *spt_stride, int *var_stride, int *dim_stride, int etype) {
  zefr_t *z = singleton();
  return z ? zefr_get_du_spts(z, ele_stride, spt_stride, var_stride, dim_stride, etype) : nullptr;
}
double *zefr_get_q_spts_d(int *ele_stride, int *spt_stride, int *var_stride, int etype) {
  zefr_t *z = singleton();
  return z ? zefr_get_u_spts_d(z, ele_stride, spt_stride, var_stride, etype) : nullptr;
}
double *zefr_get_dq_spts_d(int *ele_stride, int *spt_stride, int *var_stride, int *dim_stride, int etype) {
  zefr_t *z = singleton();
  return z ? zefr_get_du_spts_d(z, ele_stride, spt_stride, var_stride, dim_stride, etype) : nullptr;
}
double zefr_get_q_spt(int ele, int spt, int var) {
  zefr_t *z = singleton();
  return z ? zefr_get_u_spt(z, ele, spt, var) : 0.0 / 0.0;
}

}  // extern "C"
