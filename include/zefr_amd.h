/* zefr_amd.h -- C ABI of the MI355X-native ZEFR residual path.
 *
 * Two layers, both plain C (pointers, sizes, POD structs; no torch or C++ types):
 *
 *  1. Solver interface.  The entry points a SWIG/ctypes/cffi binding of ZEFR's library
 *     interface binds: the `Zefr` object (include/zefr.hpp:24-209), the `zefr`
 *     namespace functions (include/zefr_interface.hpp:89-159) and `InputStruct`
 *     (include/inputstruct.hpp:113-194).  Each declaration cites the reference
 *     interface it replaces.  `zefr_t` is the `Zefr` object; the `zefr_*` singleton
 *     functions operate on the object `zefr_initialize` created, like the reference's
 *     global `ZEFR` (src/zefr_interface.cpp:5).
 *
 *  2. Kernel interface.  The thin layer between the host classes and the HIP kernels;
 *     it replaces `gimmik_mm_gpu` (include/gimmik.h:17-21) and the `*_wrapper`
 *     launchers (include/elements_kernels.h:7-12, include/faces_kernels.h:26-37,
 *     include/solver_kernels.h:18-22), which take C++ containers in the reference.
 *
 * Every function returning int returns 0 on success and a nonzero status otherwise;
 * `zefr_last_error()` then holds a thread-local message.  No entry point aborts.  The
 * kernel-level entry points check that the descriptor's arrays a launch touches are not
 * NULL before launching.
 *
 * Device data layout (element-major, "element-blocked"): per element e the solution
 * block is U[e][var][spt], so the reference's stride-based accessors still describe it
 * exactly: ele_stride = nVars*nSpts, var_stride = nSpts, spt_stride = 1.
 */
#ifndef ZEFR_AMD_H
#define ZEFR_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- input ---------- */

#define ZEFR_MAX_BOUNDS 16
#define ZEFR_NAME_LEN 64
#define ZEFR_MAX_LEVELS 8

/* Hot-path subset of InputStruct (include/inputstruct.hpp:113-194); field names and
 * meanings as in the reference.  Fields the residual path does not use are omitted. */
typedef struct zefr_input {
  unsigned int equation;          /* 1 = EulerNS (include/input.hpp:25-28)          */
  unsigned int nDims, order;
  int viscous;
  unsigned int fconv_type;        /* 0 = Rusanov (include/input.hpp:30-33), 2 = Roe   */
  unsigned int fvisc_type;        /* 1 = LDG                                         */
  double rus_k, ldg_b, ldg_tau;
  char dt_scheme[16];             /* "Euler" or "RK44"                               */
  unsigned int nStages;
  unsigned int dt_type;           /* 0 = supplied dt (the only supported mode)       */
  double dt;
  unsigned int n_steps, report_freq, write_freq, res_type, res_field;
  unsigned int ic_type, test_case;
  double gamma, prandtl, mu, R, S, T_gas, rt, c_sth;
  double rho_fs, P_fs, v_mag_fs, mach_fs, Re_fs, L_fs, T_fs, T_tot_fs, P_tot_fs;
  double V_fs[3], norm_fs[3];
  int fix_vis, disable_nondim;
  char meshfile[512];
  char output_prefix[256];
  char spt_type[32];
  int nMeshBounds;
  char meshBoundName[ZEFR_MAX_BOUNDS][ZEFR_NAME_LEN]; /* mesh_bound <name> <bc> pairs */
  char meshBoundBC[ZEFR_MAX_BOUNDS][ZEFR_NAME_LEN];
  unsigned int iter, initIter;
  double time;
  /* p-multigrid (include/inputstruct.hpp:116-122; src/input.cpp:137-142) */
  int p_multi;
  unsigned int nLevels;                       /* entries used in mg_levels / mg_steps    */
  unsigned int mg_levels[ZEFR_MAX_LEVELS];    /* orders, finest first (mg_levels[0]=order) */
  unsigned int mg_steps[ZEFR_MAX_LEVELS];     /* smoothing steps per level               */
  char mg_cycle[8];                           /* "V" or "FMG"                            */
  unsigned int FMG_vcycles;
  double rel_fac;                             /* prolongation relaxation factor          */
  /* solution files (include/inputstruct.hpp:115-118; src/input.cpp:130-147) */
  int restart;                                /* restart_solution reads a file           */
  unsigned int restart_type;                  /* 2 = PyFR .pyfrs (1 = ParaView: rejected) */
  unsigned int restart_iter;
  char restart_case[256];                     /* reads <case>/<case>-<iter>.pyfrs        */
  int write_pyfr;                             /* write_solution writes .pyfrs            */
  int write_paraview;                         /* read for compatibility; not written     */
} zefr_input_t;

typedef struct zefr_s zefr_t;

const char *zefr_last_error(void);

/* read_input_file (src/input.cpp:31-461), including the similarity-parameter and
 * non-dimensionalisation step of src/input.cpp:431-539. */
int zefr_read_input_file(const char *input_file, zefr_input_t *out);

/* initialize_inputs (src/input.cpp:542-562): re-apply nStages/dt_scheme and the
 * non-dimensionalisation after fields were changed by the caller. */
int zefr_initialize_inputs(zefr_input_t *inp);

/* ---------------------------------------------------------------- Zefr object ---- */

/* Zefr::Zefr (src/zefr.cpp:228-257) */
zefr_t *zefr_create(void);
void zefr_destroy(zefr_t *z);
/* Zefr::read_input (src/zefr.cpp:286-307) */
int zefr_read_input(zefr_t *z, const char *input_file);
/* Zefr::get_input (include/zefr.hpp:99): live pointer, edits take effect at init_inputs */
zefr_input_t *zefr_get_input(zefr_t *z);
/* Zefr::init_inputs (src/zefr.cpp:309-312) */
int zefr_init_inputs(zefr_t *z);
/* Zefr::setup_solver (src/zefr.cpp:314-363): mesh, connectivity, operators, metrics,
 * initial condition, upload to the current HIP device. */
int zefr_setup_solver(zefr_t *z);
/* The host-only half of setup_solver (no GPU needed): everything up to the upload. */
int zefr_setup_host(zefr_t *z);
/* Zefr::do_step (src/zefr.cpp:371-383): one full time step (FRSolver::update). */
int zefr_do_step(zefr_t *z);
/* Zefr::do_rk_stage (src/zefr.cpp:385-396) */
int zefr_do_rk_stage(zefr_t *z, int iter, int stage);
/* Zefr::do_rk_stage_start / _mid / _finish (src/zefr.cpp:398-430) */
int zefr_do_rk_stage_start(zefr_t *z, int iter, int stage);
int zefr_do_rk_stage_mid(zefr_t *z, int iter, int stage);
int zefr_do_rk_stage_finish(zefr_t *z, int iter, int stage);
/* Zefr::write_residual -> FRSolver::report_residuals (src/solver.cpp:6990-7137);
 * res[nVars] receives the per-variable norm divided by nDoF (no file is written). */
int zefr_write_residual(zefr_t *z, double *res);

/* FRSolver::compute_residual (src/solver.cpp:1494-1645) and its split
 * (src/solver.cpp:1647-1840) on the solver's stream. */
int zefr_compute_residual(zefr_t *z, unsigned int stage);
int zefr_compute_residual_start(zefr_t *z, unsigned int stage);
int zefr_compute_residual_mid(zefr_t *z, unsigned int stage);
int zefr_compute_residual_finish(zefr_t *z, unsigned int stage);
/* FRSolver::update with RK44 / Euler (src/solver.cpp:2599-2646, 2865-3097) */
int zefr_update(zefr_t *z);
/* Block the host until the solver's stream is idle. */
int zefr_synchronize(zefr_t *z);

/* Zefr::get_u_spts_d / get_du_spts_d (src/zefr.cpp:607-645): device pointer + strides.
 * etype is the element-type slot (0 for single-type meshes).  get_du_spts_d materialises
 * the physical gradient of the current state first.  dim_stride = nVars*nSpts. */
double *zefr_get_u_spts_d(zefr_t *z, int *ele_stride, int *spt_stride, int *var_stride, int etype);
double *zefr_get_du_spts_d(zefr_t *z, int *ele_stride, int *spt_stride, int *var_stride,
                           int *dim_stride, int etype);
/* Zefr::get_u_spts / get_du_spts: host mirror, refreshed from the device on each call
 * (the reference returns its possibly stale host array, src/zefr.cpp:352-363). */
double *zefr_get_u_spts(zefr_t *z, int *ele_stride, int *spt_stride, int *var_stride, int etype);
double *zefr_get_du_spts(zefr_t *z, int *ele_stride, int *spt_stride, int *var_stride,
                         int *dim_stride, int etype);
/* Zefr::get_u_spt / get_grad_spt (include/zefr.hpp:142-143) */
double zefr_get_u_spt(zefr_t *z, int ele, int spt, int var);
double zefr_get_grad_spt(zefr_t *z, int ele, int spt, int dim, int var);

/* Sizes: out[0..9] = nEles nSpts nFpts nVars nDims nGfpts nGfpts_int nStages order etype */
int zefr_get_dims(zefr_t *z, int *out);
/* Host <-> device copies of whole state arrays in the device layout [E][V][nS]:
 * which = 0 U_spts, 1 U_ini, 2 divF(stage).  n = capacity in doubles. */
int zefr_copy_state(zefr_t *z, int which, int stage, double *host, long n, int to_device);
/* Connectivity in the reference layout [nFpts][nEles] (after orient_fpts): global
 * flux point and slot of every element flux point (src/geometry.cpp:1727-1742). */
int zefr_get_fpt_maps(zefr_t *z, int *fpt2gfpt, signed char *slot);
/* Host arrays for inspection, in the reference layouts: "oppE", "oppD", "oppD_fpts",
 * "oppDiv", "oppDiv_fpts" (dense, from the 1D tables the kernels use), "coord_spts",
 * "coord_fpts", "jaco_det_spts", "inv_jaco_spts", "inv_jaco_fpts", "norm", "dA",
 * "weights_spts", "ops1d", "U_init".  Returns the count; with out == NULL only the
 * count is returned. */
long zefr_get_host_array(zefr_t *z, const char *name, double *out, long n);
/* Pick the elements-per-workgroup tiles for this solver's kernels by timing the
 * candidates on the device, each of a residual's three launches separately (see
 * zefr_set_tiles).  Returns the tile of the middle launch. */
int zefr_autotune(zefr_t *z, int reps);
int zefr_set_tile(zefr_t *z, int eles_per_block);
/* Per-launch tiles of a residual: tiles[0] the trace launch, [1] the gradient / one-sided
 * flux launch, [2] the divergence / face-correction launch (what zefr_autotune picks, each
 * by its own timing; 0 = the solver-wide tile).  get returns the tiles in effect. */
int zefr_set_tiles(zefr_t *z, const int *tiles);
int zefr_get_tiles(zefr_t *z, int *tiles);
/* How FRSolver::step_RK (src/solver.cpp:2865-3097) runs: fuse_rk folds each stage's RK
 * update into the residual's last kernel (default on); use_graph captures one whole time
 * step (the U_ini copy and every launch) as a HIP graph and replays it (default off;
 * recaptured when dt or the tile change).  Under p-multigrid the settings apply to every
 * level, and use_graph captures a whole V-cycle (all levels, transfers and updates) as one
 * graph instead.  Results are identical either way. */
int zefr_set_step_mode(zefr_t *z, int fuse_rk, int use_graph);
/* p-multigrid V-cycles as one persistent launch (new; no reference counterpart; built on
 * zfr_persist_*): with on != 0 and p_multi set, do_step records the V-cycle of
 * PMGrid::cycle (src/multigrid.cpp:263-333) once and replays it as one cooperative launch
 * (recorded again when dt, rel_fac, a tile or the fusion mode change).  For inviscid runs
 * with fused steps, at least one smoothing step on every coarse level and every
 * correction's prolongation fused into its level's last step; otherwise do_step fails
 * with the reason.  Results are identical to the eager and graph cycles. */
int zefr_set_pmg_persistent(zefr_t *z, int on);
/* Run `reps` residuals (stage 0) with HIP events around each kernel on the solver's
 * stream; ms_out[0..2] = average duration of the extrapolate / gradient / divergence
 * launch (gradient is 0 for inviscid runs).  No counterpart in the reference: this is the
 * timing hook of the benchmark. */
int zefr_time_residual_kernels(zefr_t *z, int reps, double *ms_out);
/* divF = R(U) for an arbitrary device state U (device layout [E][V][nS], as
 * get_u_spts_d) on the caller's `stream` (NULL: the null stream), with the kernels of
 * compute_residual (src/solver.cpp:1494-1645) but without touching the solver's state.
 * It uses the solver's trace scratch: do not overlap it with other work of the solver
 * (its own stream is non-blocking, so synchronise in between). */
int zefr_residual_at(zefr_t *z, const double *U, double *divF, void *stream);
/* v = (dR/dU)^T w at U (all device pointers, layout [E][V][nS]) on `stream` (NULL: the
 * null stream): the residual's vector-Jacobian product (new; no reference counterpart).
 * Euler, or Navier-Stokes with LDG beta = +-1/2.  Same scratch rule as residual_at. */
int zefr_residual_vjp(zefr_t *z, const double *U, const double *w, double *v, void *stream);
/* p-multigrid (PMGrid, include/multigrid.hpp:17-44): with p_multi set, setup_solver builds
 * the coarse FRSolvers of mg_levels (PMGrid::setup, src/multigrid.cpp:18-103) and do_step
 * runs PMGrid::cycle (src/multigrid.cpp:263-333) instead of one update.  The per-level
 * arrays are copied here in the device layout [E][V][(P+1)^nDims] of level `level`
 * (1..nLevels-1; level 0 is the solver itself, see zefr_copy_state): which = 0 U_spts,
 * 2 divF(0), 3 source term, 4 stored solution, 5 correction.  n = capacity in doubles. */
int zefr_pmg_copy_state(zefr_t *z, int level, int which, double *host, long n, int to_device);
/* Zefr::write_solution (src/zefr.cpp:445-453): with write_pyfr, the current solution as a
 * PyFR file <output_prefix>/<output_prefix>-<iter>.pyfrs (FRSolver::write_solution_pyfr,
 * src/solver.cpp:4275-4668; HDF5 loaded at run time, ZEFR_HDF5_LIBRARY overrides the
 * search).  ParaView / surface plot output is not written. */
int zefr_write_solution(zefr_t *z);
/* Zefr::restart_solution (src/zefr.cpp:366-370): with restart set, read
 * <restart_case>/<restart_case>-<restart_iter>.pyfrs (FRSolver::restart_pyfr,
 * src/solver.cpp:4670-4837): mesh_uuid check, tcurr and iteration, the solution
 * (interpolated when written at another order).  restart_type must be 2. */
int zefr_restart_solution(zefr_t *z);

/* Batched residuals of independent solvers (new; no reference counterpart): the
 * compute_residual(0) of n set-up solvers as one multi-case plan (zfr_multi_create: one
 * launch per kernel, shape and flux branch over all members), captured once as a HIP graph
 * on the batch's own stream, so that one graph launch evaluates all of them.  Meant for
 * many small cases whose residuals are too small to fill the GPU or to outlast the launch
 * cost.  zefr_batch_residual launches the graph on the batch's stream and returns;
 * zefr_batch_synchronize waits for it; zefr_batch_launches returns the kernel launches in
 * the graph.  The member solvers must have no work pending when the batch is launched,
 * and must outlive the batch (the graph holds their device arrays).
 * Results land in each solver's divF(0), equal to its own compute_residual(0) to rounding. */
typedef struct zefr_batch_s zefr_batch_t;
zefr_batch_t *zefr_batch_create(zefr_t **solvers, int n);
int zefr_batch_residual(zefr_batch_t *b);
int zefr_batch_synchronize(zefr_batch_t *b);
int zefr_batch_launches(zefr_batch_t *b);
void zefr_batch_destroy(zefr_batch_t *b);

/* ---------------------------------------------------------------- singleton ------ */

/* zefr::initialize (include/zefr_interface.hpp:94, src/zefr_interface.cpp:27-33) */
int zefr_initialize(const char *input_file);
/* zefr::get_zefr_object / set_zefr_object / finalize (zefr_interface.hpp:97-102) */
zefr_t *zefr_get_zefr_object(void);
void zefr_set_zefr_object(zefr_t *z);
int zefr_finalize(void);
/* zefr::use_gpus (zefr_interface.hpp:104-111) */
int zefr_use_gpus(void);
/* zefr::get_q_spts(_d) / get_dq_spts(_d) / get_q_spt / get_grad_spt
 * (zefr_interface.hpp:122-129) on the singleton */
double *zefr_get_q_spts(int *ele_stride, int *spt_stride, int *var_stride, int etype);
double *zefr_get_dq_spts(int *ele_stride, int *spt_stride, int *var_stride, int *dim_stride, int etype);
double *zefr_get_q_spts_d(int *ele_stride, int *spt_stride, int *var_stride, int etype);
double *zefr_get_dq_spts_d(int *ele_stride, int *spt_stride, int *var_stride, int *dim_stride, int etype);
double zefr_get_q_spt(int ele, int spt, int var);

/* ---------------------------------------------------------------- kernels -------- */

/* One element-type block on the device.  All pointers are device pointers. */
typedef struct zfr_ele_desc {
  int nDims, order, nVars, nEles, nSpts, nFpts;
  int viscous;
  int eles_per_block;              /* workgroup tile (elements per workgroup), 0 = default */
  double gamma, prandtl, mu, rt, c_sth;
  int fix_vis;
  double rus_k, ldg_b, ldg_tau;
  const double *ops1d;             /* 1D tables: lE[2][N], Dss[N][N], Dsf[N][2]          */
  const double *metric;            /* [E][D*D+1][nS]: adjugate inv_jaco(d1,d2), detJ     */
  const double *fgeo;              /* [E][D+1][nF]: unit normal of the gfpt, signed dA   */
  const unsigned int *nbr;         /* [E][nF]: partner (ele << (B+1) | slot << B | fpt), B = 8 if nF <= 256, else 9 */
  double *Utr;                     /* [E][V][nF] solution trace (scratch)                */
  double *Gtr;                     /* [E][D][V][nF] gradient trace (scratch, viscous)    */
  double *Ftr;                     /* [E][V][nF] common normal flux at the flux points
                                      whose side owns it (scratch, LDG beta = +-1/2)     */
  int fconv_type;                  /* convective common flux: 0 Rusanov, 2 Roe           */
} zfr_ele_desc_t;

/* Elements::extrapolate_U (src/elements.cpp:1337-1360): U[E][V][nS] -> Utr trace. */
int zfr_extrapolate(const zfr_ele_desc_t *d, const double *U, void *stream);
/* The same trace written only at the flux points of faces the partner element owns,
 * which is all the one-sided LDG path (ldg_b = +-1/2) reads: there the owner takes the
 * other side's trace as the common solution and for its common flux, and no element reads
 * its own trace.  Half the writes of zfr_extrapolate; the other points are left as they
 * were.  No reference counterpart (the reference always writes U_fpts in full). */
int zfr_extrapolate_partner(const zfr_ele_desc_t *d, const double *U, void *stream);
/* Faces::compute_common_U + Elements::compute_dU_spts/_fpts + the gradient part of
 * Elements::compute_F + Elements::extrapolate_dU (src/faces.cpp:1366-1474,
 * src/elements.cpp:1362-1447, 1588-1699): writes the gradient trace Gtr and, when
 * dU_out is not NULL, the physical gradient dU_out[E][D][V][nS]. */
int zfr_gradient(const zfr_ele_desc_t *d, const double *U, double *dU_out, void *stream);
/* Elements::compute_F + compute_divF_spts + Faces::compute_common_F + compute_divF_fpts
 * (src/elements.cpp:1449-1501, 1588-1699; src/faces.cpp:1244-1364, 1476-1676):
 * divF[E][V][nS] from U and the traces.  Euler: from U alone (the partner traces are
 * extrapolated from the partner's U inside, replacing extrapolate_U; d->Utr is not read);
 * viscous: reads the solution trace Utr and the gradient trace Gtr (zfr_extrapolate,
 * zfr_gradient).  zfr_divergence_rk reads Utr in both cases when an RK epilogue writes
 * the U it reads (rk->U, U_ini_out or corr_out equal to U); otherwise Euler extrapolates
 * from U as here. */
int zfr_divergence(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream);
/* Navier-Stokes with LDG ldg_b = +-1/2, where one side's gradient enters each common
 * flux: corrected gradient + compute_F + compute_divF_spts (volume part of divF) and the
 * complete common normal flux (Rusanov + LDG, src/faces.cpp:1244-1614) at the flux
 * points whose gradient this element owns, written unscaled to Ftr.  Replaces
 * zfr_gradient + the volume half of zfr_divergence on this path (src/solver.cpp:1552-1610),
 * and extrapolate_U as well: the partner traces it needs (the other side's state at its
 * owned points) are extrapolated here from the partner's U, so no trace launch precedes
 * it.  d->Utr is scratch: the kernel parks those traces at its own owned points. */
int zfr_gradient_flux(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream);
/* The same, reading the partner traces from d->Utr at the partner's points instead (as
 * written by zfr_extrapolate_partner, or by the previous RK stage's fused update with
 * trace_out = d->Utr); used for RK stages 1.. whose trace is already on the device. */
int zfr_gradient_flux_traced(const zfr_ele_desc_t *d, const double *U, double *divF, void *stream);
/* Elements::compute_divF_fpts (src/elements.cpp:1476-1501) on that path:
 * divF += oppDiv_fpts Fcomm with Fcomm = owner's Ftr * signed dA of this side. */
int zfr_face_correction(const zfr_ele_desc_t *d, double *divF, void *stream);
/* RK stage update (src/solver.cpp:2897-2941): U = U_ini - (alpha*dt/detJ) * divF */
int zfr_rk_stage(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF,
                 double alpha, double dt, void *stream);
/* Final RK combination (src/solver.cpp:2997-3062): U = U_ini - sum_s (beta_s*dt/detJ)*divF_s
 * in stage order; divF holds nStages consecutive [E][V][nS] blocks. */
int zfr_rk_final(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF,
                 const double *beta, int nStages, double dt, void *stream);
/* The same two updates with a p-multigrid source term, [E][V][nS] (RK_update_source,
 * src/solver_kernels.cu:253-329, with the CPU path's arithmetic, src/solver.cpp:2946-2962,
 * 3061-3073): every divF_s is replaced by divF_s + source.  source = NULL is the plain update. */
int zfr_rk_stage_source(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF,
                        const double *source, double alpha, double dt, void *stream);
int zfr_rk_final_source(const zfr_ele_desc_t *d, double *U, const double *U_ini, const double *divF,
                        const double *source, const double *beta, int nStages, double dt, void *stream);
/* Residual norm of divF/detJ (src/solver.cpp:7005-7026): res[V] = max | sum abs |
 * sum squares (res_type 0/1/2), reduced on the device; work holds >= 1024*V doubles. */
int zfr_residual_norm(const zfr_ele_desc_t *d, const double *divF, int res_type, double *work,
                      double *res_out_device, void *stream);
/* zfr_residual_norm without the work argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  The block partials go to a fixed array in each
 * device's global memory (1024 x 8 doubles, 64 KiB, allocated with the library's code
 * object and held until the process ends); this call allocates nothing.  Rejects nVars
 * above 8.  Every call on a device shares that array: calls that overlap on one device
 * (other host threads or streams) race, so serialise them. */
int zfr_residual_norm_static(const zfr_ele_desc_t *d, const double *divF, int res_type,
                             double *res_out_device, void *stream);
/* RK update fused into the last kernel of a residual (the stage update of
 * src/solver.cpp:2897-2941 or the final combination of :2997-3062, same arithmetic as
 * zfr_rk_stage / zfr_rk_final): after divF of `stage` is formed, mode 1 writes
 * U = U_ini - alpha dt/detJ divF, mode 2 writes U = U_ini - sum_s beta_s dt/detJ divF_s
 * with the other stages read from divF0 ([nStages][E][V][nS], stage 0 first).
 * trace_out (mode 1 only; NULL for none): the same kernel also writes the new U's trace
 * there ([E][V][nF]), as the next stage's extrapolation would (zfr_extrapolate_partner for
 * zfr_face_correction_rk, read by zfr_gradient_flux_traced; zfr_extrapolate for
 * zfr_divergence_rk).  zfr_face_correction_rk reads no trace and may write d->Utr itself;
 * a zfr_divergence_rk that reads the partners' d->Utr while it writes needs its trace_out
 * in another buffer (the next stage then runs with Utr = trace_out).
 * rk->U need not be the U the residual is formed from: the update writes U from U_ini,
 * so a step may form stage 0's residual from its starting state, write the stage states
 * to a second buffer, and have the final combination write the starting state's buffer
 * (returning the starting state into the second one with U_ini_out). */
typedef struct zfr_rk_args {
  double *U;
  const double *U_ini;
  const double *divF0;
  int mode, stage, nStages;
  double alpha, dt;
  double beta[4];
  const double *source;  /* p-multigrid source term added to each divF_s (modes 1, 2, 4), or NULL */
  double *trace_out;     /* mode 1: also write the new state's trace here, or NULL  */
  /* modes 1, 2: also store a copy of U_ini (the state the update starts from, read before
   * U is written) here, or NULL.  U_ini and U may then be different buffers at every stage:
   * a step may keep its starting state in U and the stage states elsewhere. */
  double *U_ini_out;
  /* mode 2: also write corr_out = (new U) - corr_base, or NULL (p-multigrid's coarse
   * correction, src/multigrid.cpp:236-256) */
  double *corr_out;
  const double *corr_base;
  /* Residual epilogues without an update, for p-multigrid (PMGrid::compute_source_term
   * and the restricted residual, src/multigrid.cpp:105-261, 474-518): mode 3 writes
   * source_out -= divF (source_out holding the restricted residual); mode 4 writes
   * divF += source.  U and U_ini are not read or written in these modes.  Mode 1 with
   * source_out (and source): the stage first forms the source term, source_out = source -
   * divF, then updates with it (stage 0 of a coarse level's first smoothing step, whose
   * residual is the one compute_source_term forms; source_out may be source). */
  double *source_out;
  /* p-multigrid transfers fused into the same kernel (PMGrid::restrict_pmg / prolong_err,
   * src/multigrid.cpp:335-433), with zfr_pmg_transfer's arithmetic.  Restriction (modes 0
   * and 4; mode 0 only with it): the element's U and its divF (after the source term)
   * restricted with restrict_A [restrict_n][N] (N = d->order + 1, 1 <= restrict_n <= N)
   * into restrict_U and restrict_divF ([E][V][restrict_n^D]); zfr_face_correction_rk
   * reads U from restrict_src, zfr_divergence_rk restricts the U it is given.
   * Prolongation (mode 2 with corr_out): prolong_out ([E][V][prolong_n^D], prolong_n >= N)
   * += prolong_alpha * (prolong_A [prolong_n][N] applied to corr_out). */
  const double *restrict_A, *restrict_src;
  double *restrict_U, *restrict_divF;
  int restrict_n, prolong_n;
  const double *prolong_A;
  double prolong_alpha;
  double *prolong_out;
} zfr_rk_args_t;
int zfr_face_correction_rk(const zfr_ele_desc_t *d, double *divF, const zfr_rk_args_t *rk, void *stream);
int zfr_divergence_rk(const zfr_ele_desc_t *d, const double *U, double *divF, const zfr_rk_args_t *rk,
                      void *stream);
/* Vector-Jacobian product of one residual (new; no reference counterpart): for the state
 * U [E][V][nS] and a cotangent w of divF (same layout) writes v = (d divF / dU)^T w.
 * Euler and Navier-Stokes with any LDG beta.  d->Utr, Wtr and Xtr are [E][V][nF]
 * scratch buffers (the partner traces are extrapolated from U inside).  With beta other
 * than +-1/2 both sides' gradients enter every common flux: the VJP then also takes
 * d->Gtr ([E][D][V][nF]) as scratch and runs six launches (forward traces and gradient
 * traces, face weights, the two-sided face reverse, the element reverse, the gather). */
int zfr_residual_vjp(const zfr_ele_desc_t *d, const double *U, const double *w, double *v, double *Wtr,
                     double *Xtr, void *stream);
/* Multi-case residual plan (new; no reference counterpart): the residuals of n independent
 * element blocks (descs[i] with its own arrays, state U[i] and output divF[i], each
 * [E_i][V][nS_i]) in one launch per kernel, shape and flux branch instead of one launch
 * per case and kernel.  Cases are grouped by (kernel, nDims, order, LDG branch, convective
 * flux); each group's workgroups find their case by a binary search over the cases' first
 * workgroups.  Every case takes the kernels its own descriptor selects (Euler: divergence;
 * LDG beta = +-1/2: gradient_flux -> face_correction; other beta: extrapolate -> gradient
 * -> divergence), with the same arithmetic as the
 * single-case entry points (separately compiled, so they agree to rounding).  The descriptors'
 * tiles are ignored: each group takes one tile for its total element count.
 * zfr_multi_create copies the case tables to the device (hipMalloc, blocking copies);
 * zfr_multi_residual only launches, so it may be captured into a HIP graph;
 * zfr_multi_launches returns the number of launches of one residual. */
typedef struct zfr_multi_s zfr_multi_t;
int zfr_multi_create(const zfr_ele_desc_t *descs, const double *const *U, double *const *divF, int n,
                     zfr_multi_t **plan);
int zfr_multi_residual(zfr_multi_t *plan, void *stream);
int zfr_multi_launches(const zfr_multi_t *plan);
void zfr_multi_destroy(zfr_multi_t *plan);
/* Persistent plans (new; no reference counterpart; they replace the launch sequence of
 * PMGrid::v_cycle, src/multigrid.cpp:105-261, for inviscid runs): a recorded sequence of
 * Euler divergence launches (residuals, fused RK stages, source terms and the fused
 * p-multigrid transfers) replayed by ONE cooperative launch.  Its workgroups, all
 * resident, walk each recorded launch's element blocks and meet at a grid-wide barrier
 * before the next one, so results equal the recorded launches' bit for bit.
 * zfr_persist_record_begin starts recording on the calling thread: zfr_divergence and
 * zfr_divergence_rk with inviscid descriptors (order 0..5, one nDims) are then appended
 * instead of launched, and every other launching entry point fails, until
 * zfr_persist_record_end, which builds the plan (hipMalloc and a blocking copy of the
 * phase table; plan NULL discards the recording).  The plan keeps the recorded pointers
 * and scalars: replay it only while those buffers live and hold the same meaning.
 * zfr_persist_info returns the number of recorded launches and the workgroups of the
 * cooperative launch (by default one per CU, at most the largest recorded launch's
 * workgroups); zfr_persist_set_workgroups sets it, up to the workgroups that can be
 * resident at once. */
typedef struct zfr_persist_s zfr_persist_t;
int zfr_persist_record_begin(void);
int zfr_persist_record_end(zfr_persist_t **plan);
int zfr_persist_launch(const zfr_persist_t *plan, void *stream);
int zfr_persist_info(const zfr_persist_t *plan, int *phases, int *workgroups);
int zfr_persist_set_workgroups(zfr_persist_t *plan, int n);
void zfr_persist_destroy(zfr_persist_t *plan);
/* Plain device copy dst[i] = src[i], i < n (bandwidth calibration for the benchmark; no
 * counterpart in the reference). */
int zfr_stream_copy(double *dst, const double *src, long n, void *stream);
/* p-multigrid transfer between orders: the operator products of restrict_pmg, prolong_err
 * and prolong_U (src/multigrid.cpp:335-472), which the reference runs as cublasDGEMM with the
 * dense [nSpts_out][nSpts_in] operator oppRes / oppPro (src/quads.cpp:363-465,
 * src/hexas.cpp:586-704).  That operator is the tensor product of a 1D table, so this takes
 * the 1D table A [n_out][n_in] (row-major, device memory; n = order + 1 points per
 * direction, 1..10) and computes, per element and variable,
 * out = alpha * (A x A [x A]) in + beta * out, in [nEles][nVars][n_in^nDims] ->
 * [nEles][nVars][n_out^nDims].  beta = 0 does not read out. */
int zfr_pmg_transfer(int nDims, int nVars, long nEles, int n_in, int n_out, const double *A, double alpha,
                     const double *in, double beta, double *out, void *stream);
/* z[i] = a*x[i] + b*y[i], i < n (device_copy / device_add / device_subtract,
 * src/aux_kernels.cu:143-205); z may alias x or y. */
int zfr_axpby(long n, double a, const double *x, double b, const double *y, double *z, void *stream);

/* ---------------------------------------------------------------- host physics --- */

/* Physical parameters of the pointwise fluxes (InputStruct fields of the same names). */
typedef struct zfr_phys {
  double gamma, prandtl, mu, rt, c_sth;
  int fix_vis;
  double rus_k, ldg_b, ldg_tau;
  int fconv_type;                  /* 0 Rusanov, 2 Roe (Harten entropy fix, 0.1 a)       */
} zfr_phys_t;

/* Host builds of the pointwise physics the kernels run, for checking on the CPU.
 * zfr_host_point_flux: F[V][D] = Fconv(U) - Fvisc(U, dU) (include/flux.hpp:15-202;
 * dU[V][D] may be NULL for the inviscid flux).  zfr_host_common_flux: the common normal
 * flux of one side, Rusanov + LDG (src/faces.cpp:1244-1614), with each side's viscous
 * normal flux formed from its own gradient dUL / dUR [V][D] (the LDG beta in `a` decides
 * which are used).  The *_vjp variants return the reverse-mode derivatives (new; no
 * reference counterpart): for the cotangent Fb of F they write Ub = (dF/dU)^T Fb and
 * dUb = (dF/d dU)^T Fb. */
int zfr_host_point_flux(int nDims, const zfr_phys_t *a, const double *U, const double *dU, double *F);
int zfr_host_point_flux_vjp(int nDims, const zfr_phys_t *a, const double *U, const double *dU, const double *Fb,
                            double *Ub, double *dUb);
int zfr_host_common_flux(int nDims, const zfr_phys_t *a, int viscous, const double *UL, const double *UR,
                         const double *dUL, const double *dUR, const double *n, double dAs, double *F);
int zfr_host_common_flux_vjp(int nDims, const zfr_phys_t *a, int viscous, const double *UL, const double *UR,
                             const double *dUL, const double *dUR, const double *n, double dAs, const double *Fb,
                             double *ULb, double *URb, double *dULb, double *dURb);

#ifdef __cplusplus
}
#endif

#endif /* ZEFR_AMD_H */
