#!/usr/bin/env python
"""Residual throughput of the MI355X DFR path on the 3D Navier-Stokes Taylor-Green case.

Workload (one "step" = one FRSolver::compute_residual over the whole mesh, inputs
resident in HBM): periodic 64^3 uniform hex mesh on [0, 2pi]^3, p = 3 (64 solution
points per element, 16.8M points x 5 variables = 83.9M DoF), Re = 1600, Ma = 0.1,
Pr = 0.71, Roe + LDG (beta = 0.5, tau = 0.1), FP64, Taylor-Green initial state.

Prints one JSON line with the throughput in Gdof-residuals/s, the roofline of the
dominant kernel (live HIP-event timing on the solver's stream), the whole-residual
byte model, a stream-copy bandwidth measured on the same device, the residual's VJP
ratio on the 32^3 adjoint case, and a CPU baseline: the CPU oracle (oracle/, OpenMP, in
a child process) on the same 64^3 mesh and state.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--n 64] [--order 3] [--flux Roe]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gdof-residuals/s FP64 hex p=3, % HBM roofline; VJP/residual time ratio"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=64, help="elements per direction")
    p.add_argument("--order", type=int, default=3)
    p.add_argument("--flux", default="Roe", choices=["Roe", "Rusanov"],
                   help="convective common flux of the headline (the other one is timed as a leg)")
    p.add_argument("--tile", type=int, default=-1, help="elements per workgroup (-1: autotune)")
    p.add_argument("--tiles", default=None,
                   help="per-launch tiles 'trace,flux,correction' (overrides --tile; default: autotune)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-vjp", action="store_true", help="skip the VJP timing")
    p.add_argument("--no-steps", action="store_true", help="skip the RK4 step timings")
    p.add_argument("--no-pmg", action="store_true", help="skip the p-multigrid timing")
    p.add_argument("--no-other-flux", action="store_true", help="skip the leg with the other convective flux")
    p.add_argument("--no-c5", action="store_true", help="skip the 32^3 adjoint-case VJP ratio")
    p.add_argument("--no-batch", action="store_true", help="skip the batched small-case timing")
    p.add_argument("--no-c2", action="store_true", help="skip the 2D Navier-Stokes 256^2 p=4 timing")
    p.add_argument("--cpu-warmup", type=int, default=2)
    p.add_argument("--cpu-steps", type=int, default=10, help="timed CPU residuals (median reported)")
    p.add_argument("--workdir", default=None)
    p.add_argument("--cpu-worker", nargs=4, metavar=("MESH", "ORDER", "FLUX", "OUT"), default=None,
                   help=argparse.SUPPRESS)  # internal: the CPU baseline's child process
    return p.parse_args()


def write_deck(path, meshfile, order, dt=2e-4, fconv="Rusanov"):
    with open(path, "w") as f:
        f.write(f"""nDims 3
meshfile {meshfile}
mesh_bound periodic periodic
mesh_bound fluid fluid
order {order}
equation EulerNS
viscous 1
n_steps 200
dt_scheme RK44
dt {dt!r}
dt_type 0
write_freq 0
report_freq 10
res_type 2
ic_type 1
test_case 4
fconv_type {fconv}
fvisc_type LDG
rus_k 0.0
ldg_b 0.5
ldg_tau 0.1
gamma 1.4
prandtl 0.71
disable_nondim 1
fix_vis 1
mach_fs 0.1
Re_fs 1600
P_fs 1.0
rho_fs 1.0
L_fs 1.0
output_prefix tgv
""")
    return path


def kernel_bytes(D, N):
    """Algorithmic bytes per element of the two launches of one viscous residual with
    LDG beta = +-1/2 (every input read once, every output written once; FP64 values,
    uint32 partner map).  Each face has exactly one owner side, so an element owns nF/2
    flux points on average.
      gradient_flux : U (nS V), the partner's solution line through each owned fpt
                      (nF/2 N V; the partner trace is extrapolated in the kernel),
                      adjugate + detJ (nS (D^2+1)), owned normals (nF/2 D), nbr (nF)
                      -> Ftr at owned fpts (nF/2 V), divF (nS V)
      face_correction: divF (nS V) in/out, owner Ftr (nF V), signed dA (nF), nbr (nF)
    The partner traces the gradient_flux kernel parks in Utr for its own face phase
    (nF/2 V written and read back) are scratch and not counted."""
    V = D + 2
    nS = N ** D
    nF = 2 * D * N ** (D - 1)
    h = nF // 2
    gflux = 8 * (nS * V + h * N * V + nS * (D * D + 1) + h * D + h * V + nS * V) + 4 * nF
    fcorr = 8 * (2 * nS * V + nF * V + nF) + 4 * nF
    return {"gradient_flux": gflux, "face_correction": fcorr}


def kernel_bytes_read_once(D, N):
    """The same launches counted with every value read once (the scope table's rule): the
    partner's solution enters gradient_flux as one trace value per owned flux point and
    variable (nF/2 V), as if a trace launch had formed it, instead of its N-point line."""
    V = D + 2
    nF = 2 * D * N ** (D - 1)
    b = kernel_bytes(D, N)
    b["gradient_flux"] -= 8 * (nF // 2) * (N - 1) * V
    return b


FLUX_CODE = {"Rusanov": 0, "Roe": 1}  # the kernels' convective-flux template parameter (FC)
KERNEL_NAMES = {"gradient_flux": "k_visc_onesided<3,{N},1,{FC}>", "face_correction": "k_face_correction<3,{N},1>"}


def residual_model_bytes(D, N, viscous):
    """Per-element algorithmic bytes of one residual (the benchmark's byte model):
    8 [2 nS V + nS m + 2 nF V + (viscous ? 2 nF V D : 0)] + (nF/2)(D+2) 8 + nF 4."""
    V = D + 2
    nS = N ** D
    nF = 2 * D * N ** (D - 1)
    m = D * D + (1 if viscous else 0)
    return 8 * (2 * nS * V + nS * m + 2 * nF * V + (2 * nF * V * D if viscous else 0)) \
        + (nF // 2) * (D + 2) * 8 + nF * 4


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "residual_kernels_tgv_hex64_p3_hbm_traffic.md")


def profiled_traffic(kernel, n, order):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (copy-kernel calibrated), which covers the default workload (64^3, p=3) only."""
    if (n, order) != (64, 3) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return int(parts[2])
    return None


def stream_copy_gbs(L, nbytes=2 << 30, reps=20):
    """Device copy bandwidth (read + write bytes / time) with the library's copy kernel."""
    import torch
    n = nbytes // 8
    src = torch.empty(n, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    st = torch.cuda.current_stream()
    sp = C.c_void_p(st.cuda_stream)
    for _ in range(3):
        L.zfr_stream_copy(dst.data_ptr(), src.data_ptr(), n, sp)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(st)
    for _ in range(reps):
        L.zfr_stream_copy(dst.data_ptr(), src.data_ptr(), n, sp)
    b.record(st)
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    ok = bool(torch.equal(dst[:1024], src[:1024]))
    del src, dst
    torch.cuda.empty_cache()
    return 2 * n * 8 / (ms * 1e-3) / 1e9, ok


def step_timing(z, steps):
    """Mean wall time of one RK4 step (FRSolver::update) per execution mode: the plain
    launch sequence, the RK update fused into each residual's last kernel, and the fused
    step replayed from a HIP graph.  The state is restored after each mode."""
    U0 = z.state("U", layout="device")
    out = {}
    for name, fuse, graph in (("plain", False, False), ("fused", True, False), ("fused_graph", True, True)):
        z.set_step_mode(fuse, graph)
        z.set_state(U0, layout="device")
        z.do_step()
        z.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            z.do_step()
        z.synchronize()
        out[name + "_ms"] = round((time.perf_counter() - t) / steps * 1e3, 4)
    z.set_step_mode(True, False)
    z.set_state(U0, layout="device")
    return out


def small_case_steps(workdir, steps=200):
    """The CPU-runnable case: 2D Euler isentropic vortex, 32^2 quads on [-5,5]^2, p=3,
    RK44 dt=1e-3; per-step latency per execution mode (launch-bound at this size)."""
    from zefr_amd import zefr
    from zefr_amd.mesh import write_box
    mesh = write_box(os.path.join(workdir, "vortex32.msh"), (32, 32), (-5.0, -5.0), (5.0, 5.0))
    deck = os.path.join(workdir, "vortex.deck")
    with open(deck, "w") as f:
        f.write(f"""nDims 2
meshfile {mesh}
mesh_bound periodic periodic
mesh_bound fluid fluid
order 3
equation EulerNS
viscous 0
n_steps {steps}
dt_scheme RK44
dt 1e-3
dt_type 0
ic_type 1
test_case 1
fconv_type Rusanov
rus_k 0.0
gamma 1.4
disable_nondim 1
rho_fs 1.0
P_fs 1.0
mach_fs 0.1
output_prefix vortex
""")
    z = zefr.Zefr()
    z.read_input(deck)
    z.setup_solver()
    d = z.dims()
    res = step_timing(z, steps)
    res["workload"] = (f"2D Euler isentropic vortex, 32^2 quads on [-5,5]^2, p=3 "
                       f"({d['nEles'] * d['nSpts'] * d['nVars']} DoF), RK44 steps")
    return res


def c2_timing(workdir, steps, n=256, order=4):
    """The 2D Navier-Stokes configuration: 256^2 periodic quads on [0, 2 pi]^2, p = 4, the
    2D Taylor-Green field u = sin x cos y, v = -cos x sin y (Ma 0.1, Re 1600, Pr 0.72),
    Rusanov + LDG (beta 0.5, tau 0.1).  Residual throughput after autotuning, wall time of
    `steps` back-to-back residuals, and the two kernels' average durations (HIP events)."""
    import numpy as np
    from zefr_amd import zefr
    from zefr_amd._lib import lib
    from zefr_amd.mesh import write_box
    mesh = write_box(os.path.join(workdir, f"c2_quad{n}.msh"), (n, n), (0.0, 0.0), (2 * math.pi, 2 * math.pi))
    deck = os.path.join(workdir, "c2.deck")
    with open(deck, "w") as f:
        f.write(f"nDims 2\nmeshfile {mesh}\nmesh_bound periodic periodic\nmesh_bound fluid fluid\n"
                f"order {order}\nequation EulerNS\nviscous 1\nn_steps 1\ndt_scheme RK44\ndt 1e-3\n"
                f"dt_type 0\nic_type 0\nfconv_type Rusanov\nfvisc_type LDG\nrus_k 0.0\nldg_b 0.5\n"
                f"ldg_tau 0.1\ngamma 1.4\nprandtl 0.72\ndisable_nondim 1\nfix_vis 1\nmach_fs 0.1\n"
                f"Re_fs 1600\nP_fs 1.0\nrho_fs 1.0\nL_fs 1.0\nR 286.9\noutput_prefix c2\n")
    z = zefr.Zefr()
    z.read_input(deck)
    z.setup_solver()
    d = z.dims()
    X = z.host_array("coord_spts").reshape(d["nSpts"], 2, d["nEles"])  # reference layout
    x, y = X[:, 0, :], X[:, 1, :]
    g, P0, rho0 = 1.4, 1.0, 1.0
    U0 = 0.1 * math.sqrt(g * P0 / rho0)
    u = U0 * np.sin(x) * np.cos(y)
    v = -U0 * np.cos(x) * np.sin(y)
    p = P0 + rho0 * U0 * U0 / 4.0 * (np.cos(2 * x) + np.cos(2 * y))
    rho = rho0 * p / P0
    z.set_state(np.stack([rho, rho * u, rho * v, p / (g - 1.0) + 0.5 * rho * (u * u + v * v)], axis=1))
    z.autotune(3)
    dof = d["nEles"] * d["nSpts"] * d["nVars"]
    for _ in range(3):
        z.compute_residual(0)
    z.synchronize()
    ms3 = (C.c_double * 3)()
    t = time.perf_counter()
    rc = lib().zefr_time_residual_kernels(z._h, steps, ms3)
    z.synchronize()
    wall = time.perf_counter() - t
    if rc != 0:
        raise RuntimeError(lib().zefr_last_error().decode())
    tiles = z.tiles()
    del z
    return {"workload": f"2D Navier-Stokes Taylor-Green residual, {n}^2 quads on [0,2pi]^2, p={order}, "
                        "Rusanov+LDG(beta=0.5,tau=0.1)", "dof": dof,
            "value": round(dof / (wall / steps) / 1e9, 4), "unit": "Gdof-residuals/s",
            "ms_per_step": round(wall / steps * 1e3, 4), "tiles_eles_per_block": tiles,
            "kernels_ms": {"gradient_flux": round(ms3[1], 4), "face_correction": round(ms3[2], 4)}}


def pmg_timing(workdir, cycles=10):
    """p-multigrid on the quad half of the steady case: 128^2 periodic quads on [0,1]^2, p=4,
    Euler, levels 4-3-2-1-0 with one RK44 smoothing step per level and V-cycles (the
    reference's mixed tri/quad variant is outside this build).  Mean wall time of one
    V-cycle against one fine RK44 step, and the restriction kernel's bandwidth (HIP events,
    fine U -> order 3, bytes = elements x vars x (25 + 16) x 8)."""
    import numpy as np
    import torch
    from zefr_amd import zefr
    from zefr_amd._lib import lib
    from zefr_amd.mesh import write_box
    mesh = write_box(os.path.join(workdir, "quad128.msh"), (128, 128), (0.0, 0.0), (1.0, 1.0))
    deck = os.path.join(workdir, "pmg.deck")
    with open(deck, "w") as f:
        f.write(f"""nDims 2
meshfile {mesh}
mesh_bound periodic periodic
mesh_bound fluid fluid
order 4
equation EulerNS
viscous 0
n_steps 10
dt_scheme RK44
dt 1e-4
dt_type 0
ic_type 0
fconv_type Rusanov
rus_k 0.0
gamma 1.4
disable_nondim 1
rho_fs 1.0
P_fs 1.0
mach_fs 0.5
p_multi 1
mg_levels 4 3 2 1 0
mg_steps 1 1 1 1 1
mg_cycle V
output_prefix pmg
""")
    z = zefr.Zefr()
    z.read_input(deck)
    z.setup_solver()
    d = z.dims()
    E, V, nS = d["nEles"], d["nVars"], d["nSpts"]
    rng = np.random.default_rng(0)
    U0 = z.state("U", layout="device")
    U0 = U0 * (1.0 + 0.05 * rng.standard_normal(U0.shape))
    z.set_state(U0, layout="device")
    z.do_step()
    z.synchronize()
    t = time.perf_counter()
    for _ in range(cycles):
        z.do_step()
    z.synchronize()
    cyc_ms = (time.perf_counter() - t) / cycles * 1e3
    # the same cycle replayed from one captured HIP graph
    z.set_step_mode(True, True)
    z.do_step()
    z.synchronize()
    t = time.perf_counter()
    for _ in range(cycles):
        z.do_step()
    z.synchronize()
    cyc_graph_ms = (time.perf_counter() - t) / cycles * 1e3
    # the same cycle as one persistent (cooperative) launch
    z.set_step_mode(True, False)
    z.set_pmg_persistent(True)
    z.do_step()
    z.synchronize()
    t = time.perf_counter()
    for _ in range(cycles):
        z.do_step()
    z.synchronize()
    cyc_persist_ms = (time.perf_counter() - t) / cycles * 1e3
    z.set_pmg_persistent(False)
    # one fine RK44 step on the same mesh (a solver without multigrid)
    z2 = zefr.Zefr()
    z2.read_input(deck)
    z2.get_input().p_multi = 0
    z2.setup_solver()
    z2.set_state(U0, layout="device")
    z2.do_step()
    z2.synchronize()
    t = time.perf_counter()
    for _ in range(cycles):
        z2.do_step()
    z2.synchronize()
    step_ms = (time.perf_counter() - t) / cycles * 1e3
    # restriction kernel 4 -> 3 on the fine state
    from zefr_amd.zefr import check
    A1 = torch.from_numpy(_restrict_table(4, 3)).cuda()
    x = torch.from_numpy(U0).cuda()
    y = torch.empty((E, V, 16), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    reps = 50
    for _ in range(3):
        check(lib().zfr_pmg_transfer(2, V, E, 5, 4, C.c_void_p(A1.data_ptr()), 1.0, C.c_void_p(x.data_ptr()), 0.0,
                                     C.c_void_p(y.data_ptr()), C.c_void_p(st.cuda_stream)))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        lib().zfr_pmg_transfer(2, V, E, 5, 4, C.c_void_p(A1.data_ptr()), 1.0, C.c_void_p(x.data_ptr()), 0.0,
                               C.c_void_p(y.data_ptr()), C.c_void_p(st.cuda_stream))
    e1.record(st)
    e1.synchronize()
    kms = e0.elapsed_time(e1) / reps
    kbytes = E * V * (nS + 16) * 8
    del z, z2
    return {"workload": "p-multigrid V-cycle, 128^2 periodic quads on [0,1]^2, p=4, Euler, levels 4-3-2-1-0, "
                        "one RK44 step per level",
            "vcycle_ms": round(cyc_ms, 4), "vcycle_graph_ms": round(cyc_graph_ms, 4),
            "vcycle_persistent_ms": round(cyc_persist_ms, 4),
            "fine_rk44_step_ms": round(step_ms, 4),
            "vcycle_over_fine_step": round(cyc_ms / step_ms, 3),
            "restrict_kernel": {"name": "k_pmg_transfer<2>", "avg_ms": round(kms, 5), "bytes": kbytes,
                                "achieved_GBs": round(kbytes / (kms * 1e-3) / 1e9, 1)}}


def _restrict_table(p, r):
    """1D restriction table [r+1][p+1] between Gauss-Legendre point sets (order p -> r),
    chained through the intermediate orders as the solver builds it."""
    import numpy as np

    def step(nf, nt):
        xf, _ = np.polynomial.legendre.leggauss(nf)
        xt, _ = np.polynomial.legendre.leggauss(nt)
        A = np.ones((nt, nf))
        for j in range(nf):
            for m in range(nf):
                if m != j:
                    A[:, j] *= (xt - xf[m]) / (xf[j] - xf[m])
        return A
    R = step(r + 2, r + 1)
    for P in range(r + 1, p):
        R = R @ step(P + 2, P + 1)
    return np.ascontiguousarray(R)


def batch_timing(workdir, count=64, reps=20, seed=2024):
    """The batched half of the multigrid configuration: 64 independent periodic quad cases,
    seeded sizes n x n (n in 8..64) and orders 1..5, Euler, seeded perturbed freestream
    (5% Gaussian point noise).  One residual of every case per batch: eager launches on
    the solvers' own streams against replays of the batch graph (the members' residuals as a
    multi-case plan: one launch per order; an Euler residual is one kernel).  Each leg runs `reps` batches back to
    back and synchronises once."""
    import numpy as np
    from zefr_amd import zefr
    from zefr_amd.mesh import write_box
    rng = np.random.default_rng(seed)
    zs, dof = [], 0
    for i in range(count):
        n, p = int(rng.integers(8, 65)), int(rng.integers(1, 6))
        mesh = write_box(os.path.join(workdir, f"batch{i}.msh"), (n, n), (0.0, 0.0), (1.0, 1.0))
        deck = os.path.join(workdir, f"batch{i}.deck")
        with open(deck, "w") as f:
            f.write(f"nDims 2\nmeshfile {mesh}\nmesh_bound periodic periodic\nmesh_bound fluid fluid\n"
                    f"order {p}\nequation EulerNS\nviscous 0\nn_steps 1\ndt_scheme RK44\ndt 1e-4\n"
                    f"dt_type 0\nic_type 0\nfconv_type Rusanov\ngamma 1.4\ndisable_nondim 1\nrho_fs 1.0\n"
                    f"P_fs 1.0\nmach_fs 0.5\noutput_prefix batch\n")
        z = zefr.Zefr()
        z.read_input(deck)
        z.setup_solver()
        U = z.state("U", layout="device")
        z.set_state(U * (1.0 + 0.05 * rng.standard_normal(U.shape)), layout="device")
        d = z.dims()
        dof += d["nEles"] * d["nSpts"] * d["nVars"]
        zs.append(z)
    b = zefr.Batch(zs)
    from zefr_amd._lib import check, lib

    def eager(n):
        for _ in range(n):
            for z in zs:
                z.compute_residual(0)
        for z in zs:
            z.synchronize()

    def graph(n):
        for _ in range(n):
            check(lib().zefr_batch_residual(b._h))
        b.synchronize()

    out = {}
    for name, fn in (("eager", eager), ("graph", graph)):
        fn(2)
        t = time.perf_counter()
        fn(reps)
        out[name + "_ms"] = round((time.perf_counter() - t) / reps * 1e3, 4)
    launches = b.launches
    del b, zs
    return {"workload": f"{count} independent periodic quad cases (n in 8..64, p in 1..5, seed {seed}), Euler, "
                        "one residual each", "dof": dof, "eager_ms": out["eager_ms"], "eager_launches": count,
            "graph_ms": out["graph_ms"], "graph_launches": launches,
            "graph_Gdof_res_per_s": round(dof / (out["graph_ms"] * 1e-3) / 1e9, 3)}


def residual_timing(z, steps):
    """Wall time of `steps` back-to-back residuals on the solver's stream (bracketed by
    device synchronisation) and the average duration of each of its kernels (HIP events
    recorded on that stream)."""
    from zefr_amd._lib import lib
    import torch
    z.synchronize()
    torch.cuda.synchronize()
    ms3 = (C.c_double * 3)()
    t = time.perf_counter()
    rc = lib().zefr_time_residual_kernels(z._h, steps, ms3)
    z.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    if rc != 0:
        raise RuntimeError(lib().zefr_last_error().decode())
    return wall / steps * 1e3, {"gradient_flux": ms3[1], "face_correction": ms3[2]}


def tgv_solver(workdir, mesh, order, flux, tag):
    from zefr_amd import zefr
    deck = write_deck(os.path.join(workdir, f"tgv_{tag}.deck"), mesh, order, fconv=flux)
    z = zefr.Zefr()
    z.read_input(deck)
    z.setup_solver()
    return z


def flux_leg(workdir, mesh, order, steps, flux):
    """The same Taylor-Green workload with the other convective flux: residual
    throughput, its kernels' average durations and the VJP against the forward."""
    z = tgv_solver(workdir, mesh, order, flux, flux.lower())
    z.autotune(3)
    d = z.dims()
    dof = d["nEles"] * d["nSpts"] * d["nVars"]
    for _ in range(3):
        z.compute_residual(0)
    ms, kms = residual_timing(z, steps)
    vjp = vjp_timing(z, steps)
    tiles = z.tiles()
    del z
    return {"workload": f"3D Navier-Stokes Taylor-Green residual with {flux}+LDG(beta=0.5,tau=0.1), same mesh",
            "value": round(dof / (ms * 1e-3) / 1e9, 4), "unit": "Gdof-residuals/s", "ms_per_step": round(ms, 4),
            "tiles_eles_per_block": tiles,
            "kernels_ms": {k: round(v, 4) for k, v in kms.items()}, "vjp": vjp}


def vjp_timing(z, reps):
    """The residual's vector-Jacobian product on the solver's mesh and state against the
    forward residual (residual_at), both on the torch stream, each call bracketed by HIP
    events: medians over `reps` calls (the VJP ratio is a median over a median)."""
    import statistics
    import torch
    U = z.state_tensor()
    w = torch.randn_like(U)
    v = torch.empty_like(U)
    R = torch.empty_like(U)
    st = torch.cuda.current_stream()
    for _ in range(2):
        z.residual_at(U, R)
        z.residual_vjp(U, w, v)
    torch.cuda.synchronize()
    ef = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ea = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ef[0].record(st)
    for i in range(reps):
        z.residual_at(U, R)
        ef[i + 1].record(st)
    ea[0].record(st)
    for i in range(reps):
        z.residual_vjp(U, w, v)
        ea[i + 1].record(st)
    ea[-1].synchronize()
    fwd = statistics.median(ef[i].elapsed_time(ef[i + 1]) for i in range(reps))
    adj = statistics.median(ea[i].elapsed_time(ea[i + 1]) for i in range(reps))
    ok = bool(torch.isfinite(v).all())
    del U, w, v, R
    torch.cuda.empty_cache()
    return {"forward_ms": round(fwd, 4), "vjp_ms": round(adj, 4), "vjp_residual_ratio": round(adj / fwd, 3),
            "statistic": f"median of {reps}", "finite": ok}


def c5_vjp(workdir, order, reps, flux):
    """The adjoint configuration's VJP/residual ratio: 32^3 hexes p=3 Taylor-Green."""
    from zefr_amd.mesh import write_box
    mesh = write_box(os.path.join(workdir, "hex32.msh"), (32, 32, 32), (0.0,) * 3, (2 * math.pi,) * 3)
    out = {"workload": f"3D Navier-Stokes Taylor-Green, 32^3 hexes on [0,2pi]^3, p={order}"}
    for fl in (flux, "Rusanov" if flux == "Roe" else "Roe"):
        z = tgv_solver(workdir, mesh, order, fl, f"c5_{fl.lower()}")
        z.autotune(3)
        out[fl] = vjp_timing(z, reps)
        del z
    out["vjp_residual_ratio"] = out[flux]["vjp_residual_ratio"]
    return out


def cpu_topology():
    """Host CPU description: lscpu's model and counts, the CPUs this process may run on and
    the physical cores among them (sysfs topology), and the OpenMP thread count to use:
    one per available physical core, capped by an inherited OMP_NUM_THREADS."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, ValueError):
        pass
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    phys = len(cores)
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(phys, cap) if cap > 0 else phys
    info.update({"cpus_available": len(cpus), "physical_cores_available": phys,
                 "omp_num_threads_inherited": cap or None})
    return threads, info


def cpu_worker(mesh, order, flux, out_path, warm, reps):
    """Child process of the CPU baseline: the CPU oracle on the benchmark's mesh and state,
    `warm` untimed then `reps` timed residuals; writes the per-residual times (s)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle.oracle import Oracle, make_phys
    from common import tgv3d_state
    U0 = 0.1 * math.sqrt(1.4)
    mu = 1.0 * U0 * 1.0 / 1600.0
    t0 = time.perf_counter()
    o = Oracle(mesh, 3, order, make_phys(viscous=True, gamma=1.4, prandtl=0.71, mu=mu, fix_vis=True,
                                         rus_k=0.0, ldg_b=0.5, ldg_tau=0.1, dt=2e-4, fconv=2 if flux == "Roe" else 0))
    x = o.coord_spts()
    o.set_U(tgv3d_state(x[:, 0, :], x[:, 1, :], x[:, 2, :]))
    setup = time.perf_counter() - t0
    for _ in range(warm):
        o.compute_residual(0)
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        o.compute_residual(0)
        ts.append(time.perf_counter() - t)
    with open(out_path, "w") as f:
        json.dump({"times_s": ts, "setup_s": setup, "dof": o.nEles * o.nSpts * o.nVars, "nEles": o.nEles}, f)


def cpu_baseline(workdir, mesh, n, order, flux, warm, reps):
    """The CPU oracle (oracle/zefr_oracle.cpp, OpenMP) on the same mesh, state and flux as
    the headline, run in a child process so that its OpenMP settings take effect:
    OMP_NUM_THREADS = the physical cores available, OMP_PROC_BIND=close, OMP_PLACES=cores;
    median of `reps` residuals after `warm` untimed ones."""
    import statistics
    import subprocess
    threads, topo = cpu_topology()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out_path = os.path.join(workdir, "cpu_baseline.json")
    subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", mesh, str(order), flux, out_path,
                    "--cpu-warmup", str(warm), "--cpu-steps", str(reps)], env=env, check=True)
    with open(out_path) as f:
        r = json.load(f)
    med = statistics.median(r["times_s"])
    return {"value": r["dof"] / med / 1e9, "unit": "Gdof-residuals/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (oracle/zefr_oracle.cpp, OpenMP, {threads} threads, OMP_PROC_BIND=close, "
                      f"OMP_PLACES=cores) on the headline workload itself: {n}^3 hexes p={order}, {flux}+LDG, "
                      f"{r['dof'] / 1e6:.1f}M DoF; median of {reps} residuals after {warm} warm-up "
                      f"({med * 1e3:.0f} ms each, {sum(r['times_s']):.1f} s timed)",
            "ms_per_residual": round(med * 1e3, 2), "times_ms": [round(t * 1e3, 1) for t in r["times_s"]],
            "setup_s": round(r["setup_s"], 1), "host": topo}


def main():
    args = parse()
    if args.cpu_worker:
        mesh, order, flux, out_path = args.cpu_worker
        cpu_worker(mesh, int(order), flux, out_path, args.cpu_warmup, args.cpu_steps)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is a design, see the sharding notes)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    from zefr_amd._lib import lib
    from zefr_amd.mesh import write_box

    workdir = args.workdir or tempfile.mkdtemp(prefix="zefr_bench_")
    n, order, flux = args.n, args.order, args.flux
    other = "Rusanov" if flux == "Roe" else "Roe"
    D, N, FC = 3, order + 1, FLUX_CODE[flux]
    t0 = time.perf_counter()
    mesh = write_box(os.path.join(workdir, f"hex{n}.msh"), (n, n, n), (0.0,) * 3, (2 * math.pi,) * 3)
    z = tgv_solver(workdir, mesh, order, flux, "main")
    setup_s = time.perf_counter() - t0
    d = z.dims()
    E, nS, V = d["nEles"], d["nSpts"], d["nVars"]
    dof = E * nS * V

    if args.tiles:
        z.set_tiles([int(x) for x in args.tiles.split(",")])
    elif args.tile >= 0:
        z.set_tile(args.tile)
    else:
        z.autotune(3)
    tile = z.tiles()

    for _ in range(args.warmup):
        z.compute_residual(0)
    ms_step, kms = residual_timing(z, args.steps)
    value = dof / (ms_step * 1e-3) / 1e9

    kb = kernel_bytes(D, N)
    names = ["gradient_flux", "face_correction"]  # the one-sided residual has no trace launch
    dom = max(names, key=lambda k: kms[k])
    kname = {k: KERNEL_NAMES[k].format(N=N, FC=FC) for k in names}
    achieved = kb[dom] * E / (kms[dom] * 1e-3) / 1e9
    per_kernel = {kname[k]: {"avg_ms": round(kms[k], 4), "alg_bytes_per_element": kb[k],
                             "achieved_GBs": round(kb[k] * E / (kms[k] * 1e-3) / 1e9, 1) if kms[k] > 0 else None}
                  for k in names}
    read_once = kernel_bytes_read_once(D, N)
    model = residual_model_bytes(D, N, True)
    res_gbs = model * E / (ms_step * 1e-3) / 1e9
    copy_gbs, copy_ok = stream_copy_gbs(lib())
    vjp = None if args.no_vjp else vjp_timing(z, args.steps)
    rk = None
    if not args.no_steps:
        rk = {f"tgv_hex{n}_p{order}": step_timing(z, 5), "vortex_quad32_p3": small_case_steps(workdir)}
    del z
    torch.cuda.empty_cache()
    c5 = None if (args.no_c5 or args.no_vjp) else c5_vjp(workdir, order, args.steps, flux)
    c2 = None if args.no_c2 else c2_timing(workdir, args.steps)
    pmg = None if args.no_pmg else pmg_timing(workdir)
    oth = None if args.no_other_flux else flux_leg(workdir, mesh, order, args.steps, other)
    batch = None if args.no_batch else batch_timing(workdir)

    rec = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gdof-residuals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: generated periodic hex mesh, Taylor-Green initial state",
        "config": {"workload": f"3D Navier-Stokes Taylor-Green residual, {n}^3 hexes on [0,2pi]^3, p={order}",
                   "elements": E, "spts_per_element": nS, "vars": V, "dof": dof, "Re": 1600, "Ma": 0.1,
                   "Pr": 0.71, "flux": f"{flux}+LDG(beta=0.5,tau=0.1)", "tiles_eles_per_block": tile,
                   "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(kname[dom], n, order),
                     "kernel": kname[dom], "timing": "live: HIP events on the solver's stream, mean launch",
                     "alg_bytes_per_launch": kb[dom] * E, "avg_launch_ms": round(kms[dom], 4),
                     "read_once": {"alg_bytes_per_launch": read_once[dom] * E,
                                   "achieved": round(read_once[dom] * E / (kms[dom] * 1e-3) / 1e9, 1),
                                   "frac": round(read_once[dom] * E / (kms[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
        "roofline_residual": {"bytes_per_element_model": model, "model_GBs": round(res_gbs, 1),
                              "model_frac_of_peak": round(res_gbs / HBM_PEAK_GBS, 4),
                              "note": "byte model of the reference's eleven-launch data flow, not a bandwidth",
                              "stream_copy_GBs": round(copy_gbs, 1), "stream_copy_ok": copy_ok},
        "kernels": per_kernel,
        "vjp": vjp,
        "vjp_c5": c5,
        "rk4_step": rk,
        "c2_quad256_p4": c2,
        "pmg": pmg,
        other.lower(): oth,
        "batch": batch,
    }
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(workdir, mesh, n, order, flux, args.cpu_warmup, args.cpu_steps)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
