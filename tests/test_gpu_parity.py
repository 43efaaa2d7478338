"""GPU parity: the HIP residual path (through the C ABI) against the CPU oracle on the same
mesh, order and initial state.

Bars (floating point, L-inf taken norm-wise per variable v):
  * one residual: max|divF - divF_ref|_v <= max(1e-12 * max|divF_ref|_v, 64 eps * max|S|_v)
    where S = |oppDiv| |F| + |oppDiv_fpts| |Fcomm| is the magnitude of the terms summed
    into divF.  The first bound is the 1e-12 relative bar; the second only takes over where
    divF is a cancellation of much larger terms (the Taylor-Green initial states are near
    equilibrium: |divF| ~ 1e-4 |S|), where no two FP64 evaluations with different rounding
    can agree to 1e-12 of |divF|.  The assertion message reports the plain relative error
    and the cancellation factor kappa = max|S| / max|divF|.
  * after RK steps (state U, well conditioned): relative <= 1e-10
Integer maps are compared bit-exactly in test_host_setup.py.
"""

import math

import numpy as np
import pytest

from common import TWO_PI, paired, rel_err, residual_parity

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
RK_TOL = 1e-10

CASES = [
    dict(name="q_euler_p3_vortex", n=(8, 8), lo=(-5.0, -5.0), hi=(5.0, 5.0), order=3, viscous=False,
         state="vortex"),
    dict(name="q_ns_p4_tgv", n=(6, 6), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=4, viscous=True,
         state="tgv2d"),
    dict(name="h_ns_p3_tgv", n=(3, 3, 3), lo=(0.0, 0.0, 0.0), hi=(TWO_PI,) * 3, order=3, viscous=True,
         state="tgv3d"),
    dict(name="h_ns_p2_curved_central", n=(3, 3, 2), lo=(0.0, 0.0, 0.0), hi=(1.0, 2.0, 1.5), order=2,
         viscous=True, state="smooth", perturb=0.04, extra={"ldg_b": 0.0, "ldg_tau": 1.0}),
    dict(name="q_ns_p3_curved_sutherland", n=(5, 4), lo=(0.0, 0.0), hi=(2.0, 1.0), order=3, viscous=True,
         state="smooth", perturb=0.05,
         extra={"ldg_b": -0.5, "rus_k": 0.3, "fix_vis": 0, "disable_nondim": 0, "mach_fs": 0.3,
                "Re_fs": 100.0, "rho_fs": 1.2}),
    dict(name="q_euler_p5_ragged", n=(5, 3), lo=(0.0, 0.0), hi=(1.0, 1.0), order=5, viscous=False,
         state="smooth"),
    dict(name="h_euler_p1_ragged", n=(5, 3, 3), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), order=1,
         viscous=False, state="smooth", perturb=0.03),
    dict(name="q_ns_p0", n=(4, 3), lo=(0.0, 0.0), hi=(1.0, 1.0), order=0, viscous=True, state="smooth"),
    dict(name="h_ns_p4_curved", n=(2, 2, 3), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), order=4, viscous=True,
         state="smooth", perturb=0.03),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_one_residual(case):
    """Against the FP64 oracle to the bar above, and against the extended-precision oracle:
    the HIP residual is within 4x of the FP64 oracle's own distance from the exact answer."""
    z, o, _ = paired(case, exact=True)
    z.compute_residual(0)
    o.compute_residual(0)
    err, kappa, ok = residual_parity(z.state("divF", 0), o)
    assert ok, f"{case['name']}: relative L-inf {err:.3e} (cancellation factor {kappa:.1e})"
    if kappa < 100:  # well-conditioned states meet the plain relative bar
        assert err <= RES_TOL, f"{case['name']}: relative L-inf {err:.3e}"


# The benchmark configurations' shapes and parameters on small meshes: 3D Navier-Stokes
# Taylor-Green (hex p=3, Re 1600, Ma 0.1, Pr 0.71) with Rusanov + LDG and Roe + LDG, and 2D
# Navier-Stokes Taylor-Green (quad p=4, Pr 0.72).
CONFIG_CASES = [
    dict(name="c3_hex4_p3_tgv_rusanov", n=(4, 4, 4), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True,
         state="tgv3d", dt=2e-4),
    dict(name="c3_hex4_p3_tgv_roe", n=(4, 4, 4), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True,
         state="tgv3d", dt=2e-4, extra={"fconv_type": "Roe"}),
    dict(name="c2_quad8_p4_tgv", n=(8, 8), lo=(0.0,) * 2, hi=(TWO_PI,) * 2, order=4, viscous=True,
         state="tgv2d", dt=2e-4, extra={"prandtl": 0.72}),
]


@pytest.mark.parametrize("case", CONFIG_CASES, ids=[c["name"] for c in CONFIG_CASES])
def test_config_residual_vs_exact(case):
    """The Taylor-Green states are near equilibrium, so |divF| is a cancellation of terms
    ~1e3-1e5 times larger and no FP64 evaluation meets 1e-12 of |divF| against the exact
    answer -- the FP64 oracle (the reference's algorithm) included.  Bar: the HIP residual's
    distance from the extended-precision oracle is at most 4x the FP64 oracle's, per variable."""
    z, o, _ = paired(case, exact=True)
    z.compute_residual(0)
    o.compute_residual(0)
    got = z.state("divF", 0)
    err, kappa, ok = residual_parity(got, o)
    from common import exact_errors
    eg, eo = exact_errors(got, o)
    print(f"{case['name']}: vs oracle {err:.2e} (kappa {kappa:.1e}); vs exact: product {eg.max():.2e}, "
          f"FP64 oracle {eo.max():.2e}")
    assert ok, (err, kappa, eg, eo)


@pytest.mark.parametrize("case", CASES[:5], ids=[c["name"] for c in CASES[:5]])
def test_rk4_steps(case):
    z, o, _ = paired(case)
    nsteps = 10
    for _ in range(nsteps):
        z.do_step()
    o.step_rk(nsteps)
    err = rel_err(z.state("U"), o.U())
    assert err <= RK_TOL, f"{case['name']}: relative L-inf after {nsteps} steps {err:.3e}"


LONG_CASES = CONFIG_CASES + [CASES[3], CASES[4]]  # + general LDG beta (hex p=2), beta = -1/2 Sutherland quads


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "separate"])
@pytest.mark.parametrize("case", LONG_CASES, ids=[c["name"] for c in LONG_CASES])
def test_rk4_100_steps(case, fused):
    """100 RK4 steps against the oracle at 1e-10 (relative L-inf of the state), with the RK
    update fused into each residual's last kernel (stages 1.. then read the trace the
    previous stage's update wrote) and with separate update kernels."""
    z, o, _ = paired(case)
    z.set_step_mode(fuse_rk=fused, use_graph=False)
    nsteps = 100
    for _ in range(nsteps):
        z.do_step()
    o.step_rk(nsteps)
    err = rel_err(z.state("U"), o.U())
    print(f"{case['name']} ({'fused' if fused else 'separate'}): relative L-inf after {nsteps} steps {err:.2e}")
    assert err <= RK_TOL, f"{case['name']}: relative L-inf after {nsteps} steps {err:.3e}"


@pytest.mark.parametrize("order,viscous,tiles", [(5, True, (3, 2, 4)), (4, True, (2, 2, 3)), (5, False, (4, 4, 3))],
                         ids=["q_ns_p5", "q_ns_p4", "q_euler_p5"])
def test_rk4_steps_multi_element_tiles(order, viscous, tiles):
    """Fused RK stages write the next stage's trace from the last kernel's LDS.  Quads at
    p >= 4 have more solution than flux points per element; tiles of several elements per
    workgroup check that the per-element blocks stay apart (the default tile is one element
    per workgroup on meshes this small)."""
    case = dict(name=f"q_tiled_p{order}", n=(6, 5), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=order,
                viscous=viscous, state="tgv2d")
    z, o, _ = paired(case)
    z.set_tiles(tiles)
    nsteps = 10
    for _ in range(nsteps):
        z.do_step()
    o.step_rk(nsteps)
    err = rel_err(z.state("U"), o.U())
    assert err <= RK_TOL, f"{case['name']}: relative L-inf after {nsteps} steps {err:.3e}"


@pytest.mark.parametrize("case", [CASES[2], CASES[0], CASES[3]], ids=["onesided_hex_p3", "euler_quad_p3",
                                                                      "general_beta_hex_p2"])
def test_autotuned_tiles_keep_parity(case):
    """The autotuner picks one workgroup tile per launch from timings; the residual and two
    RK4 steps at the tiles it picks still match the oracle.  Paths without a trace launch
    (one-sided LDG, Euler) leave that launch's tile unset, so it reports the solver-wide
    tile (the one-sided / divergence launch's)."""
    z, o, _ = paired(case)
    z.autotune(2)
    tiles = z.tiles()  # effective tiles: an unset one falls back to the solver-wide tile
    has_trace_launch = case["viscous"] and case.get("extra", {}).get("ldg_b", 0.5) not in (0.5, -0.5)
    assert min(tiles) > 0, tiles
    if not has_trace_launch:
        assert tiles[0] == tiles[1], tiles
    if not case["viscous"]:  # no middle launch either: every phase takes the divergence launch's tile
        assert tiles[1] == tiles[2], tiles
    z.compute_residual(0)
    o.compute_residual(0)
    err, kappa, ok = residual_parity(z.state("divF", 0), o)
    assert ok, f"{case['name']} at tiles {tiles}: relative L-inf {err:.3e} (kappa {kappa:.1e})"
    for _ in range(2):
        z.do_step()
    o.step_rk(2)
    assert rel_err(z.state("U"), o.U()) <= RK_TOL


def test_vortex_100_rk4_steps():
    """The CPU-runnable benchmark case: 2D Euler vortex, 32x32 quads on [-5,5]^2, p=3,
    Rusanov, RK4 dt=1e-3, 100 steps."""
    case = dict(name="c1_vortex", n=(32, 32), lo=(-5.0, -5.0), hi=(5.0, 5.0), order=3, viscous=False,
                state="vortex")
    z, o, _ = paired(case)
    z.compute_residual(0)
    o.compute_residual(0)
    err, kappa, ok = residual_parity(z.state("divF", 0), o)
    assert ok and err <= RES_TOL, (err, kappa)
    z.set_state(o.U())
    for _ in range(100):
        z.do_step()
    o.step_rk(100)
    err = rel_err(z.state("U"), o.U())
    assert err <= RK_TOL, f"relative L-inf after 100 steps {err:.3e}"
    res_z = z.write_residual()
    res_o = o.report_residuals()
    assert np.allclose(res_z, res_o, rtol=1e-10, atol=0), (res_z, res_o)


def test_forward_euler_scheme():
    case = dict(name="q_euler_fe", n=(6, 6), lo=(-5.0, -5.0), hi=(5.0, 5.0), order=2, viscous=False,
                state="vortex", dt_scheme="Euler")
    z, o, _ = paired(case)
    for _ in range(5):
        z.do_step()
    o.step_rk(5)
    assert rel_err(z.state("U"), o.U()) <= RK_TOL


def test_rk_stage_interface_matches_do_step():
    """do_rk_stage_start/mid/finish over all stages == do_step (src/zefr.cpp:385-430)."""
    case = dict(name="q_ns_stages", n=(4, 4), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=3, viscous=True,
                state="tgv2d")
    z1, o, U0 = paired(case)
    z2, _, _ = paired(dict(case, name="q_ns_stages_b"))
    for it in range(3):
        z1.do_step()
        for st in range(4):
            z2.do_rk_stage_start(it, st)
            z2.do_rk_stage_mid(it, st)
            z2.do_rk_stage_finish(it, st)
    np.testing.assert_array_equal(z1.state("U"), z2.state("U"))
    o.step_rk(3)
    assert rel_err(z1.state("U"), o.U()) <= RK_TOL


def test_residual_norm_types():
    case = dict(name="h_norms", n=(3, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=2, viscous=True,
                state="tgv3d")
    z, o, _ = paired(case)
    z.compute_residual(0)
    o.compute_residual(0)
    for rt in (0, 1, 2):
        z.get_input().res_type = rt
        o_res = _oracle_norm(o, rt)
        assert np.allclose(z.write_residual(), o_res, rtol=1e-12, atol=0)


def _oracle_norm(o, res_type):
    d = o.divF(0)
    j = o.jaco_det()[:, None, :]
    x = d / j
    if res_type == 0:
        r = np.abs(x).max(axis=(0, 2))
    elif res_type == 1:
        r = np.abs(x).sum(axis=(0, 2))
    else:
        r = np.sqrt((x * x).sum(axis=(0, 2)))
    return r / (o.nSpts * o.nEles)


def test_freestream_preserved_on_curved_mesh():
    case = dict(name="h_free", n=(3, 3, 3), lo=(0.0,) * 3, hi=(1.0,) * 3, order=3, viscous=True,
                state="smooth", perturb=0.05)
    z, o, U0 = paired(case)
    U = np.zeros_like(U0)
    U[:, 0] = 1.0; U[:, 1] = 0.3; U[:, 2] = -0.2; U[:, 3] = 0.1; U[:, 4] = 2.6
    z.set_state(U)
    z.compute_residual(0)
    assert np.abs(z.state("divF", 0)).max() < 1e-11


def test_determinism():
    case = dict(name="h_det", n=(3, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True,
                state="tgv3d")
    z, _, _ = paired(case)
    z.compute_residual(0)
    a = z.state("divF", 0)
    z.compute_residual(0)
    b = z.state("divF", 0)
    np.testing.assert_array_equal(a, b)


def test_gradient_export_matches_oracle():
    """get_du_spts: physical gradient at spts (the reference writes it into dU_spts in
    compute_F, src/elements.cpp:1630-1650)."""
    case = dict(name="q_grad", n=(4, 3), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=3, viscous=True,
                state="tgv2d", perturb=0.03)
    z, o, _ = paired(case)
    o.compute_residual(0)
    dU, es, ss, vs, ds = z.get_du_spts()
    ref = o.array("dU_spts", (o.nDims, o.nSpts, o.nVars, o.nEles))
    got = dU.transpose(1, 3, 2, 0)  # [E][D][V][nS] -> [D][nS][V][E]
    assert rel_err(got, ref, axis_var=2) <= RES_TOL
    assert (es, ss, vs, ds) == (o.nDims * o.nVars * o.nSpts, 1, o.nSpts, o.nVars * o.nSpts)


def test_tiles_agree():
    """Every workgroup tile gives the same residual bit for bit (same per-element arithmetic)."""
    case = dict(name="h_tiles", n=(5, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True,
                state="tgv3d")
    z, _, _ = paired(case)
    outs = []
    for eb in (1, 2, 3, 4):
        z.set_tile(eb)
        z.compute_residual(0)
        outs.append(z.state("divF", 0))
    for a in outs[1:]:
        np.testing.assert_array_equal(a, outs[0])
