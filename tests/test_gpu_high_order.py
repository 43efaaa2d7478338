"""Quads of order 6..9 and hexes of order 6 and 7 on the GPU against the oracle (the
reference's Gauss-Legendre tables run to 10 points, src/points.cpp:9-209; the p = 9 set
keeps the reference's 10-point node +-0.433895394129247, as the oracle does).

* one residual per order and LDG branch (Euler, one-sided beta = +-1/2, general beta,
  curved meshes, Sutherland, Roe) at the one-residual bar of test_gpu_parity.py, and
  against the extended-precision oracle;
* 100 RK4 steps at 1e-10 with fused and separate updates;
* the residual VJP against fourth-order differences of the oracle's residual;
* p-multigrid V-cycles from p = 9 (a prolongation whose table does not fit the fused
  epilogue's region runs as its own transfer kernel);
* a batch whose members of order 6..9 get launches of their own shape beside the
  mixed-order launches.
"""

from __future__ import annotations

import numpy as np
import pytest

from common import TWO_PI, paired, rel_err, residual_parity

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
RK_TOL = 1e-10

CASES = [
    dict(name="q_euler_p6_vortex", n=(6, 5), lo=(-5.0, -5.0), hi=(5.0, 5.0), order=6, viscous=False,
         state="vortex"),
    dict(name="q_ns_p7_tgv", n=(4, 4), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=7, viscous=True,
         state="tgv2d"),
    dict(name="q_ns_p8_curved_sutherland_bm", n=(4, 3), lo=(0.0, 0.0), hi=(2.0, 1.0), order=8, viscous=True,
         state="smooth", perturb=0.04,
         extra={"ldg_b": -0.5, "rus_k": 0.3, "fix_vis": 0, "disable_nondim": 0, "mach_fs": 0.3,
                "Re_fs": 100.0, "rho_fs": 1.2}),
    dict(name="q_ns_p9_curved_beta02", n=(3, 4), lo=(0.0, 0.0), hi=(1.0, 1.5), order=9, viscous=True,
         state="smooth", perturb=0.03, extra={"ldg_b": 0.2}),
    dict(name="q_ns_p9_roe", n=(3, 3), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=9, viscous=True,
         state="smooth", extra={"fconv_type": "Roe"}),
    dict(name="q_euler_p9_ragged", n=(5, 3), lo=(0.0, 0.0), hi=(1.0, 1.0), order=9, viscous=False,
         state="smooth", perturb=0.03),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_one_residual(case):
    print("[high order]", case["name"], flush=True)
    z, o, _ = paired(case, exact=True)
    z.compute_residual(0)
    o.compute_residual(0)
    err, kappa, ok = residual_parity(z.state("divF", 0), o)
    print(f"{case['name']}: relative L-inf {err:.2e} (kappa {kappa:.1e})")
    assert ok, f"{case['name']}: relative L-inf {err:.3e} (cancellation factor {kappa:.1e})"
    if kappa < 100:
        assert err <= RES_TOL, f"{case['name']}: relative L-inf {err:.3e}"


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "separate"])
@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[3]], ids=lambda c: c["name"])
def test_rk4_100_steps(case, fused):
    print("[high order rk]", case["name"], fused, flush=True)
    z, o, _ = paired(dict(case, dt=2e-4))
    z.set_step_mode(fuse_rk=fused, use_graph=False)
    nsteps = 100
    for _ in range(nsteps):
        z.do_step()
    o.step_rk(nsteps)
    err = rel_err(z.state("U"), o.U())
    print(f"{case['name']} ({'fused' if fused else 'separate'}): relative L-inf after {nsteps} steps {err:.2e}")
    assert err <= RK_TOL, f"{case['name']}: relative L-inf after {nsteps} steps {err:.3e}"


@pytest.mark.parametrize("case", [CASES[1], CASES[3], CASES[4]], ids=lambda c: c["name"])
def test_vjp_directional_vs_oracle_finite_differences(case):
    """<d, J^T w> from the VJP against <w, J d> from fourth-order central differences of the
    oracle's residual.  Extrapolating pointwise noise to the faces of a high-order element
    amplifies it (about tenfold at p = 9), so some directions put a face state within the
    differencing step of a Rusanov switching point (|u.n| = 0 or tied wave speeds): a
    difference is used only where it agrees with the one at half the step to 1e-9 (no
    switch crossed), and two such directions are required."""
    import torch
    from test_gpu_vjp import fd4, noisy_state, oracle_R, to_dev
    print("[high order vjp]", case["name"], flush=True)
    z, o, _ = paired(dict(case, name=case["name"] + "_vjp"))
    rng = np.random.default_rng(5)
    U = noisy_state(o, 6, amp=0.01)
    Ut = torch.from_numpy(to_dev(U)).cuda()
    R = lambda x: oracle_R(o, x)  # noqa: E731
    checked = 0
    for trial in range(6):
        w = rng.standard_normal(U.shape)
        d = rng.standard_normal(U.shape)
        jd = None
        for h in (1e-5, 5e-6, 2.5e-6):
            a, b = fd4(R, U, d, h), fd4(R, U, d, h / 2)
            if abs(np.sum(w * a) - np.sum(w * b)) <= 1e-9 * np.sum(np.abs(w * a)):
                jd = b
                break
        if jd is None:
            continue
        v = z.residual_vjp(Ut, torch.from_numpy(to_dev(w)).cuda()).cpu().numpy()
        lhs = np.sum(to_dev(d) * v)
        rhs = np.sum(w * jd)
        scale = np.sum(np.abs(w * jd)) + 1e-300
        assert abs(lhs - rhs) <= 1e-8 * scale, (trial, lhs, rhs)
        checked += 1
        if checked == 2:
            break
    assert checked == 2, f"{case['name']}: fewer than two directions without a switching point"


@pytest.mark.parametrize("levels,steps,viscous", [([9, 7, 4, 1, 0], [1, 1, 1, 2, 2], False),
                                                  ([9, 2, 0], [1, 2, 2], True)],
                         ids=["euler_9_7_4_1_0", "ns_9_2_0"])
def test_v_cycles_from_order_9(levels, steps, viscous):
    """9 -> 7 -> 4 -> 1 -> 0: every correction prolonged by its level's last step.
    9 -> 2 -> 0: the [10][3] prolongation table of level 2 exceeds the 24 doubles of the
    p = 2 kernels' table region, so that correction is prolonged by the transfer kernel."""
    from test_gpu_pmg import _check_cycle, _paired_pmg, _pmg_case
    case = _pmg_case(f"pmg_q_p9_{'ns' if viscous else 'euler'}", (4, 4), (0.0, 0.0), (TWO_PI, TWO_PI), 9,
                     viscous, "noisy:3" if not viscous else "smooth", levels, steps, dt=5e-5)
    print("[high order pmg]", case["name"], flush=True)
    z, o, p = _paired_pmg(case)
    for it in range(3):
        z.do_step()
        p.cycle(1)
        _check_cycle(z, o, p, case, it)
    # fused and separate smoothing steps agree bit for bit
    out = {}
    for fuse in (True, False):
        z2, _, _ = paired(dict(case, name=case["name"] + ("_f" if fuse else "_s")))
        z2.set_step_mode(fuse, False)
        z2.do_step()
        out[fuse] = [z2.state("U")] + [z2.pmg_state(lv, "correction") for lv in range(1, len(levels))]
    for a, b in zip(out[True], out[False]):
        assert np.array_equal(a, b)


def test_batch_with_high_orders():
    """Members of order 6..9 run in launches of their own shape; the others keep the
    mixed-order launches.  Every member's batched residual equals its eager one."""
    from zefr_amd.zefr import Batch
    members = [CASES[0], CASES[1], CASES[5],
               dict(name="hb_q_euler_p3", n=(5, 5), lo=(0.0, 0.0), hi=(1.0, 1.0), order=3, viscous=False,
                    state="noisy:2"),
               dict(name="hb_q_ns_p4", n=(4, 4), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=4, viscous=True,
                    state="tgv2d")]
    zs, refs = [], []
    for case in members:
        print("[high order batch] member", case["name"], flush=True)
        z, _, _ = paired(dict(case, name=case["name"] + "_b"))
        z.compute_residual(0)
        refs.append(z.state("divF", 0, layout="device").copy())
        zs.append(z)
    b = Batch(zs)
    # Euler: mixed (p = 3) + p = 6 + p = 9; one-sided: mixed (p = 4) and p = 7, each with
    # its face correction
    assert b.launches == 7, b.launches
    for rep in range(2):
        b.residual()
        for case, z, ref in zip(members, zs, refs):
            got = z.state("divF", 0, layout="device")
            err = float(np.abs(got - ref).max() / np.abs(ref).max())
            assert err <= 1e-15, (rep, case["name"], err)
    del b


# Hexes of order 6 to 8: one element per workgroup of 343 / 512 / 729 threads (up to 160
# KiB of LDS), partner-map entries with 9 flux-point bits (294 / 384 / 486 flux points).
HEX_CASES = [
    dict(name="h_euler_p6_curved", n=(2, 2, 3), lo=(0.0,) * 3, hi=(1.0,) * 3, order=6, viscous=False,
         state="smooth", perturb=0.02),
    dict(name="h_ns_p7_tgv", n=(2, 2, 2), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=7, viscous=True,
         state="tgv3d"),
    dict(name="h_ns_p6_beta02_curved", n=(2, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=6, viscous=True,
         state="smooth", perturb=0.02, extra={"ldg_b": 0.2}),
    dict(name="h_ns_p7_roe_bm", n=(2, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=7, viscous=True,
         state="smooth", extra={"fconv_type": "Roe", "ldg_b": -0.5}),
    dict(name="h_euler_p8", n=(2, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=8, viscous=False,
         state="smooth", perturb=0.02),
    dict(name="h_ns_p8_tgv", n=(2, 2, 2), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=8, viscous=True,
         state="tgv3d"),
    dict(name="h_ns_p8_beta03", n=(2, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=8, viscous=True,
         state="smooth", extra={"ldg_b": 0.3}),
]


@pytest.mark.parametrize("case", HEX_CASES, ids=[c["name"] for c in HEX_CASES])
def test_hex_one_residual(case):
    test_one_residual(case)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "separate"])
@pytest.mark.parametrize("case", [HEX_CASES[0], HEX_CASES[1], HEX_CASES[5]], ids=lambda c: c["name"])
def test_hex_rk4_50_steps(case, fused):
    print("[high order rk]", case["name"], fused, flush=True)
    z, o, _ = paired(dict(case, name=case["name"] + ("_f" if fused else "_s"), dt=1e-4))
    z.set_step_mode(fuse_rk=fused, use_graph=False)
    nsteps = 50
    for _ in range(nsteps):
        z.do_step()
    o.step_rk(nsteps)
    err = rel_err(z.state("U"), o.U())
    print(f"{case['name']} ({'fused' if fused else 'separate'}): relative L-inf after {nsteps} steps {err:.2e}")
    assert err <= RK_TOL, f"{case['name']}: relative L-inf after {nsteps} steps {err:.3e}"


@pytest.mark.parametrize("case", [HEX_CASES[1], HEX_CASES[2]], ids=lambda c: c["name"])
def test_hex_vjp_directional_vs_oracle_finite_differences(case):
    test_vjp_directional_vs_oracle_finite_differences(case)


def test_hex_v_cycles_from_order_7():
    """7 -> 4 -> 1 -> 0 on hexes (Euler): every correction prolonged by its level's last step."""
    from test_gpu_pmg import _check_cycle, _paired_pmg, _pmg_case
    case = _pmg_case("pmg_h_p7", (2, 2, 2), (0.0,) * 3, (1.0,) * 3, 7, False, "noisy:4", [7, 4, 1, 0],
                     [1, 1, 2, 2], dt=2e-5)
    print("[high order pmg]", case["name"], flush=True)
    z, o, p = _paired_pmg(case)
    for it in range(2):
        z.do_step()
        p.cycle(1)
        _check_cycle(z, o, p, case, it)


def test_batch_with_high_order_hexes():
    """Hex members of order 6 and 7 run in launches of their own shape beside a mixed-order
    launch; every member's batched residual equals its eager one."""
    from zefr_amd.zefr import Batch
    members = [HEX_CASES[0], HEX_CASES[1],
               dict(name="hb_h_euler_p2", n=(3, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=2, viscous=False,
                    state="noisy:9"),
               dict(name="hb_h_ns_p3", n=(2, 2, 2), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True,
                    state="tgv3d")]
    zs, refs = [], []
    for case in members:
        print("[high order batch] member", case["name"], flush=True)
        z, _, _ = paired(dict(case, name=case["name"] + "_hb"))
        z.compute_residual(0)
        refs.append(z.state("divF", 0, layout="device").copy())
        zs.append(z)
    b = Batch(zs)
    # Euler: p = 6 alone, p = 2 mixed; one-sided: p = 7 alone and p = 3 mixed, each with a
    # face correction
    assert b.launches == 6, b.launches
    for rep in range(2):
        b.residual()
        for case, z, ref in zip(members, zs, refs):
            got = z.state("divF", 0, layout="device")
            err = float(np.abs(got - ref).max() / np.abs(ref).max())
            assert err <= 1e-15, (rep, case["name"], err)
    del b


def test_hex_p8_vjp_is_refused_with_the_reason():
    """At hex order 8 one element's viscous adjoint needs 199 KiB of LDS, more than a
    workgroup can allocate: the VJP fails with that reason, and the forward residual of
    the same solver still works afterwards."""
    import torch
    from test_gpu_vjp import noisy_state, to_dev
    z, o, _ = paired(dict(HEX_CASES[5], name="h_ns_p8_tgv_vjp"))
    U = torch.from_numpy(to_dev(noisy_state(o, 2, amp=0.01))).cuda()
    with pytest.raises(RuntimeError, match="hexes to order 7"):
        z.residual_vjp(U, torch.ones_like(U))
    z.compute_residual(0)
    assert np.isfinite(z.state("divF", 0)).all()
