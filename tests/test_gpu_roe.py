"""The Roe flux on the GPU against the oracle's textbook Roe (the reference has no Roe flux;
both sides are pinned by tests/test_roe.py).  One residual and ten RK4 steps on the Euler,
one-sided LDG (beta = +-1/2) and general-beta LDG kernel paths, including the benchmark's
Taylor-Green field on hexes at p=3; the residual VJP with Roe against differences of the
oracle's Roe residual."""

from __future__ import annotations

import pytest

from common import TWO_PI, paired, rel_err, residual_parity

pytestmark = pytest.mark.gpu

ROE = {"fconv_type": "Roe"}

CASES = [
    dict(name="roe_q_euler_p3_vortex", n=(8, 8), lo=(-5.0, -5.0), hi=(5.0, 5.0), order=3, viscous=False,
         state="vortex", extra=ROE),
    dict(name="roe_h_ns_p3_tgv", n=(3, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True,
         state="tgv3d", extra=ROE),
    dict(name="roe_q_ns_p4_bm", n=(5, 4), lo=(0.0, 0.0), hi=(2.0, 1.0), order=4, viscous=True, state="smooth",
         perturb=0.04, extra=dict(ROE, ldg_b=-0.5)),
    dict(name="roe_h_ns_p2_central", n=(3, 3, 2), lo=(0.0,) * 3, hi=(1.0, 2.0, 1.5), order=2, viscous=True,
         state="smooth", perturb=0.04, extra=dict(ROE, ldg_b=0.0, ldg_tau=1.0)),
    dict(name="roe_h_euler_p1_noisy", n=(4, 3, 3), lo=(0.0,) * 3, hi=(1.0,) * 3, order=1, viscous=False,
         state="noisy:5", extra=ROE),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_roe_residual_vs_oracle(case):
    print("[roe]", case["name"], flush=True)
    z, o, _ = paired(case)
    assert z.get_input().fconv_type == 2
    z.compute_residual(0)
    o.compute_residual(0)
    err, kappa, ok = residual_parity(z.state("divF", 0), o)
    assert ok, f"{case['name']}: relative L-inf {err:.3e} (cancellation factor {kappa:.1e})"


@pytest.mark.parametrize("case", CASES[:3], ids=[c["name"] for c in CASES[:3]])
def test_roe_rk4_steps_vs_oracle(case):
    z, o, _ = paired(case)
    for _ in range(10):
        z.do_step()
    o.step_rk(10)
    assert rel_err(z.state("U"), o.U()) <= 1e-10


@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: c["name"])
def test_roe_vjp_vs_oracle_finite_differences(case):
    """<J^T w, d> from the residual VJP with Roe against fourth-order central differences of
    the oracle's Roe residual along d (states away from the flux's kinks)."""
    import numpy as np
    import torch
    from test_gpu_vjp import fd4, noisy_state, oracle_R, to_dev
    z, o, _ = paired(case)
    rng = np.random.default_rng(3)
    U = noisy_state(o, 4)
    Ut = torch.from_numpy(to_dev(U)).cuda()
    for trial in range(2):
        w = rng.standard_normal(U.shape)
        d = rng.standard_normal(U.shape)
        v = z.residual_vjp(Ut, torch.from_numpy(to_dev(w)).cuda()).cpu().numpy()
        lhs = np.sum(to_dev(d) * v)
        jd = fd4(lambda x: oracle_R(o, x), U, d, 1e-5)
        rhs = np.sum(w * jd)
        assert abs(lhs - rhs) <= 1e-8 * (np.sum(np.abs(w * jd)) + 1e-300), (trial, lhs, rhs)
