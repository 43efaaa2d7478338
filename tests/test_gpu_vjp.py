"""The residual's vector-Jacobian product on the GPU (zefr_residual_vjp through the C ABI)
against finite differences of the residual: the oracle's (CPU) and the product's own.

* directional test: <w, J d> by fourth-order central differences of R along d vs.
  <J^T w, d> from the VJP, for random w and d;
* a Jacobian-transpose assembled column by column by finite differences on a tiny mesh,
  compared with VJP rows (one-sided LDG beta = +-1/2 and general beta);
* torch autograd through R and through RK4 steps of the kinetic energy vs. central
  finite differences (eps = 1e-7) of the same functional.
Test states are random perturbations of a uniform flow, so interface jumps keep the
Rusanov wave-speed maximum and |u.n| away from their switching points.
"""

from __future__ import annotations

import numpy as np
import pytest

from common import TWO_PI, make_mesh, paired, tmp, write_deck, product
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def noisy_state(o, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    D, V = o.nDims, o.nVars
    vel = np.array([0.3, 0.2, 0.1][:D])
    rho = 1.0 + amp * rng.standard_normal((o.nSpts, o.nEles))
    mom = rho[:, None, :] * (vel[None, :, None] + amp * rng.standard_normal((o.nSpts, D, o.nEles)))
    p = 1.0 + amp * rng.standard_normal((o.nSpts, o.nEles))
    E = p / 0.4 + 0.5 * (mom ** 2).sum(axis=1) / rho
    return np.concatenate([rho[:, None, :], mom, E[:, None, :]], axis=1)


def to_dev(a):
    """[nSpts][nVars][nEles] -> [nEles][nVars][nSpts]."""
    return np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0))


def to_ref(a):
    return np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0))


def oracle_R(o, U_ref):
    o.set_U(U_ref)
    o.compute_residual(0)
    return o.divF(0).copy()


def fd4(f, x, d, h):
    return (-f(x + 2 * h * d) + 8 * f(x + h * d) - 8 * f(x - h * d) + f(x - 2 * h * d)) / (12 * h)


CASES = [
    dict(name="vjp_q_euler_p3", n=(4, 3), lo=(0.0, 0.0), hi=(1.0, 1.0), order=3, viscous=False, state="smooth"),
    dict(name="vjp_q_ns_p2_curved", n=(3, 4), lo=(0.0, 0.0), hi=(1.0, 1.5), order=2, viscous=True,
         state="smooth", perturb=0.04),
    dict(name="vjp_q_ns_p3_bm", n=(3, 3), lo=(0.0, 0.0), hi=(1.0, 1.0), order=3, viscous=True, state="smooth",
         extra={"ldg_b": -0.5, "fix_vis": 0, "disable_nondim": 0, "rus_k": 0.2}),
    dict(name="vjp_h_ns_p2", n=(3, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=2, viscous=True, state="smooth",
         perturb=0.03),
    dict(name="vjp_h_ns_p3", n=(2, 2, 2), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True, state="smooth"),
    dict(name="vjp_h_euler_p1", n=(3, 3, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=1, viscous=False, state="smooth"),
    # general LDG beta: both sides' gradients in every common flux, the common solution
    # a weighted mix of both traces (src/faces.cpp:1366-1474, 1520-1545)
    dict(name="vjp_q_ns_p3_beta02_curved", n=(3, 4), lo=(0.0, 0.0), hi=(1.0, 1.5), order=3, viscous=True,
         state="smooth", perturb=0.04, extra={"ldg_b": 0.2}),
    dict(name="vjp_h_ns_p2_beta_m03_sutherland", n=(2, 2, 3), lo=(0.0,) * 3, hi=(1.0,) * 3, order=2, viscous=True,
         state="smooth", extra={"ldg_b": -0.3, "fix_vis": 0, "disable_nondim": 0, "rus_k": 0.1}),
    dict(name="vjp_h_ns_p3_beta0_roe", n=(2, 2, 2), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True,
         state="smooth", extra={"ldg_b": 0.0, "ldg_tau": 0.5, "fconv_type": "Roe"}),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_vjp_directional_vs_oracle_finite_differences(case):
    import torch
    z, o, _ = paired(case)
    print(f"[vjp] {case['name']}", flush=True)
    rng = np.random.default_rng(1)
    U = noisy_state(o, 2)
    Ut = torch.from_numpy(to_dev(U)).cuda()
    for trial in range(3):
        w = rng.standard_normal(U.shape)
        d = rng.standard_normal(U.shape)
        v = z.residual_vjp(Ut, torch.from_numpy(to_dev(w)).cuda()).cpu().numpy()
        lhs = np.sum(to_dev(d) * v)
        # h small enough that no Rusanov wave-speed maximum switches sides along d
        jd = fd4(lambda x: oracle_R(o, x), U, d, 1e-5)
        rhs = np.sum(w * jd)
        scale = np.sum(np.abs(w * jd)) + 1e-300
        assert abs(lhs - rhs) <= 1e-8 * scale, (trial, lhs, rhs)


@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[3]], ids=lambda c: c["name"])
def test_forward_residual_at_matches_solver(case):
    """residual_at(U) equals compute_residual on the same state and the oracle."""
    import torch
    z, o, _ = paired(case)
    U = noisy_state(o, 3)
    z.set_state(U)
    z.compute_residual(0)
    ref = z.state("divF", 0)
    got = z.residual_at(torch.from_numpy(to_dev(U)).cuda()).cpu().numpy()
    np.testing.assert_array_equal(to_ref(got), ref)
    want = oracle_R(o, U)
    assert np.abs(to_ref(got) - want).max() <= 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("viscous,beta", [(False, 0.5), (True, 0.5), (True, -0.5), (True, 0.2), (True, 0.0)])
def test_vjp_equals_column_assembled_jacobian_transpose(viscous, beta):
    """Tiny mesh (2x2 quads, p=1: 64 unknowns): J assembled column by column from central
    differences of the product's own forward residual; every VJP row e_i^T J^T ... is
    compared (all 64 unit cotangents)."""
    import torch
    case = dict(name=f"vjp_cols_{int(viscous)}_{beta}", n=(2, 2), lo=(0.0, 0.0), hi=(1.0, 1.0), order=1,
                viscous=viscous, state="smooth", extra={"ldg_b": beta})
    z, o, _ = paired(case)
    U = noisy_state(o, 4, amp=0.08)
    Ud = to_dev(U)
    n = Ud.size
    R = lambda x: z.residual_at(torch.from_numpy(x.reshape(Ud.shape)).cuda()).cpu().numpy().ravel()  # noqa: E731
    J = np.zeros((n, n))
    h = 1e-5
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        J[:, k] = fd4(R, Ud.ravel(), e, h)
    Ut = torch.from_numpy(Ud).cuda()
    JT = np.zeros((n, n))
    for i in range(n):
        w = np.zeros(n)
        w[i] = 1.0
        JT[:, i] = z.residual_vjp(Ut, torch.from_numpy(w.reshape(Ud.shape)).cuda()).cpu().numpy().ravel()
    assert np.abs(JT - J.T).max() <= 1e-7 * np.abs(J).max()


@pytest.mark.parametrize("beta", [0.5, 0.2])
def test_vjp_dot_product_with_forward_linearisation(beta):
    """<w, J d> = <J^T w, d> with J d from central differences of the product's own
    forward residual (eps 1e-6, second order) on 3^3 hexes at p=2: the VJP is the exact
    transpose of the forward linearisation, whatever the LDG branch."""
    import torch
    case = dict(name=f"vjp_dot_{beta}", n=(3, 3, 3), lo=(0.0,) * 3, hi=(1.0,) * 3, order=2, viscous=True,
                state="smooth", perturb=0.02, extra={"ldg_b": beta})
    z, o, _ = paired(case)
    U = to_dev(noisy_state(o, 11))
    R = lambda x: z.residual_at(torch.from_numpy(x).cuda()).cpu().numpy()  # noqa: E731
    rng = np.random.default_rng(12)
    for _ in range(2):
        w = rng.standard_normal(U.shape)
        d = rng.standard_normal(U.shape)
        jd = fd4(R, U, d, 1e-5)
        v = z.residual_vjp(torch.from_numpy(U).cuda(), torch.from_numpy(w).cuda()).cpu().numpy()
        lhs, rhs = np.sum(v * d), np.sum(w * jd)
        assert abs(lhs - rhs) <= 1e-8 * (np.sum(np.abs(w * jd)) + 1e-300), (lhs, rhs)


def test_autograd_residual_backward():
    import torch
    from zefr_amd.autograd import residual
    case = CASES[3]
    z, o, _ = paired(case)
    U = torch.from_numpy(to_dev(noisy_state(o, 5))).cuda().requires_grad_()
    w = torch.randn_like(U)
    (residual(z, U) * w).sum().backward()
    v = z.residual_vjp(U.detach(), w)
    assert torch.equal(U.grad, v)


def _tgv_torch(X, Y, Z, Ma, gamma=1.4, P0=1.0, rho0=1.0, R=286.9, drift=(0.02, 0.015, 0.01)):
    """test_case 4 (src/funcs.cpp:125-155) as a torch function of Ma, plus a small uniform
    drift so that no face normal velocity sits exactly on the |u.n| kink of Rusanov."""
    import torch
    U0 = Ma * (gamma * P0 / rho0) ** 0.5
    u = U0 * torch.sin(X) * torch.cos(Y) * torch.cos(Z) + drift[0]
    v = -U0 * torch.cos(X) * torch.sin(Y) * torch.cos(Z) + drift[1]
    w = torch.zeros_like(u) + drift[2]
    p = P0 + rho0 * U0 * U0 / 16.0 * (torch.cos(2 * X) + torch.cos(2 * Y)) * (torch.cos(2 * Z) + 2)
    rho = p / (R * (P0 / (rho0 * R)))
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return torch.stack([rho, rho * u, rho * v, rho * w, E], dim=1)


def _ke_case():
    import torch
    case = dict(name="vjp_ke", n=(4, 4, 4), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=2, viscous=True,
                state="tgv3d", dt=2e-3)
    z, o, _ = paired(case)
    x = torch.from_numpy(np.ascontiguousarray(o.coord_spts().transpose(2, 1, 0))).cuda()  # [E][D][nS]
    return z, x[:, 0], x[:, 1], x[:, 2]


def test_unsteady_adjoint_initial_state_vs_finite_differences():
    """d/dU0 of the kinetic energy after 3 RK4 steps (3D NS, 4^3 hexes, p=2): autograd
    through the VJP, with checkpointed blocks, against central differences (eps = 1e-7)."""
    import torch
    from zefr_amd.autograd import checkpointed_rollout, kinetic_energy, rollout
    z, X, Y, Z = _ke_case()
    dt, steps = 2e-3, 3
    U = _tgv_torch(X, Y, Z, 0.1).contiguous().requires_grad_()
    J = kinetic_energy(z, checkpointed_rollout(z, U, dt, steps, every=2))
    J.backward()
    g = U.grad.detach().cpu().numpy()
    rng = np.random.default_rng(9)
    for _ in range(3):
        d = rng.standard_normal(g.shape)
        dt_ = torch.from_numpy(d).cuda()
        with torch.no_grad():
            Jp = kinetic_energy(z, rollout(z, U.detach() + 1e-7 * dt_, dt, steps)).item()
            Jm = kinetic_energy(z, rollout(z, U.detach() - 1e-7 * dt_, dt, steps)).item()
        fd = (Jp - Jm) / 2e-7
        assert abs(fd - np.sum(g * d)) <= 1e-6 * max(abs(fd), 1e-12)


def test_unsteady_adjoint_mach_number_vs_finite_differences():
    """dJ/dMa with the Mach number entering the initial state and the viscosity
    mu = rho U0 L / Re (disable_nondim), through autograd and the residual's mu-derivative."""
    import torch
    from zefr_amd.autograd import checkpointed_rollout, kinetic_energy, rollout
    z, X, Y, Z = _ke_case()
    dt, steps, Re = 2e-3, 3, 1600.0

    def J_of(Ma, grad):
        mu = Ma * 1.4 ** 0.5 / Re
        U = _tgv_torch(X, Y, Z, Ma).contiguous()
        run = checkpointed_rollout if grad else rollout
        kw = dict(every=2) if grad else {}
        return kinetic_energy(z, run(z, U, dt, steps, mu=mu, **kw))
    Ma = torch.tensor(0.1, dtype=torch.float64, device="cuda", requires_grad=True)
    J = J_of(Ma, True)
    J.backward()
    with torch.no_grad():
        eps = 1e-7
        fd = (J_of(torch.tensor(0.1 + eps, dtype=torch.float64, device="cuda"), False).item()
              - J_of(torch.tensor(0.1 - eps, dtype=torch.float64, device="cuda"), False).item()) / (2 * eps)
    assert abs(fd - Ma.grad.item()) <= 1e-6 * abs(fd)
