"""The benchmark configurations at their full sizes on the GPU.

* 2D Navier-Stokes Taylor-Green, 256^2 quads on [0,2pi]^2, p=4: one residual and five
  RK4 steps against the oracle on the same mesh and state.
* 3D Navier-Stokes Taylor-Green, 64^3 hexes on [0,2pi]^3, p=3 (too large for the oracle
  in a test): size-independent properties - discrete conservation of divF, freestream
  preservation, bit-reproducibility, and the VJP dot-product identity against central
  differences of the product's own residual.
* Adjoint: kinetic energy after 50 RK4 steps on 32^3 hexes, p=3, with checkpoints every
  10 steps; d/dU0 along random directions and d/dMa against central differences
  (eps = 1e-5; the 4^3-hex check of test_gpu_vjp.py uses 1e-7).
The 2D vortex case (32^2 quads, p=3, 100 RK4 steps) is in test_gpu_parity.py.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from common import TWO_PI, paired, product, residual_parity, tmp, write_deck, make_mesh

pytestmark = pytest.mark.gpu


def test_c2_quad_p4_tgv_256_vs_oracle():
    case = dict(name="c2_tgv2d_256_p4", n=(256, 256), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=4,
                viscous=True, state="tgv2d", extra={"prandtl": 0.72}, dt=1e-3)
    print("[full] c2 setup", flush=True)
    z, o, _ = paired(case)
    z.compute_residual(0)
    o.compute_residual(0)
    err, kappa, ok = residual_parity(z.state("divF", 0), o)
    assert ok, (err, kappa)
    for _ in range(5):
        z.do_step()
    o.step_rk(5)
    ref = o.U()
    assert np.abs(z.state("U") - ref).max() <= 1e-10 * np.abs(ref).max()


def _c3():
    mesh = make_mesh("c3_hex64.msh", (64, 64, 64), (0.0,) * 3, (TWO_PI,) * 3)
    deck = write_deck(tmp("c3.deck"), mesh, 3, 3, True, dt=2e-4, ic_type=1, test_case=4)
    return product(deck)


def test_c3_hex_p3_64_properties():
    import torch
    print("[full] c3 setup", flush=True)
    z = _c3()
    d = z.dims()
    E, nS, V = d["nEles"], d["nSpts"], d["nVars"]
    w = z.host_array("weights_spts")
    # discrete conservation: sum_e sum_spt w divF = 0 on a periodic mesh
    z.compute_residual(0)
    dv = z.state("divF", 0, layout="device")  # [E][V][nS]
    tot = np.einsum("s,evs->v", w, dv)
    scale = np.einsum("s,evs->v", w, np.abs(dv))
    # each divF value carries rounding ~eps of its largest terms (the pressure flux, up to
    # 2e5 x |divF| for this near-equilibrium state), summed over 16.8M points
    assert (np.abs(tot) <= 1e-10 * scale).all(), (tot, scale)
    # bit-reproducible
    z.compute_residual(0)
    assert np.array_equal(z.state("divF", 0, layout="device"), dv)
    # VJP dot-product identity against central differences of the product's residual,
    # at a perturbed state (the TGV field has u.n = 0 exactly on some faces, a kink of
    # the Rusanov flux that a difference quotient would straddle)
    rng = np.random.default_rng(3)
    U = z.state_tensor()
    U = U + 1e-3 * torch.from_numpy(rng.standard_normal((E, V, nS))).cuda()
    wt = torch.from_numpy(rng.standard_normal((E, V, nS))).cuda()
    dt = torch.from_numpy(rng.standard_normal((E, V, nS)) * 1e-3).cuda()
    v = z.residual_vjp(U, wt)
    h = 1e-3
    jd = (z.residual_at(U + h * dt) - z.residual_at(U - h * dt)) / (2 * h)
    lhs = float((v * dt).sum())
    rhs = float((wt * jd).sum())
    assert abs(lhs - rhs) <= 1e-6 * float((wt * jd).abs().sum())
    # freestream preservation
    Uf = np.zeros((E, V, nS))
    Uf[:, 0] = 1.0
    Uf[:, 1] = 0.1
    Uf[:, 2] = -0.05
    Uf[:, V - 1] = 1.0 / 0.4 + 0.5 * (0.1 ** 2 + 0.05 ** 2)
    z.set_state(Uf, layout="device")
    z.compute_residual(0)
    assert np.abs(z.state("divF", 0, layout="device")).max() < 1e-11


def test_c3_hex_p3_64_200_steps_conservation_and_decay():
    """The 3D Taylor-Green configuration as specified: 64^3 hexes, p=3, Re 1600, Ma 0.1,
    Pr 0.71, Roe + LDG, RK4 dt = 2e-4, 200 steps.  No oracle at this size; checks:
      * mass, x-momentum and energy integrals (sum of w detJ U) stay constant to rounding
        (DFR on Gauss-Legendre points is discretely conservative on a periodic mesh);
      * the kinetic energy decreases monotonically at the viscous rate of the initial field:
        for u = U0 (sin x cos y cos z, -cos x sin y cos z, 0) on the 2 pi box,
        dK/dt / K = -mu int |grad u|^2 / (rho0 int |u|^2 / 2) = -6 mu / rho0 (incompressible
        limit; Ma 0.1 adds O(Ma^2) pressure work), to 5%."""
    import torch
    from zefr_amd.autograd import device_field
    print("[full] c3 200-step setup", flush=True)
    mesh = make_mesh("c3_hex64.msh", (64, 64, 64), (0.0,) * 3, (TWO_PI,) * 3)
    deck = write_deck(tmp("c3_roe.deck"), mesh, 3, 3, True, dt=2e-4, ic_type=1, test_case=4,
                      extra={"fconv_type": "Roe"})
    z = product(deck)
    inp = z.get_input()
    mu, rho0 = inp.mu, inp.rho_fs
    w = torch.from_numpy(z.host_array("weights_spts")).cuda()
    detJ = device_field(z, "jaco_det_spts")[:, 0, :]  # [E][nS]

    def integrals():
        U = z.state_tensor()  # [E][V][nS]
        wd = w[None, :] * detJ
        tot = (U * wd[:, None, :]).sum(dim=(0, 2))
        scale = (U.abs() * wd[:, None, :]).sum(dim=(0, 2))
        ke = (0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / U[:, 0] * wd).sum()
        return tot.cpu().numpy(), scale.cpu().numpy(), float(ke)

    tot0, scale0, ke0 = integrals()
    kes = [ke0]
    for step in range(200):
        z.do_step()
        if (step + 1) % 20 == 0:
            tot, scale, ke = integrals()
            kes.append(ke)
            assert (np.abs(tot - tot0)[[0, 1, 4]] <= 1e-13 * scale0[[0, 1, 4]]).all(), (step, tot - tot0, scale0)
    print(f"[full] c3 200 steps: K {ke0:.12e} -> {kes[-1]:.12e}", flush=True)
    assert all(b < a for a, b in zip(kes, kes[1:])), kes
    t = 200 * inp.dt
    rate = (kes[-1] - ke0) / (t * ke0)
    want = -6.0 * mu / rho0
    print(f"[full] c3 decay rate {rate:.4e} vs {want:.4e}", flush=True)
    assert abs(rate - want) <= 0.05 * abs(want), (rate, want)


def _tgv_torch(X, Y, Z, Ma, gamma=1.4, P0=1.0, rho0=1.0, R=286.9, drift=(0.02, 0.015, 0.01)):
    import torch
    U0 = Ma * (gamma * P0 / rho0) ** 0.5
    u = U0 * torch.sin(X) * torch.cos(Y) * torch.cos(Z) + drift[0]
    v = -U0 * torch.cos(X) * torch.sin(Y) * torch.cos(Z) + drift[1]
    w = torch.zeros_like(u) + drift[2]
    p = P0 + rho0 * U0 * U0 / 16.0 * (torch.cos(2 * X) + torch.cos(2 * Y)) * (torch.cos(2 * Z) + 2)
    rho = p / (R * (P0 / (rho0 * R)))
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return torch.stack([rho, rho * u, rho * v, rho * w, E], dim=1)


def test_c5_adjoint_32_p3_50_steps_checkpointed():
    """d(kinetic energy after 50 RK4 steps)/d(U0, Ma) on 32^3 hexes, p=3.  The difference
    step is 1e-5 here: J sums 10.5M terms, and with 1e-7 its rounding (~1e-16 J) alone
    is 1e-6 of the directional derivative."""
    import torch
    from zefr_amd.autograd import checkpointed_rollout, kinetic_energy, rollout
    print("[full] c5 setup", flush=True)
    mesh = make_mesh("c5_hex32.msh", (32, 32, 32), (0.0,) * 3, (TWO_PI,) * 3)
    deck = write_deck(tmp("c5.deck"), mesh, 3, 3, True, dt=2e-4)
    z = product(deck)
    x = torch.from_numpy(np.ascontiguousarray(
        z.host_array("coord_spts").reshape(z.dims()["nSpts"], 3, z.dims()["nEles"]).transpose(2, 1, 0))).cuda()
    X, Y, Zc = x[:, 0], x[:, 1], x[:, 2]
    dt, steps, Re = 2e-4, 50, 1600.0
    Ma = torch.tensor(0.1, dtype=torch.float64, device="cuda", requires_grad=True)
    U0 = _tgv_torch(X, Y, Zc, Ma).contiguous()
    U0.retain_grad()
    mu = Ma * 1.4 ** 0.5 / Re
    J = kinetic_energy(z, checkpointed_rollout(z, U0, dt, steps, every=10, mu=mu))
    J.backward()
    g = U0.grad.detach()
    rng = np.random.default_rng(5)
    with torch.no_grad():
        Ub = U0.detach()
        for _ in range(2):
            d = torch.from_numpy(rng.standard_normal(tuple(Ub.shape))).cuda()
            Jp = kinetic_energy(z, rollout(z, Ub + 1e-5 * d, dt, steps, mu=mu.detach())).item()
            Jm = kinetic_energy(z, rollout(z, Ub - 1e-5 * d, dt, steps, mu=mu.detach())).item()
            fd = (Jp - Jm) / 2e-5
            assert abs(fd - float((g * d).sum())) <= 1e-6 * abs(fd)

        def J_of(m):
            m = torch.tensor(m, dtype=torch.float64, device="cuda")
            U = _tgv_torch(X, Y, Zc, m).contiguous()
            return kinetic_energy(z, rollout(z, U, dt, steps, mu=m * 1.4 ** 0.5 / Re)).item()
        fdm = (J_of(0.1 + 1e-5) - J_of(0.1 - 1e-5)) / 2e-5
    assert abs(fdm - Ma.grad.item()) <= 1e-6 * abs(fdm)
