"""Pointwise physics of the kernels, built for the host from the same templates
(zefr_amd/csrc/physics.hpp via libzefr_amd.so's zfr_host_* entry points), on the CPU:

* the forward fluxes against the oracle's point flux (a restatement of
  include/flux.hpp) and against a textbook Rusanov + LDG common flux;
* the hand-written reverse-mode derivatives (vector-Jacobian products) against
  fourth-order central finite differences of the forward functions, and the dot-product
  identity <Fb, J d> = <J^T Fb, d>.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle.oracle import make_phys, point_flux
from zefr_amd import _lib

DP = C.POINTER(C.c_double)


def dp(a):
    return a.ctypes.data_as(DP)


def phys(gamma=1.4, prandtl=0.72, mu=2e-2, rt=0.7, c_sth=0.3, fix_vis=1, rus_k=0.0, ldg_b=0.5, ldg_tau=0.1,
         fconv_type=0):
    return _lib.ZfrPhys(gamma, prandtl, mu, rt, c_sth, fix_vis, rus_k, ldg_b, ldg_tau, fconv_type)


def state(D, rng):
    rho = 0.8 + 0.5 * rng.random()
    u = 0.4 * rng.standard_normal(D)
    p = 0.7 + 0.6 * rng.random()
    E = p / 0.4 + 0.5 * rho * u @ u
    return np.concatenate([[rho], rho * u, [E]])


def host_flux(D, a, U, dU=None):
    F = np.zeros((D + 2, D))
    _lib.check(_lib.lib().zfr_host_point_flux(D, C.byref(a), dp(U), dp(dU) if dU is not None else None, dp(F)))
    return F


def host_flux_vjp(D, a, U, dU, Fb):
    Ub = np.zeros(D + 2)
    dUb = np.zeros((D + 2, D))
    _lib.check(_lib.lib().zfr_host_point_flux_vjp(D, C.byref(a), dp(U), dp(dU) if dU is not None else None,
                                                  dp(np.ascontiguousarray(Fb)), dp(Ub), dp(dUb)))
    return Ub, dUb


def host_common(D, a, visc, UL, UR, dUL, dUR, n, dAs):
    F = np.zeros(D + 2)
    _lib.check(_lib.lib().zfr_host_common_flux(D, C.byref(a), int(visc), dp(UL), dp(UR), dp(dUL), dp(dUR),
                                               dp(n), dAs, dp(F)))
    return F


def host_common_vjp(D, a, visc, UL, UR, dUL, dUR, n, dAs, Fb):
    V = D + 2
    ULb, URb = np.zeros(V), np.zeros(V)
    dULb, dURb = np.zeros((V, D)), np.zeros((V, D))
    _lib.check(_lib.lib().zfr_host_common_flux_vjp(D, C.byref(a), int(visc), dp(UL), dp(UR), dp(dUL), dp(dUR),
                                                   dp(n), dAs, dp(np.ascontiguousarray(Fb)), dp(ULb), dp(URb),
                                                   dp(dULb), dp(dURb)))
    return ULb, URb, dULb, dURb


def fd4(f, x, d, h):
    """Fourth-order central difference of f at x along d."""
    return (-f(x + 2 * h * d) + 8 * f(x + h * d) - 8 * f(x - h * d) + f(x - 2 * h * d)) / (12 * h)


def oracle_phys(a, viscous=True):
    return make_phys(viscous=viscous, gamma=a.gamma, prandtl=a.prandtl, mu=a.mu, rt=a.rt, c_sth=a.c_sth,
                     fix_vis=bool(a.fix_vis), rus_k=a.rus_k, ldg_b=a.ldg_b, ldg_tau=a.ldg_tau)


# ---------------------------------------------------------------- forward
@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("fix_vis", [1, 0])
def test_point_flux_matches_oracle(D, fix_vis):
    rng = np.random.default_rng(D + 10 * fix_vis)
    a = phys(fix_vis=fix_vis)
    for _ in range(10):
        U = state(D, rng)
        dU = rng.standard_normal((D + 2, D))
        np.testing.assert_allclose(host_flux(D, a, U), point_flux(D, oracle_phys(a, False), U),
                                   rtol=1e-14, atol=1e-15)
        # Sutherland: r * sqrt(r) here, pow(r, 1.5) in the oracle -> an ulp or two
        np.testing.assert_allclose(host_flux(D, a, U, dU), point_flux(D, oracle_phys(a), U, dU),
                                   rtol=1e-13, atol=1e-14)


def _textbook_common(D, a, visc, UL, UR, dUL, dUR, n, dAs):
    """src/faces.cpp:1244-1364 (Rusanov) and 1476-1614 (LDG) from the oracle's point flux."""
    FL = point_flux(D, oracle_phys(a, False), UL) @ n
    FR = point_flux(D, oracle_phys(a, False), UR) @ n

    def lam(U):
        rho = U[0]
        u = U[1:1 + D] / rho
        p = (a.gamma - 1) * (U[-1] - 0.5 * rho * u @ u)
        return abs(u @ n) + np.sqrt(a.gamma * p / rho)
    eig = max(lam(UL), lam(UR))
    F = 0.5 * (FL + FR) - 0.5 * eig * (1 - a.rus_k) * (UR - UL)
    if visc:
        GL = (point_flux(D, oracle_phys(a), UL, dUL) - point_flux(D, oracle_phys(a, False), UL)) @ n
        GR = (point_flux(D, oracle_phys(a), UR, dUR) - point_flux(D, oracle_phys(a, False), UR)) @ n
        F = F + 0.5 * (GL + GR) + a.ldg_tau * (UL - UR) + a.ldg_b * (GL - GR)
    return F * dAs


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("visc,beta", [(False, 0.5), (True, 0.5), (True, -0.5), (True, 0.2)])
def test_common_flux_textbook(D, visc, beta):
    rng = np.random.default_rng(7 + D)
    a = phys(ldg_b=beta, rus_k=0.2)
    for _ in range(8):
        UL, UR = state(D, rng), state(D, rng)
        dUL, dUR = rng.standard_normal((D + 2, D)), rng.standard_normal((D + 2, D))
        n = rng.standard_normal(D)
        n /= np.linalg.norm(n)
        dAs = 0.3 + rng.random()
        got = host_common(D, a, visc, UL, UR, dUL, dUR, n, dAs)
        want = _textbook_common(D, a, visc, UL, UR, dUL, dUR, n, dAs)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)


# ---------------------------------------------------------------- reverse mode
@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("viscous,fix_vis", [(False, 1), (True, 1), (True, 0)])
def test_point_flux_vjp_matches_finite_differences(D, viscous, fix_vis):
    rng = np.random.default_rng(3 * D + fix_vis)
    a = phys(fix_vis=fix_vis)
    V = D + 2
    for _ in range(6):
        U = state(D, rng)
        dU = rng.standard_normal((V, D)) if viscous else None
        Fb = rng.standard_normal((V, D))
        Ub, dUb = host_flux_vjp(D, a, U, dU, Fb)
        for k in range(V):  # column by column in U
            e = np.zeros(V)
            e[k] = 1.0
            col = fd4(lambda x: host_flux(D, a, x, dU), U, e, 1e-3)
            assert abs(np.sum(Fb * col) - Ub[k]) < 1e-9 * max(1.0, abs(Ub[k]))
        if viscous:
            for k in range(V * D):
                e = np.zeros(V * D)
                e[k] = 1.0
                col = fd4(lambda x: host_flux(D, a, U, x.reshape(V, D)), dU.ravel(), e, 1e-3)
                assert abs(np.sum(Fb * col) - dUb.ravel()[k]) < 1e-9 * max(1.0, abs(dUb.ravel()[k]))


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("visc,beta,fix_vis,fconv", [(False, 0.5, 1, 0), (True, 0.5, 1, 0), (True, -0.5, 0, 0),
                                                     (True, 0.3, 1, 0), (False, 0.5, 1, 2), (True, 0.5, 1, 2),
                                                     (True, -0.5, 0, 2)])
def test_common_flux_vjp_matches_finite_differences(D, visc, beta, fix_vis, fconv):
    """Rusanov (fconv 0) and Roe (fconv 2, the reference has none) common fluxes."""
    rng = np.random.default_rng(11 * D + int(10 * beta) + fconv)
    a = phys(ldg_b=beta, fix_vis=fix_vis, rus_k=0.1, fconv_type=fconv)
    V = D + 2
    for trial in range(6):
        UL, UR = state(D, rng), state(D, rng)
        dUL, dUR = rng.standard_normal((V, D)), rng.standard_normal((V, D))
        n = rng.standard_normal(D)
        n /= np.linalg.norm(n)
        dAs = 0.5 + rng.random()
        Fb = rng.standard_normal(V)
        ULb, URb, dULb, dURb = host_common_vjp(D, a, visc, UL, UR, dUL, dUR, n, dAs, Fb)
        x0 = np.concatenate([UL, UR, dUL.ravel(), dUR.ravel()])

        def f(x):
            return host_common(D, a, visc, x[:V], x[V:2 * V], x[2 * V:2 * V + V * D].reshape(V, D),
                               x[2 * V + V * D:].reshape(V, D), n, dAs)
        grad = np.concatenate([ULb, URb, dULb.ravel(), dURb.ravel()])
        # dot-product identity along random directions, and every column
        for k in range(x0.size):
            e = np.zeros(x0.size)
            e[k] = 1.0
            col = fd4(f, x0, e, 1e-4)
            assert abs(Fb @ col - grad[k]) < 1e-8 * max(1.0, abs(grad[k])), (trial, k)
        d = rng.standard_normal(x0.size)
        assert abs(Fb @ fd4(f, x0, d, 1e-4) - grad @ d) < 1e-8 * max(1.0, abs(grad @ d))


@pytest.mark.parametrize("D", [2, 3])
def test_roe_vjp_inside_the_entropy_fix(D):
    """Transonic states put an acoustic speed inside Harten's fix (|lambda| < 0.1 a), where
    the flux is smooth but follows the parabola: the VJP still matches differences."""
    rng = np.random.default_rng(70 + D)
    a = phys(fconv_type=2)
    V = D + 2
    for _ in range(4):
        n = rng.standard_normal(D)
        n /= np.linalg.norm(n)
        c = np.sqrt(1.4)
        UL = np.concatenate([[1.0], (0.995 * c) * n, [1.0 / 0.4 + 0.5 * 0.995 ** 2 * 1.4]])
        UR = np.concatenate([[1.01], 1.01 * (1.002 * c) * n, [1.005 / 0.4 + 0.5 * 1.01 * 1.002 ** 2 * 1.4]])
        Fb = rng.standard_normal(V)
        z = np.zeros((V, D))
        ULb, URb, _, _ = host_common_vjp(D, a, False, UL, UR, z, z, n, 1.0, Fb)
        x0 = np.concatenate([UL, UR])

        def f(x):
            return host_common(D, a, False, x[:V], x[V:], z, z, n, 1.0)
        grad = np.concatenate([ULb, URb])
        for k in range(2 * V):
            e = np.zeros(2 * V)
            e[k] = 1.0
            col = fd4(f, x0, e, 1e-6)
            assert abs(Fb @ col - grad[k]) < 1e-7 * max(1.0, abs(grad[k])), k


def test_vjp_check_rejects_a_planted_sign_error():
    """The finite-difference check separates the derivative from a wrong one."""
    D, V = 3, 5
    rng = np.random.default_rng(5)
    a = phys()
    U = state(D, rng)
    dU = rng.standard_normal((V, D))
    Fb = rng.standard_normal((V, D))
    Ub, _ = host_flux_vjp(D, a, U, dU, Fb)
    e = np.zeros(V)
    e[0] = 1.0
    col = fd4(lambda x: host_flux(D, a, x, dU), U, e, 1e-3)
    assert abs(np.sum(Fb * col) - Ub[0]) < 1e-9 * max(1.0, abs(Ub[0]))
    assert abs(np.sum(Fb * col) + Ub[0]) > 1e-6  # a flipped sign would fail
