"""Roe's flux on the CPU (the reference has no Roe flux, so nothing of its own pins it).

The oracle's textbook restatement (oracle/zefr_oracle.cpp: roe_flux) is checked against an
independent construction: |A~| of the Euler flux Jacobian (taken by complex-step
differentiation) at the Roe-averaged state, formed by Sylvester's formula over its
eigenvalues un - a, un, un + a (checked against numpy's eigvals), with Harten's entropy fix
(|lambda| < delta = 0.1 a~ -> (lambda^2 + delta^2) / (2 delta)) on the two acoustic
eigenvalues.  The product's flux (physics.hpp through zfr_host_common_flux) is checked
against the oracle and against the textbook properties: consistency F(U, U) = F(U).n, pure
upwinding for supersonic normal flow (Roe's property A~ dU = dF), symmetry, rotational
invariance.  Planted mistakes (no entropy fix; no dissipation on the entropy and shear
waves) are rejected.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle.oracle import roe_flux as oracle_roe

GAMMA = 1.4


def _cons(rho, vel, p):
    vel = np.asarray(vel, dtype=complex if np.iscomplexobj(vel) else float)
    return np.concatenate([[rho], rho * vel, [p / (GAMMA - 1) + 0.5 * rho * np.dot(vel, vel)]])


def _flux_n(U, n):
    rho = U[0]
    vel = U[1:-1] / rho
    p = (GAMMA - 1) * (U[-1] - 0.5 * rho * np.dot(vel, vel))
    un = np.dot(vel, n)
    return np.concatenate([[rho * un], rho * vel * un + p * n, [(U[-1] + p) * un]])


def _jacobian(U, n):
    V = U.size
    J = np.zeros((V, V))
    h = 1e-30
    for j in range(V):
        Uc = U.astype(complex)
        Uc[j] += 1j * h
        J[:, j] = _flux_n(Uc, n).imag / h
    return J


def _numpy_roe(UL, UR, n, fix=True, drop_linear=False):
    """Roe flux from the Jacobian: |A~| by Sylvester's formula over the three distinct
    eigenvalues un - a, un, un + a of the (diagonalisable) Jacobian at the Roe state."""
    rl, rr = np.sqrt(UL[0]), np.sqrt(UR[0])
    uL, uR = UL[1:-1] / UL[0], UR[1:-1] / UR[0]
    pL = (GAMMA - 1) * (UL[-1] - 0.5 * UL[0] * uL @ uL)
    pR = (GAMMA - 1) * (UR[-1] - 0.5 * UR[0] * uR @ uR)
    HL, HR = (UL[-1] + pL) / UL[0], (UR[-1] + pR) / UR[0]
    u = (rl * uL + rr * uR) / (rl + rr)
    H = (rl * HL + rr * HR) / (rl + rr)
    rho = rl * rr
    E = (rho * H + (GAMMA - 1) * 0.5 * rho * u @ u) / GAMMA  # the state with velocity u, enthalpy H
    Ut = np.concatenate([[rho], rho * u, [E]])
    A = _jacobian(Ut, n)
    a = np.sqrt((GAMMA - 1) * (H - 0.5 * u @ u))
    un = u @ n
    mu = np.array([un - a, un, un + a])
    ev = np.sort(np.linalg.eigvals(A).real)
    assert abs(ev[0] - mu[0]) < 1e-9 * (1 + a) and abs(ev[-1] - mu[2]) < 1e-9 * (1 + a)
    delta = 0.1 * a
    f = np.abs(mu)
    if fix:
        for k in (0, 2):
            if f[k] < delta:
                f[k] = (mu[k] ** 2 + delta ** 2) / (2 * delta)
    if drop_linear:  # a plausible mistake: no dissipation on the entropy and shear waves
        f[1] = 0.0
    I = np.eye(A.shape[0])
    absA = np.zeros_like(A)
    for k in range(3):
        P = I.copy()
        for j in range(3):
            if j != k:
                P = P @ (A - mu[j] * I) / (mu[k] - mu[j])
        absA += f[k] * P
    return 0.5 * (_flux_n(UL, n) + _flux_n(UR, n)) - 0.5 * absA @ (UR - UL)


def _states(D, rng, mach=0.5, jump=0.2):
    n = rng.standard_normal(D)
    n /= np.linalg.norm(n)
    v = rng.standard_normal(D)
    v *= mach * np.sqrt(GAMMA) / np.linalg.norm(v)
    UL = _cons(1.0, v, 1.0)
    UR = _cons(1.0 + jump * rng.uniform(-1, 1), v + jump * rng.standard_normal(D), 1.0 + jump * rng.uniform(-1, 1))
    return UL, UR, n


def _product_roe(D, UL, UR, n):
    from zefr_amd import _lib
    phys = _lib.ZfrPhys(gamma=GAMMA, prandtl=0.72, mu=0.0, rt=0.0, c_sth=0.0, fix_vis=1, rus_k=0.0, ldg_b=0.5,
                        ldg_tau=0.1, fconv_type=2)
    F = np.zeros(D + 2)
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))
    ULc, URc, nc = (np.ascontiguousarray(x, np.float64) for x in (UL, UR, n))
    _lib.check(_lib.lib().zfr_host_common_flux(D, C.byref(phys), 0, dp(ULc), dp(URc), None, None, dp(nc), 1.0,
                                               F.ctypes.data_as(C.POINTER(C.c_double))))
    return F


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("mach", [0.05, 0.5, 0.98, 1.02, 2.5])
def test_oracle_roe_matches_eigen_decomposition(D, mach):
    rng = np.random.default_rng(int(mach * 100) + D)
    for _ in range(20):
        UL, UR, n = _states(D, rng, mach)
        ref = _numpy_roe(UL, UR, n)
        got = oracle_roe(D, UL, UR, n)
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_entropy_fix_is_active_and_checked():
    """At a sonic state the fix changes the flux; the numpy check without the fix, or without
    the entropy and shear waves' dissipation, then disagrees with the oracle."""
    rng = np.random.default_rng(9)
    D = 2
    n = np.array([1.0, 0.0])
    a = np.sqrt(GAMMA)
    UL = _cons(1.0, [0.999 * a, 0.3], 1.0)
    UR = _cons(1.02, [1.001 * a, 0.1], 1.03)
    got = oracle_roe(D, UL, UR, n)
    assert np.abs(got - _numpy_roe(UL, UR, n)).max() <= 1e-12 * np.abs(got).max()
    assert np.abs(got - _numpy_roe(UL, UR, n, fix=False)).max() > 1e-6
    UL, UR, n = _states(D, rng, 0.4)
    got = oracle_roe(D, UL, UR, n)
    assert np.abs(got - _numpy_roe(UL, UR, n, drop_linear=True)).max() > 1e-6


@pytest.mark.parametrize("D", [2, 3])
def test_product_roe_matches_oracle(D):
    rng = np.random.default_rng(20 + D)
    for mach in (0.1, 0.9, 1.0, 3.0):
        for _ in range(10):
            UL, UR, n = _states(D, rng, mach)
            ref = oracle_roe(D, UL, UR, n)
            assert np.abs(_product_roe(D, UL, UR, n) - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("D", [2, 3])
def test_product_roe_textbook_properties(D):
    rng = np.random.default_rng(40 + D)
    for _ in range(10):
        UL, UR, n = _states(D, rng, 0.6)
        # consistency
        F = _product_roe(D, UL, UL, n)
        assert np.abs(F - _flux_n(UL, n)).max() <= 1e-14 * np.abs(F).max()
        # symmetry: F(UL, UR, n) = -F(UR, UL, -n)
        assert np.abs(_product_roe(D, UL, UR, n) + _product_roe(D, UR, UL, -n)).max() <= 1e-13 * np.abs(F).max()
    # pure upwinding for supersonic normal flow in either direction
    for sgn in (1.0, -1.0):
        for _ in range(10):
            n = rng.standard_normal(D)
            n /= np.linalg.norm(n)
            v = sgn * 3.0 * np.sqrt(GAMMA) * n + 0.2 * rng.standard_normal(D)
            UL = _cons(1.0, v, 1.0)
            UR = _cons(1.1, v * 1.05, 1.08)
            F = _product_roe(D, UL, UR, n)
            up = _flux_n(UL if sgn > 0 else UR, n)
            assert np.abs(F - up).max() <= 1e-12 * np.abs(up).max()
    # rotational invariance in the plane of the first two axes
    th = 0.7
    Rm = np.eye(D)
    Rm[:2, :2] = [[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]
    UL, UR, n = _states(D, rng, 0.8)
    rot = lambda U: np.concatenate([[U[0]], Rm @ U[1:-1], [U[-1]]])
    F = _product_roe(D, UL, UR, n)
    Fr = _product_roe(D, rot(UL), rot(UR), Rm @ n)
    assert np.abs(Fr - rot(F)).max() <= 1e-13 * np.abs(F).max()
