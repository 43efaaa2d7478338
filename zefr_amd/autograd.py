"""PyTorch autograd for the DFR residual: gradients flow through R(U) by the library's
hand-written vector-Jacobian product (zefr_residual_vjp), and through the RK44 stages
that combine residuals by torch itself.

    from zefr_amd import zefr
    from zefr_amd.autograd import residual, rk4_step, kinetic_energy
    z = zefr.Zefr(); z.read_input(deck); z.setup_solver()
    U = z.state_tensor().requires_grad_()       # [nEles][nVars][nSpts] float64 on the device
    R = residual(z, U)                          # = divF, differentiable in U
    J = kinetic_energy(z, rk4_step(z, U, dt))   # any scalar; J.backward() fills U.grad

The time derivative the reference integrates is dU/dt = -divF / detJ
(src/solver.cpp:2897-2941); rk4_step applies the classical RK44 stages of
src/solver.cpp:713-735 to it.  `checkpointed_rollout` recomputes each block of steps in
the backward pass (torch.utils.checkpoint), so memory holds one state per block.
"""

from __future__ import annotations

import numpy as np
import torch

__all__ = ["Residual", "residual", "rk4_step", "rollout", "checkpointed_rollout", "kinetic_energy",
           "device_field"]


class Residual(torch.autograd.Function):
    """divF = R(U) with backward v = (dR/dU)^T g computed on the device."""

    @staticmethod
    def forward(ctx, U, z):
        ctx.z = z
        ctx.save_for_backward(U)
        return z.residual_at(U.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        (U,) = ctx.saved_tensors
        v = ctx.z.residual_vjp(U.detach().contiguous(), g.contiguous())
        return v, None


class ResidualMu(torch.autograd.Function):
    """divF = R(U; mu) differentiable in U and in the viscosity coefficient mu (a 0-d
    tensor).  R is affine in mu for constant viscosity and for Sutherland's law (both
    scale the stress and heat flux by mu; the LDG penalty does not depend on it), so
    dR/dmu = R(U; 1) - R(U; 0) exactly."""

    @staticmethod
    def forward(ctx, U, mu, z):
        ctx.z = z
        ctx.save_for_backward(U, mu)
        with _viscosity(z, float(mu)):
            return z.residual_at(U.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        U, mu = ctx.saved_tensors
        Uc = U.detach().contiguous()
        with _viscosity(ctx.z, float(mu)):
            v = ctx.z.residual_vjp(Uc, g.contiguous())
        with _viscosity(ctx.z, 1.0):
            r1 = ctx.z.residual_at(Uc)
        with _viscosity(ctx.z, 0.0):
            r0 = ctx.z.residual_at(Uc)
        dmu = (g * (r1 - r0)).sum().reshape(mu.shape)
        return v, dmu, None


class _viscosity:
    """Temporarily set the solver's viscosity coefficient (InputStruct.mu)."""

    def __init__(self, z, mu):
        self.z, self.mu = z, mu

    def __enter__(self):
        inp = self.z.get_input()
        self.saved = inp.mu
        inp.mu = self.mu

    def __exit__(self, *exc):
        self.z.get_input().mu = self.saved


def residual(z, U, mu=None):
    """R(U); with a tensor `mu`, R(U; mu) differentiable in mu as well."""
    if mu is None:
        return Residual.apply(U, z)
    return ResidualMu.apply(U, mu, z)


def device_field(z, name):
    """A host geometry array of the solver as a device tensor in the device layout:
    'jaco_det_spts' -> [nEles][1][nSpts]; 'weights_spts' -> [1][1][nSpts]."""
    d = z.dims()
    E, nS = d["nEles"], d["nSpts"]
    a = z.host_array(name)
    if name == "jaco_det_spts":
        t = a.reshape(nS, E).T.reshape(E, 1, nS)
    elif name == "weights_spts":
        t = a.reshape(1, 1, nS)
    else:
        raise ValueError(name)
    return torch.from_numpy(np.ascontiguousarray(t)).to("cuda")


def rk4_step(z, U, dt, detJ=None, mu=None):
    """One classical RK4 step of dU/dt = -R(U)/detJ (src/solver.cpp:2865-3097)."""
    if detJ is None:
        detJ = device_field(z, "jaco_det_spts")
    f = lambda X: -residual(z, X, mu) / detJ  # noqa: E731
    k1 = f(U)
    k2 = f(U + 0.5 * dt * k1)
    k3 = f(U + 0.5 * dt * k2)
    k4 = f(U + dt * k3)
    return U + dt * (k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6)


def rollout(z, U, dt, steps, detJ=None, mu=None):
    detJ = device_field(z, "jaco_det_spts") if detJ is None else detJ
    for _ in range(steps):
        U = rk4_step(z, U, dt, detJ, mu)
    return U


def checkpointed_rollout(z, U, dt, steps, every=10, detJ=None, mu=None):
    """`steps` RK4 steps, storing only the state at the start of each block of `every`
    steps; the backward pass recomputes the block's stages from it."""
    from torch.utils.checkpoint import checkpoint
    detJ = device_field(z, "jaco_det_spts") if detJ is None else detJ
    done = 0
    while done < steps:
        n = min(every, steps - done)
        if mu is None:
            U = checkpoint(lambda X, n=n: rollout(z, X, dt, n, detJ), U, use_reentrant=False)
        else:
            U = checkpoint(lambda X, m, n=n: rollout(z, X, dt, n, detJ, m), U, mu, use_reentrant=False)
        done += n
    return U


def kinetic_energy(z, U, detJ=None, weights=None):
    """J = sum_e sum_spt w_spt detJ |m|^2 / (2 rho): the volume integral of the kinetic
    energy by the solution points' quadrature."""
    detJ = device_field(z, "jaco_det_spts") if detJ is None else detJ
    w = device_field(z, "weights_spts") if weights is None else weights
    D = U.shape[1] - 2
    rho = U[:, 0:1, :]
    m2 = (U[:, 1:1 + D, :] ** 2).sum(dim=1, keepdim=True)
    return (w * detJ * 0.5 * m2 / rho).sum()
