"""Differentiable CCSD residual and energy (``torch.autograd.Function`` wrappers).

``residual(cc, t1, t2, F=None)`` returns ``(r1, r2) = cc.residuals(F, t1, t2)`` as autograd outputs.
Its forward-mode derivative (``jvp``) is the CCSD Jacobian on HBAR (``pycc/ccderiv.py:149-180``). Its
backward is the vector-Jacobian product kernel set of the library: the Lambda-equation
contractions of ``pycc/cclambda.py:345-369`` (r_L1) and ``:468-496`` (r_L2) without their source
terms, evaluated at l1 = c1, l2 = c2 + c2^T for the cotangent (c1, c2), on an HBAR
(``pycc/cchbar.py:128-823``) built at the forward's amplitudes. The t2 gradient is pair-symmetrised:
t2_ijab and t2_jiba are one parameter of the CCSD equations, so for a pair-symmetric perturbation
dt2 the directional derivative is sum(g2 * dt2).

``energy(cc, t1, t2, F=None)`` returns ``cc.cc_energy`` as a 0-dim tensor; its gradient is
(2 Hov, L), the source terms of the Lambda equations (cclambda.py:348, :468). With both,
``E + sum(l1 r1) + 1/2 sum(l2 r2)`` is the CCSD Lagrangian and its gradient is the Lambda residual.

When ``F`` requires grad, both Functions also return dF (and the residual's forward mode takes an F
tangent: the residual is affine in F at fixed amplitudes): the one-particle density of
``pycc/ccdensity.py:335-470`` at l = (c1, c2 + c2^T) minus its l = 0 part for the residual, at l = 0
times the incoming gradient for the energy (``CCwfn.integral_vjp`` / ``CCwfn.energy_densities``). The
two-electron integrals live in the context, not in a tensor argument: their gradient is the
two-particle density returned by ``CCwfn.integral_vjp``.
"""
from __future__ import annotations

import torch


class CCSDResidual(torch.autograd.Function):
    @staticmethod
    def forward(ctx, t1, t2, cc, F, store_vvvv):
        r1, r2 = cc.residuals(F.detach(), t1.detach(), t2.detach())
        ctx.cc = cc
        ctx.F = F
        ctx.store_vvvv = store_vvvv
        ctx.save_for_backward(t1, t2)
        ctx.save_for_forward(t1, t2)
        return r1, r2

    @staticmethod
    def jvp(ctx, dt1, dt2, _cc, dF, _store):
        t1, t2 = ctx.saved_tensors
        t1, t2 = t1.detach(), t2.detach()
        dt1 = torch.zeros_like(t1) if dt1 is None else dt1.contiguous()
        dt2 = torch.zeros_like(t2) if dt2 is None else dt2.contiguous()
        j1, j2 = ctx.cc.jvp(t1, t2, dt1, dt2, F=ctx.F.detach(), store_vvvv=ctx.store_vvvv)
        if dF is not None:
            # the residual is affine in F at fixed amplitudes: dr = r(dF, t) - r(0, t)
            a1, a2 = ctx.cc.residuals(dF.detach().to(torch.float64).contiguous(), t1, t2)
            b1, b2 = ctx.cc.residuals(torch.zeros_like(ctx.F, dtype=torch.float64), t1, t2)
            j1, j2 = j1 + (a1 - b1), j2 + (a2 - b2)
        return j1, j2

    @staticmethod
    def backward(ctx, c1, c2):
        t1, t2 = ctx.saved_tensors
        cc = ctx.cc
        c1 = torch.zeros_like(t1) if c1 is None else c1.contiguous()
        c2 = torch.zeros_like(t2) if c2 is None else c2.contiguous()
        F = ctx.F.detach()
        g1, g2 = cc.vjp(t1.detach(), t2.detach(), c1, c2, F=F, store_vvvv=ctx.store_vvvv)
        gF = None
        if ctx.needs_input_grad[3]:
            gF = cc.integral_vjp(t1.detach(), t2.detach(), c1, c2, fock_only=True)["D"].to(ctx.F.dtype)
        return g1, g2, None, gF, None


class CCSDEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, t1, t2, cc, F):
        e = cc.cc_energy(F=F.detach(), t1=t1.detach(), t2=t2.detach())
        ctx.cc = cc
        ctx.F = F
        ctx.save_for_backward(t1, t2)
        return torch.tensor(e, dtype=torch.float64, device=t1.device)

    @staticmethod
    def backward(ctx, ge):
        t1, t2 = ctx.saved_tensors
        d1, d2 = ctx.cc.energy_gradient(t1.detach(), t2.detach(), F=ctx.F.detach())
        gF = None
        if ctx.needs_input_grad[3]:
            gF = ctx.cc.energy_densities(t1.detach(), t2.detach(), fock_only=True)["D"] * ge
        return d1 * ge, d2 * ge, None, gF


def residual(cc, t1, t2, F=None, store_vvvv=False):
    """(r1, r2) of ``cc.residuals(F, t1, t2)`` with the Lambda-equation vJp as backward."""
    return CCSDResidual.apply(t1, t2, cc, cc.H.F if F is None else F, bool(store_vvvv))


def energy(cc, t1, t2, F=None):
    """``cc.cc_energy`` at (t1, t2) with gradient (2 Hov, L)."""
    return CCSDEnergy.apply(t1, t2, cc, cc.H.F if F is None else F)


def lagrangian(cc, t1, t2, l1, l2, F=None, store_vvvv=False):
    """E(t) + sum(l1 r1(t)) + 1/2 sum(l2 r2(t)); its gradient at fixed l is the Lambda residual."""
    r1, r2 = residual(cc, t1, t2, F, store_vvvv)
    return energy(cc, t1, t2, F) + (l1 * r1).sum() + 0.5 * (l2 * r2).sum()
