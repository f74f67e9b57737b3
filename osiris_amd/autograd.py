"""torch.autograd front end of the flash: differentiable phase split.

``flash(ctx, T, P, z)`` returns (x, y, L) of SuccFalsh (compsim/
Init_Calcs.cpp:571-738) computed by the HIP flash; backward calls the
adjoint kernel (osr_flash_adjoint_device) and forward-mode products call the
tangent kernel (osr_flash_tangent_device).  Both linearise the converged
state by the implicit-function theorem on the fugacity equations; gradients
flow to P and z (T is a parameter).
"""
from __future__ import annotations

import torch

from . import runtime as rt


class FlashFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, octx, T, mode, P, z):
        out = rt.flash(octx, T, P.detach(), z.detach(), mode=mode)
        ctx.P_dtype, ctx.P_device, ctx.z_dtype, ctx.z_device = P.dtype, P.device, z.dtype, z.device
        ctx.octx = octx
        ctx.T = T
        ctx.save_for_backward(P.detach(), z.detach(), out["K"], out["label"])
        ctx.mark_non_differentiable(out["label"])
        return out["x"], out["y"], out["L"], out["label"]

    @staticmethod
    def backward(ctx, xbar, ybar, Lbar, _lab):
        P, z, K, label = ctx.saved_tensors
        Pbar, zbar = rt.flash_adjoint(ctx.octx, ctx.T, P, z, K, label,
                                      xbar=None if xbar is None else xbar.contiguous(),
                                      ybar=None if ybar is None else ybar.contiguous(),
                                      Lbar=None if Lbar is None else Lbar.contiguous())
        # the kernels work in float64 on the GPU: hand the gradients back in
        # the dtype and on the device of the inputs
        return None, None, None, Pbar.to(dtype=ctx.P_dtype, device=ctx.P_device), \
            zbar.to(dtype=ctx.z_dtype, device=ctx.z_device)


def flash(octx, T, P, z, mode="persistent"):
    """Differentiable flash: returns x, y [nc, n], L [n], label [n]."""
    return FlashFn.apply(octx, T, mode, P, z)


def tangent(octx, T, P, z, dP=None, dz=None, mode="persistent"):
    """Forward-mode product: the flash outputs and their tangents
    (x, y, L, label), (dx, dy, dL) for the perturbation (dP, dz)."""
    out = rt.flash(octx, T, P, z, mode=mode)
    dx, dy, dL = rt.flash_tangent(octx, T, P, z, out["K"], out["label"], dP=dP, dz=dz)
    return (out["x"], out["y"], out["L"], out["label"]), (dx, dy, dL)
