"""Differentiable SOCS imaging for mask optimisation (OPC / inverse lithography).

``aerial_image(mask, kernels)`` is the SOCS intensity of
``CalculateAerialImage_SOCS`` (``CalculateAerialImage_SOCS.py:3-11``) for a
batch of real masks [B, Ny, Nx] (channel 0 of the reference's ScatterField),
as a ``torch.autograd.Function``: the forward pass runs ``oi_socs_image`` and
the backward pass the hand-written adjoint ``oi_socs_adjoint`` (conjugate-
kernel correlation of the image cotangent with per-kernel fields recomputed on
the coarse grid).  The reference has no gradient path of its own beyond torch
autograd through its complex64 ops.

``ilt`` runs plain gradient steps of the pixel-based ILT objective
||I(sigmoid(theta)) - T||^2.
"""
from __future__ import annotations

import torch

from . import _native as nat
from . import engine

__all__ = ["SOCSImage", "aerial_image", "kernel_set", "ilt_loss", "ilt"]


def kernel_set(kernels) -> engine.KernelSet:
    """KernelSet from DecomposeTCC_SOCS output (dense (Ny, Nx, K)) or a KernelSet."""
    return engine.kernel_set_of(kernels)


class SOCSImage(torch.autograd.Function):
    """I = SOCS image of real masks; backward = oi_socs_adjoint."""

    @staticmethod
    def forward(ctx, mask: torch.Tensor, ks: engine.KernelSet) -> torch.Tensor:
        m = mask.detach().to(torch.float32).contiguous()
        B = int(m.shape[0])
        plan = engine.plan_for(ks, 1, B)
        out = engine.socs_image(m, ks, plan=plan)
        ctx.save_for_backward(m)
        ctx.ks = ks
        ctx.in_dtype = mask.dtype
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (m,) = ctx.saved_tensors
        g = engine.socs_adjoint(m, grad_out.detach().to(torch.float32).contiguous(), ctx.ks)
        return g.to(ctx.in_dtype), None


def aerial_image(mask: torch.Tensor, kernels) -> torch.Tensor:
    """Differentiable SOCS intensity of masks [B, Ny, Nx] (or [Ny, Nx]) on the device."""
    ks = kernel_set(kernels)
    squeeze = mask.dim() == 2
    m = mask.unsqueeze(0) if squeeze else mask
    if m.dim() != 3:
        raise ValueError("mask must be (Ny, Nx) or (B, Ny, Nx)")
    out = SOCSImage.apply(m, ks)
    return out[0] if squeeze else out


def ilt_loss(theta: torch.Tensor, target: torch.Tensor, kernels) -> torch.Tensor:
    """||I(sigmoid(theta)) - target||^2 summed over the batch."""
    img = aerial_image(torch.sigmoid(theta), kernels)
    return ((img - target) ** 2).sum()


def ilt(theta: torch.Tensor, target: torch.Tensor, kernels, steps: int = 50, lr: float = 0.5):
    """Normalised gradient descent on theta [B, Ny, Nx] (each image's step is
    lr times its gradient divided by the gradient's RMS); returns (theta,
    per-step losses tensor on the device).

    The gradient is the one ``ilt_loss`` gives through autograd, written out.
    Where the plan has the band form, each step is one library call
    (oi_ilt_band_step): the image is band-limited, so the spectrum of the
    cotangent 2 (I - T) is 2 (band of I - band of T) and the loss follows by
    Parseval from the band and a per-target constant (the targets' band and
    constant are computed once, oi_ilt_target_band); the image is not formed,
    and m = sigmoid(theta) is taken as theta is loaded, never stored.  The
    call then applies dL/dtheta = dL/dm m (1 - m), its per-image RMS and the
    update of theta.  Elsewhere a step is oi_socs_residual_adjoint (direct
    plans) or oi_socs_image, oi_ilt_residual and oi_socs_adjoint (plans
    without the fused image pass), then oi_ilt_update.  The loop stays on the
    device (no host synchronisation per step)."""
    ks = kernel_set(kernels)
    theta = theta.detach().to(torch.float32).contiguous().clone()
    if theta.dim() != 3:
        raise ValueError("theta must be (B, Ny, Nx)")
    target = target.detach().to(device=theta.device, dtype=torch.float32).contiguous()
    if target.shape != theta.shape:
        raise ValueError("target must have theta's shape")
    B = int(theta.shape[0])
    losses = torch.zeros(steps, dtype=torch.float32, device=theta.device)
    if B == 0:
        return theta, losses
    npix = int(theta[0].numel())
    plan = engine.plan_for(ks, 1, B)
    lib = nat.load()
    g = torch.empty_like(theta)
    awork = engine.adjoint_work(ks, B, plan)
    tband = engine.ilt_target_band(target, ks, awork, plan)
    if tband is not None:
        # the whole step from theta, the image band and the targets' band
        for i in range(steps):
            engine.ilt_band_step(theta, ks, tband, lr, g, awork, losses[i:i + 1], plan=plan)
        return theta, losses
    # d = 2 (I - T), the loss and g = dL/dm: in the pass that produces I, with
    # the mask spectrum shared with the adjoint, where the plan has that pass;
    # otherwise the image, the residual pass and the adjoint one after another
    work = torch.empty(lib.oi_ilt_work_floats(B), dtype=torch.float32, device=theta.device)
    img = None                     # the image itself, only where the residual is not fused
    d = torch.empty_like(theta)
    m = torch.sigmoid(theta)
    st = nat.stream()
    fused = True
    for i in range(steps):
        if fused:
            fused = engine.socs_residual_adjoint(m, ks, target, d, g, awork, losses[i:i + 1],
                                                 plan=plan)
        if not fused:
            if img is None:
                img = torch.empty_like(theta)
            engine.socs_image(m, ks, out=img, plan=plan)
            nat.check(lib.oi_ilt_residual(nat.ptr(img), nat.ptr(target), nat.ptr(d),
                                          nat.ptr(losses[i:i + 1]), B, npix, nat.ptr(work), st))
            engine.socs_adjoint(m, d, ks, out=g, plan=plan)
        nat.check(lib.oi_ilt_update(nat.ptr(theta), nat.ptr(g), nat.ptr(m), float(lr), B, npix,
                                    nat.ptr(work), nat.ptr(m), st))
    return theta, losses
