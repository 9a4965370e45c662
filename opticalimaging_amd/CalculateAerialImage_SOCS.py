"""SOCS aerial image on the HIP device (CalculateAerialImage_SOCS.py:3-11).

I = fftshift(sum_k sum_c |df dg fft2(phi_k * ifftshift(Hk_c))|^2) is evaluated
as I[j] = sum_k sum_c |sum_kappa phi_k(kappa) spec_c(kappa) exp(-2 pi i kappa.j/N)|^2
with spec the windowed spectrum (all shifts and the p_x p_y df dg factors
cancel), on a coarse grid followed by exact Fourier interpolation
(``oi_socs_image``).
"""
from __future__ import annotations

import math

import torch

from . import _native as nat
from . import engine
from .Scatter import _pitches, pupil_image

__all__ = ["CalculateAerialImage_SOCS", "CalculateSOCS", "kernel_set_of", "socs_field_image"]

def kernel_set_of(kernels) -> engine.KernelSet:
    """Windowed kernel set for dense SOCS kernels (Ny, Nx, K): the one
    attached by DecomposeTCC_SOCS while the tensor is unmodified, else
    extracted from the exact support (engine.kernel_set_of)."""
    return engine.kernel_set_of(kernels)


def socs_field_image(field, ks: engine.KernelSet) -> torch.Tensor:
    """SOCS intensity (Ny, Nx) float32 on the device for one ScatterField."""
    tens, kind, cin, live = engine.field_encoding(field)
    return engine.socs_image(tens, ks, kind=kind, cin=cin, chans=live)[0]


def CalculateAerialImage_SOCS(scatter, TCCMatrix_SOCS, source, projector):
    """Returns (intensity (Ny, Nx) float32, PupilImage), on the ScatterField's
    device."""
    h = torch.as_tensor(scatter.ScatterField)
    out_dev = h.device
    ks = kernel_set_of(TCCMatrix_SOCS)
    if tuple(h.shape[:2]) != (ks.ny, ks.nx):
        raise ValueError("ScatterField grid does not match the SOCS kernel grid")
    intensity = socs_field_image(h, ks)
    px, py = _pitches(scatter, projector, source)
    pupil = pupil_image(h, px, py)
    return intensity.to(out_dev), pupil.to(out_dev)


def CalculateSOCS(FloquetMode, TCCMatrix_SOCS, numerics):
    """SOCS image from Floquet modes (CalculateAerialImage_SOCS.py:13-32):
    I[x, y] = sum_{c,k} |sum_o exp(2 pi i (ox (ix + 1/2)/Gx + oy (iy + 1/2)/Gy))
    FM[o, c] phi_k[o]|^2 on the ScatterGrid_X x ScatterGrid_Y grid, orders
    o in [-ScatterOrder_X, ScatterOrder_X] x [-ScatterOrder_Y, ScatterOrder_Y].

    With kappa = -o this is the coherent sum of oi_coherent_image on a
    Gx x Gy grid: spectrum window FM (orders reversed, times the half-pixel
    phase exp(-pi i (kappa_x/Gx + kappa_y/Gy))), kernels phi_k(-kappa), so it
    runs on the band-limited coarse-grid kernels instead of the direct
    (grid x orders x kernels) sum.  Returns float32 (Gx, Gy) on the device of
    FloquetMode (CPU in, CPU out)."""
    gx, gy = int(numerics.ScatterGrid_X), int(numerics.ScatterGrid_Y)
    ox, oy = int(numerics.ScatterOrder_X), int(numerics.ScatterOrder_Y)
    n0, n1 = 2 * ox + 1, 2 * oy + 1
    fm_in = torch.as_tensor(FloquetMode)
    out_dev = fm_in.device
    dev = nat.device()
    fm = fm_in.to(device=dev, dtype=torch.complex64).reshape(-1, 3)
    kern = torch.as_tensor(TCCMatrix_SOCS).to(device=dev, dtype=torch.complex64)
    if kern.dim() != 3:
        raise ValueError("SOCS kernels must be a (Ny, Nx, K) tensor")
    K = int(kern.shape[2])
    if fm.shape[0] != n0 * n1 or kern.shape[0] * kern.shape[1] != n0 * n1:
        # the reference's reshape to the order count fails the same way
        raise RuntimeError("FloquetMode / kernel sizes do not match the scatter orders")
    if gx < n0 or gy < n1:
        raise RuntimeError("ScatterGrid must hold at least 2*ScatterOrder+1 points per axis")
    kc = torch.fft.ifftshift(kern.transpose(0, 1), dim=(0, 1)).reshape(n0, n1, K)
    fmc = fm.reshape(n0, n1, 3)
    # window index r <-> kappa = -o_max + r <-> order o = o_max - r (both axes reversed)
    kwin = kc.flip(0, 1).permute(2, 0, 1).contiguous()              # [K][n0][n1]
    kx = torch.arange(-ox, ox + 1, device=dev, dtype=torch.float64)
    ky = torch.arange(-oy, oy + 1, device=dev, dtype=torch.float64)
    ph = torch.exp(-1j * math.pi * (kx[:, None] / gx + ky[None, :] / gy)).to(torch.complex64)
    spec = fmc.flip(0, 1).permute(2, 0, 1)                            # [3][n0][n1]
    live = [c for c in range(3) if bool((spec[c] != 0).any())] or [0]
    spec = (spec[live] * ph[None]).unsqueeze(0).contiguous()          # [1][C][n0][n1]
    plan = engine.PLANS.get(ny=gx, nx=gy, k0r=-ox, wr=n0, k0c=-oy, wc=n1, nchan=len(live),
                            max_batch=1, band_r=0, band_c=0, max_kernels=K)
    out = torch.empty((1, gx, gy), dtype=torch.float32, device=dev)
    nat.check(nat.load().oi_coherent_image(plan.handle, nat.ptr(spec), 1, nat.ptr(kwin), K,
                                           nat.ptr(out), nat.stream()))
    return out[0].to(out_dev)
