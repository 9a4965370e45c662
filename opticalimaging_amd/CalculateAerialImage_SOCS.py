"""SOCS aerial image on the HIP device (CalculateAerialImage_SOCS.py:3-11).

I = fftshift(sum_k sum_c |df dg fft2(phi_k * ifftshift(Hk_c))|^2) is evaluated
as I[j] = sum_k sum_c |sum_kappa phi_k(kappa) spec_c(kappa) exp(-2 pi i kappa.j/N)|^2
with spec the windowed spectrum (all shifts and the p_x p_y df dg factors
cancel), on a coarse grid followed by exact Fourier interpolation
(``oi_socs_image``).
"""
from __future__ import annotations

import torch

from . import _native as nat
from . import engine
from .Scatter import _pitches, pupil_image

__all__ = ["CalculateAerialImage_SOCS", "kernel_set_of", "socs_field_image"]

_cache = {}


def kernel_set_of(kernels) -> engine.KernelSet:
    """Windowed kernel set for dense SOCS kernels (Ny, Nx, K): the one
    attached by DecomposeTCC_SOCS, else extracted from the exact support (cached
    per tensor storage and version)."""
    if isinstance(kernels, engine.KernelSet):
        return kernels
    ks = getattr(kernels, "_oi_kernels", None)
    if ks is not None:
        return ks
    t = torch.as_tensor(kernels)
    key = (t.data_ptr(), t.device.type, tuple(t.shape), getattr(t, "_version", 0))
    ks = _cache.get(key)
    if ks is None:
        ks = engine.KernelSet.from_dense(t)
        if len(_cache) > 16:
            _cache.clear()
        _cache[key] = ks
    return ks


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
