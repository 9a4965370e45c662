"""Float64 torch restatement of the SOCS image for reference autograd.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).  Same operations as
``CalculateAerialImage_SOCS`` (``CalculateAerialImage_SOCS.py:3-11``) with the
spectrum of ``Scatter.CalculateSpectrum`` (``Scatter.py:20-23``), in
complex128, for one real mask in channel 0; torch autograd through it gives the
float64 mask gradient the hand-written adjoint is checked against.  Pinned by
``tests/test_oracle_pins.py`` against the numpy oracle's image and against
central finite differences of that image.
"""
from __future__ import annotations

import numpy as np
import torch


def socs_image_t(mask: torch.Tensor, kernels: torch.Tensor, period_x: float, period_y: float,
                 NA: float, wavelength: float) -> torch.Tensor:
    """mask: float64 (Ny, Nx); kernels: complex128 (Ny, Nx, K).  Returns I (Ny, Nx)."""
    ny, nx = mask.shape
    px = period_x / (wavelength / NA)
    py = period_y / (wavelength / NA)
    h = mask.to(torch.complex128)
    hk = torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(h)), dim=(0, 1)) * (px * py)
    sp = torch.fft.ifftshift(hk, dim=(0, 1))
    scale = (1.0 / px) * (1.0 / py)
    E = scale * torch.fft.fft2(kernels * sp[:, :, None], dim=(0, 1))
    I = (E.real ** 2 + E.imag ** 2).sum(dim=2)
    return torch.fft.fftshift(I, dim=(0, 1))


def socs_mask_gradient(mask: np.ndarray, grad_image: np.ndarray, kernels: np.ndarray,
                       period_x: float, period_y: float, NA: float, wavelength: float) -> np.ndarray:
    """d/d mask of sum(grad_image * I(mask)) by float64 torch autograd."""
    m = torch.tensor(np.asarray(mask, dtype=np.float64), requires_grad=True)
    k = torch.as_tensor(np.asarray(kernels).astype(np.complex128))
    g = torch.as_tensor(np.asarray(grad_image, dtype=np.float64))
    with torch.enable_grad():
        I = socs_image_t(m, k, period_x, period_y, NA, wavelength)
        (grad,) = torch.autograd.grad((g * I).sum(), m)
    return grad.numpy()
