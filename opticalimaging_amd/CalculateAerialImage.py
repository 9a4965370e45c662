"""Aerial-image entry points on the HIP device (CalculateAerialImage.py).

``CalculateAbbeImage``: partially coherent Abbe source-point summation
(CalculateAerialImage.py:18-103).  Each lit source point s and polarisation
component e is one coherent kernel sqrt(n V_s / W) P_se(kappa) on the
spectrum window (``oi_pupil_kernels``, float64, rho^2 <= 1 decided in
float64); the image is the coherent sum over kernels and field channels
(``oi_coherent_image``), evaluated on a coarse grid sized by one shifted
pupil (2 floor(2 p) + 1) and Fourier-interpolated exactly to the image grid.

``CalculateHopkinsImage``: per focus, TCC -> SOCS kernels -> SOCS image
(CalculateAerialImage.py:105-130).

``CalculateOptimized``: per focus, TCC -> SOCS kernels on the Floquet-order
grid -> ``CalculateSOCS`` on the ScatterGrid (CalculateAerialImage.py:131-152).
"""
from __future__ import annotations

import torch

from . import _native as nat
from . import engine, optics
from .CalculateAerialImage_SOCS import CalculateAerialImage_SOCS, CalculateSOCS
from .CalculateTCCMatrix import CalculateTCCMatrix
from .DecomposeTCC_SOCS import DecomposeTCC_SOCS
from .Scatter import _pitches, pupil_image

__all__ = ["ImageData", "cartesian_to_polar", "CalculateAbbeImage", "CalculateHopkinsImage",
           "CalculateOptimized", "abbe_images"]


class ImageData:
    """Result record (CalculateAerialImage.py:7-11)."""
    Intensity: torch.Tensor = None
    ImageX: torch.Tensor = None
    ImageY: torch.Tensor = None
    ImageZ: torch.Tensor = None


def cartesian_to_polar(x, y):
    """CalculateAerialImage.py:13-16."""
    return torch.sqrt(x ** 2 + y ** 2), torch.atan2(y, x)


def abbe_images(field, source, scatter, projector, numerics, encoding=None) -> torch.Tensor:
    """Abbe intensities (nFocus, Ny, Nx) float32 on the device.  ``encoding``
    is engine.field_encoding(field) when the caller already has it."""
    vector = optics.imaging_mode(numerics)
    dev = nat.device()
    h = torch.as_tensor(field)
    ny, nx = int(h.shape[0]), int(h.shape[1])
    px, py = _pitches(scatter, projector, source)
    src = optics.source_set(source)
    k0r, wr, k0c, wc, br, bc = optics.abbe_window(src, py, px, ny, nx)
    tens, kind, cin, live = encoding if encoding is not None else engine.field_encoding(h)
    ncomp = 3 if vector else 1
    plan = engine.PLANS.get(ny=ny, nx=nx, k0r=k0r, wr=wr, k0c=k0c, wc=wc, nchan=len(live),
                            max_batch=1, band_r=br, band_c=bc, max_kernels=ncomp * src.n)
    lib = nat.load()
    spec = torch.empty((1, len(live), wr, wc), dtype=torch.complex64, device=dev)
    cptr, keep = nat.int_array(live)
    nat.check(lib.oi_spectrum_window(plan.handle, nat.ptr(tens), kind, cin, cptr, 1, nat.ptr(spec),
                                     nat.stream()))
    del keep
    scale = src.kernel_scale(projector.IndexImage)
    foci = optics.focus_values(projector)
    out = torch.empty((len(foci), ny, nx), dtype=torch.float32, device=dev)
    opt = optics.optics_record(source, projector, 0.0, py, px, vector)
    # every focus's kernels from one pass, in slices of at most 1 GiB
    per_focus = ncomp * src.n * wr * wc * 8
    step = max(1, min(len(foci), (1 << 30) // max(per_focus, 1)))
    for i0 in range(0, len(foci), step):
        G = optics.pupil_kernels_foci(src, opt, foci[i0:i0 + step], scale, k0r, wr, k0c, wc)
        for k in range(G.shape[0]):
            nat.check(lib.oi_coherent_image(plan.handle, nat.ptr(spec), 1, nat.ptr(G[k]),
                                            G.shape[1] * G.shape[2], nat.ptr(out[i0 + k]),
                                            nat.stream()))
    return out


def _image_axes(scatter, nx, ny, projector):
    d = ImageData()
    d.ImageX = torch.linspace(-scatter.Period_X / 2, scatter.Period_X / 2, nx)
    d.ImageY = torch.linspace(-scatter.Period_Y / 2, scatter.Period_Y / 2, ny)
    d.ImageZ = projector.FocusRange
    return d


def CalculateAbbeImage(source, scatter, projector, numerics):
    """Returns (ImageData, PupilImage); Intensity is (nFocus, Ny, Nx) float32
    on the ScatterField's device (CalculateAerialImage.py:18-103)."""
    h = torch.as_tensor(scatter.ScatterField)
    ny, nx = int(h.shape[0]), int(h.shape[1])
    enc = engine.field_encoding(h)          # one device-to-host transfer per call
    intensity = abbe_images(h, source, scatter, projector, numerics, encoding=enc)
    px, py = _pitches(scatter, projector, source)
    pupil = pupil_image(h, px, py, encoding=enc)
    d = _image_axes(scatter, nx, ny, projector)
    d.Intensity = intensity.to(h.device)
    return d, pupil.to(h.device)


def CalculateHopkinsImage(source, scatter, projector, numerics):
    """Returns (ImageData, PupilImage) (CalculateAerialImage.py:105-130)."""
    h = torch.as_tensor(scatter.ScatterField)
    ny, nx = int(h.shape[0]), int(h.shape[1])
    pitchxy = [scatter.Period_Y, scatter.Period_X]
    nfg = [nx, ny]
    foci = optics.focus_values(projector)
    intensity = torch.zeros((len(foci), ny, nx))
    pupil = None
    for i, z in enumerate(foci):
        tcc, fg = CalculateTCCMatrix(source, pitchxy, projector, z, numerics)
        kern = DecomposeTCC_SOCS(tcc, fg, nfg, numerics)
        img, pupil = CalculateAerialImage_SOCS(scatter, kern, source, projector)
        intensity[i] = img.cpu()
        del tcc
    d = _image_axes(scatter, nx, ny, projector)
    d.Intensity = intensity.to(h.device)
    return d, pupil


def CalculateOptimized(source, scatter, projector, numerics):
    """Returns (ImageData, empty tensor) (CalculateAerialImage.py:131-152).

    scatter.FloquetMode: complex (2 ScatterOrder_X + 1, 2 ScatterOrder_Y + 1, 3)
    Floquet amplitudes of the near field; the kernels of each focus live on
    that order grid (DecomposeTCC_SOCS with Nfg = FloquetMode.shape[:2])."""
    fm = torch.as_tensor(scatter.FloquetMode)
    pitchxy = [scatter.Period_Y, scatter.Period_X]
    nfg = [int(fm.shape[0]), int(fm.shape[1])]
    foci = optics.focus_values(projector)
    gx, gy = int(numerics.ScatterGrid_X), int(numerics.ScatterGrid_Y)
    intensity = torch.zeros((len(foci), gx, gy))
    for i, z in enumerate(foci):
        tcc, fg = CalculateTCCMatrix(source, pitchxy, projector, z, numerics)
        kern = DecomposeTCC_SOCS(tcc, fg, nfg, numerics)
        intensity[i] = CalculateSOCS(fm, kern, numerics).cpu()
        del tcc
    d = ImageData()
    d.Intensity = intensity.to(fm.device)
    d.ImageX = torch.linspace(-scatter.Period_X / 2, scatter.Period_X / 2, gx)
    d.ImageY = torch.linspace(-scatter.Period_Y / 2, scatter.Period_Y / 2, gy)
    d.ImageZ = projector.FocusRange
    return d, torch.tensor([])
