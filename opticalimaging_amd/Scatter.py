"""Mask / near-field container, attribute-compatible with the reference's
``Scatter`` (Scatter.py:3-8).  ``CalculateSpectrum`` runs on the HIP device."""
from __future__ import annotations

import collections
import ctypes
import threading

import torch

from . import _native as nat
from . import engine


def _pitches(scatter, po, sr):
    na, lam = float(po.NA), float(sr.Wavelength)
    return scatter.Period_X / (lam / na), scatter.Period_Y / (lam / na)


def frequency_axes(nx, ny, px, py):
    """f (columns) and g (rows) normalised frequency axes (Scatter.py:18-19)."""
    f = (1 / px) * torch.arange(-(nx // 2), (nx + 1) // 2)
    g = (1 / py) * torch.arange(-(ny // 2), (ny + 1) // 2)
    return f, g


def full_spectrum(field: torch.Tensor, px: float, py: float) -> torch.Tensor:
    """Hk = px py fftshift(ifft2(ifftshift(h))) over dims (0, 1) of a
    (Ny, Nx, C) field, as c64 on the device (Scatter.py:14-23)."""
    dev = nat.device()
    h = torch.as_tensor(field).to(dev)
    if h.dim() != 3:
        raise ValueError("ScatterField must be (Ny, Nx, C)")
    ny, nx, cin = h.shape
    h = h.to(torch.complex64).contiguous()
    chans = list(range(cin))
    cptr, keep = nat.int_array(chans)
    work = torch.empty((cin, ny, nx), dtype=torch.complex64, device=dev)
    hk = torch.empty((ny, nx, cin), dtype=torch.complex64, device=dev)
    nat.check(nat.load().oi_spectrum_full(nat.ptr(h), nat.OI_FIELD_C64_CHLAST, cin, cptr, cin, 1,
                                          ny, nx, ctypes.c_float(px * py), nat.ptr(work),
                                          nat.ptr(hk), nat.stream()))
    del keep
    return hk


# Per-grid constants of the windowed pupil image (the fftshift phase of the
# crop and its unit-circle mask), kept on the device for the last few grids so
# that a repeated call launches no table set-up and copies nothing to the
# device.  The cached tensors are read, never written in place.
_PUPIL_TABLES = collections.OrderedDict()
_PUPIL_TABLES_SIZE = 16
_PUPIL_LOCK = threading.Lock()


def _pupil_tables(ny, nx, px, py, dev):
    key = (ny, nx, float(px), float(py), dev.index)
    with _PUPIL_LOCK:
        hit = _PUPIL_TABLES.get(key)
        if hit is not None:
            _PUPIL_TABLES.move_to_end(key)
            return hit
    ox, oy = int(px) + 3, int(py) + 3
    f, g = frequency_axes(nx, ny, px, py)
    kr = torch.arange(-oy, oy, device=dev, dtype=torch.float64)
    kc = torch.arange(-ox, ox, device=dev, dtype=torch.float64)
    # exp(-2 pi i kappa (N//2)/N): the fftshift phase of Scatter.py:22
    ph = torch.exp(-2j * torch.pi * (kr[:, None] * (ny // 2) / ny + kc[None, :] * (nx // 2) / nx))
    gm, fm = torch.meshgrid(g[ny // 2 - oy: ny // 2 + oy].to(dev),
                            f[nx // 2 - ox: nx // 2 + ox].to(dev), indexing="ij")
    entry = (ph, (gm ** 2 + fm ** 2) > 1)
    with _PUPIL_LOCK:
        _PUPIL_TABLES[key] = entry
        while len(_PUPIL_TABLES) > _PUPIL_TABLES_SIZE:
            _PUPIL_TABLES.popitem(last=False)
    return entry


def pupil_image(field: torch.Tensor, px: float, py: float, hk: torch.Tensor | None = None,
                encoding=None):
    """PupilImage of Scatter.CalculateSpectrum (Scatter.py:25-35): Hk with NaN
    outside the unit circle, cropped to +-(int(p)+3) samples about the centre.

    When the crop lies inside the grid it is evaluated from the windowed
    spectrum alone (a pruned transform); otherwise from the full spectrum with
    the reference's slicing."""
    dev = nat.device()
    h = torch.as_tensor(field).to(dev)
    ny, nx, cin = h.shape
    ox, oy = int(px) + 3, int(py) + 3
    inside = (ny // 2 - oy >= 0 and nx // 2 - ox >= 0 and oy <= ny - ny // 2 and ox <= nx - nx // 2)
    if inside and hk is None:
        enc = encoding if encoding is not None else engine.field_encoding(h)
        tens, kind, cin_, live = enc
        plan = engine.PLANS.get(ny=ny, nx=nx, k0r=-oy, wr=2 * oy, k0c=-ox, wc=2 * ox,
                                nchan=len(live), max_batch=1, max_kernels=1)
        spec = torch.empty((1, len(live), 2 * oy, 2 * ox), dtype=torch.complex64, device=dev)
        cptr, keep = nat.int_array(live)
        nat.check(nat.load().oi_spectrum_window(plan.handle, nat.ptr(tens), kind, cin_, cptr, 1,
                                                nat.ptr(spec), nat.stream()))
        del keep
        # Hk = px py exp(-2 pi i kappa (N//2)/N) spec
        ph, outside = _pupil_tables(ny, nx, px, py, dev)
        win = torch.zeros((2 * oy, 2 * ox, cin), dtype=torch.complex64, device=dev)
        vals = (spec[0].to(torch.complex128) * ph[None] * (px * py)).to(torch.complex64)
        for i, c in enumerate(enc.source_channels):
            win[:, :, c] = vals[i]
        return win.masked_fill_(outside[..., None], float("nan"))   # no host round trip
    f, g = frequency_axes(nx, ny, px, py)
    if hk is None:
        hk = full_spectrum(h, px, py)
    win = hk[(ny // 2) - oy:(ny // 2) + oy, (nx // 2) - ox:(nx // 2) + ox].clone()
    gsel = g[(ny // 2) - oy:(ny // 2) + oy].to(dev)
    fsel = f[(nx // 2) - ox:(nx // 2) + ox].to(dev)
    gm, fm = torch.meshgrid(gsel, fsel, indexing="ij")
    outside = (gm ** 2 + fm ** 2) > 1
    return win.masked_fill_(outside[..., None], float("nan"))   # no host round trip


class Scatter:
    """Mask / near-field parameters (Scatter.py:3-8)."""

    def __init__(self):
        self.Period_X = 1000  # nanometer
        self.Period_Y = 1000  # nanometer
        self.ScatterField = []
        self.FloquetMode = []

    def CalculateSpectrum(self, po, sr):
        """Returns (Hk, f, g, PupilImage) as the reference (Scatter.py:10-36);
        Hk and PupilImage are computed on the HIP device and returned on the
        ScatterField's device."""
        h = torch.as_tensor(self.ScatterField)
        out_dev = h.device
        ny, nx, _ = h.shape
        px, py = _pitches(self, po, sr)
        f, g = frequency_axes(nx, ny, px, py)
        hk = full_spectrum(h, px, py)
        pupil = pupil_image(h, px, py, hk=hk)
        return hk.to(out_dev), f, g, pupil.to(out_dev)
