"""Optics helpers shared by the Abbe, TCC and SOCS entry points: source points
on the device, the oi_optics record, shifted-pupil kernels and grid sizes."""
from __future__ import annotations

import collections
import ctypes
import math
import threading
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as nat
from . import fingerprint


def imaging_mode(numerics) -> bool:
    """True for vector imaging.  Only the lower-case mode strings work in the
    reference (CalculateAerialImage.py:50-53,68; CalculateTCCMatrix.py:39,52)."""
    mode = numerics.ImageCalculationMode
    if mode == "vector":
        return True
    if mode == "scalar":
        return False
    raise ValueError("ImageCalculationMode must be 'scalar' or 'vector'")


def focus_values(projector):
    fr = projector.FocusRange
    if isinstance(fr, torch.Tensor):
        return [float(v) for v in fr.reshape(-1).tolist()]
    return [float(v) for v in np.asarray(fr).reshape(-1)]


@dataclass
class SourceSet:
    """Lit source points on the device (float64 holding the float32 values the
    reference selects) with per-point polarisation.  The extents and the
    total weight are kept as host numbers, taken before the upload, so that
    window sizing and normalisation need no device-to-host transfer."""
    X: torch.Tensor
    Y: torch.Tensor
    V: torch.Tensor
    PX: torch.Tensor  # c128
    PY: torch.Tensor
    extent: tuple = None       # (xmin, xmax, ymin, ymax)
    total: float = None        # sum of V
    _scales: dict = field(default_factory=dict, repr=False, compare=False)

    @property
    def n(self) -> int:
        return int(self.V.numel())

    def kernel_scale(self, index_image: float) -> torch.Tensor:
        """sqrt(n V_s / W) per point, the weight of each source point's
        coherent kernel (CalculateAerialImage.py:89-92 folded into the
        kernels); kept per image-side index.  Read only."""
        key = float(index_image)
        sc = self._scales.get(key)
        if sc is None:
            sc = self._scales[key] = torch.sqrt(key * self.V / self.weight)
        return sc

    @property
    def weight(self) -> float:
        return float(self.V.sum()) if self.total is None else self.total

    def bounds(self):
        if self.extent is None:
            return float(self.X.min()), float(self.X.max()), float(self.Y.min()), float(self.Y.max())
        return self.extent


# Source sets by the source's parameters (fingerprint.digest) and device:
# Calc_SourceSimple and the upload run once per distinct source, and a hit
# restores the state the call leaves on the source.  The cached device
# tensors are shared: callers read them and never modify them in place.
_SOURCE_CACHE = collections.OrderedDict()
_SOURCE_CACHE_SIZE = 8
_SOURCE_LOCK = threading.Lock()


def source_set(source) -> SourceSet:
    """Calc_SourceSimple (Source.py:47-54) plus polarisation, moved to the
    device.  Leaves ``source.source_data`` populated as the reference does.
    Repeated calls with unchanged source parameters reuse the device set."""
    dev = nat.device()
    key = (fingerprint.digest(source, exclude=("source_data",)), dev.index)
    with _SOURCE_LOCK:
        hit = _SOURCE_CACHE.get(key)
        if hit is not None:
            _SOURCE_CACHE.move_to_end(key)
    if hit is not None:
        holder, X, Y, V, sset = hit
        # the state Calc_SourceSimple leaves: source.source_data is the
        # SourceData record (class attributes, Source.py:293-295) holding the
        # lit points
        holder.X, holder.Y, holder.Value = X.clone(), Y.clone(), V.clone()
        source.source_data = holder
        return sset
    sset, sd = _source_set_uncached(source, dev)
    # cache only the reference's record form: the SourceData class itself,
    # which every Source shares, so a hit hands no object of another source
    if isinstance(sd, type):
        entry = (sd, sd.X.clone(), sd.Y.clone(), sd.Value.clone(), sset)
        with _SOURCE_LOCK:
            _SOURCE_CACHE[key] = entry
            while len(_SOURCE_CACHE) > _SOURCE_CACHE_SIZE:
                _SOURCE_CACHE.popitem(last=False)
    return sset


def _source_set_uncached(source, dev):
    # host-side selection: the drop-in leaves source.source_data on the host
    # as the reference does (Source.Calc_SourceSimple(device=...) gives the
    # same lit set on the device for callers that keep it there)
    sd = source.Calc_SourceSimple()
    X = sd.X.reshape(-1).to(torch.float64)
    Y = sd.Y.reshape(-1).to(torch.float64)
    V = sd.Value.reshape(-1).to(torch.float64)
    if V.numel() == 0:
        raise ValueError("the source has no lit points")
    if hasattr(source, "polarization_f64"):
        PX, PY = source.polarization_f64(X, Y)
    else:
        th = torch.atan2(Y, X)
        rho = torch.sqrt(X * X + Y * Y)
        PX, PY = source.Calc_PolarizationMap(th, rho)
        PX = torch.as_tensor(PX).to(torch.complex128)
        PY = torch.as_tensor(PY).to(torch.complex128)
    extent = (float(X.min()), float(X.max()), float(Y.min()), float(Y.max()))
    sset = SourceSet(X.to(dev), Y.to(dev), V.to(dev), PX.to(dev).contiguous(),
                     PY.to(dev).contiguous(), extent, float(V.sum()))
    return sset, source.source_data


def optics_record(source, projector, focus, pitch_r, pitch_c, vector) -> nat.Optics:
    o = nat.Optics()
    o.wavelength = float(source.Wavelength)
    o.na = float(projector.NA)
    o.magnification = float(projector.Magnification)
    o.index_image = float(projector.IndexImage)
    o.focus = float(focus)
    coeffs = projector.zernike_coefficients(0.0) if hasattr(projector, "zernike_coefficients") \
        else [float(v) for v in torch.as_tensor(projector.Aberration_Zernike, dtype=torch.float64).reshape(-1)]
    if len(coeffs) != 37:
        raise ValueError("Aberration_Zernike must hold 37 Fringe coefficients")
    for i, c in enumerate(coeffs):
        o.zernike[i] = c
    o.pitch_r = float(pitch_r)
    o.pitch_c = float(pitch_c)
    o.vector = 1 if vector else 0
    return o


def pupil_kernels(src: SourceSet, opt: nat.Optics, scale: torch.Tensor, k0r, wr, k0c, wc):
    """G[e][s][i][j] (c64): scale[s] x pupil x polarisation at kappa =
    (k0r+i, k0c+j) for every lit source point (oi_pupil_kernels)."""
    dev = src.V.device
    ncomp = 3 if opt.vector else 1
    G = torch.empty((ncomp, src.n, wr, wc), dtype=torch.complex64, device=dev)
    scale = scale.to(device=dev, dtype=torch.float64).contiguous()
    px = torch.view_as_real(src.PX).contiguous() if opt.vector else None
    py = torch.view_as_real(src.PY).contiguous() if opt.vector else None
    nat.check(nat.load().oi_pupil_kernels(ctypes.byref(opt), nat.ptr(src.X), nat.ptr(src.Y),
                                          nat.ptr(scale), nat.ptr(px), nat.ptr(py), src.n, k0r, wr,
                                          k0c, wc, nat.ptr(G), nat.stream()))
    return G


def pupil_kernels_foci(src: SourceSet, opt: nat.Optics, foci, scale: torch.Tensor, k0r, wr, k0c,
                       wc):
    """G[f][e][s][i][j] (c64): pupil_kernels at each focus of `foci` (opt.focus
    is not read), the focus-independent factors evaluated once
    (oi_pupil_kernels_foci)."""
    dev = src.V.device
    ncomp = 3 if opt.vector else 1
    zs = [float(z) for z in foci]
    G = torch.empty((len(zs), ncomp, src.n, wr, wc), dtype=torch.complex64, device=dev)
    scale = scale.to(device=dev, dtype=torch.float64).contiguous()
    px = torch.view_as_real(src.PX).contiguous() if opt.vector else None
    py = torch.view_as_real(src.PY).contiguous() if opt.vector else None
    zarr = (ctypes.c_double * len(zs))(*zs)
    nat.check(nat.load().oi_pupil_kernels_foci(ctypes.byref(opt), zarr, len(zs), nat.ptr(src.X),
                                               nat.ptr(src.Y), nat.ptr(scale), nat.ptr(px),
                                               nat.ptr(py), src.n, k0r, wr, k0c, wc, nat.ptr(G),
                                               nat.stream()))
    return G


def tcc_halfwidths(xypitch, NA, wavelength, numerics=None):
    """(Nf, Ng) of the TCC grid: ceil(2 p) decided in float32 as
    CalculateTCCMatrix.py:17-20 does, unless numerics.Hopkins_GridHalfWidth
    is set."""
    hw = getattr(numerics, "Hopkins_GridHalfWidth", None) if numerics is not None else None
    if hw is not None:
        return int(hw), int(hw)
    px = np.float32(xypitch[0] / (wavelength / NA))
    py = np.float32(xypitch[1] / (wavelength / NA))
    return int(math.ceil(np.float32(2) * px)), int(math.ceil(np.float32(2) * py))


def abbe_window(src: SourceSet, p_r: float, p_c: float, ny: int, nx: int):
    """Spectrum window covering every shifted pupil |kappa/p + s| <= 1, clipped
    to the grid's centred index range, and the per-pupil support extents."""
    xmin, xmax, ymin, ymax = src.bounds()
    k0r = math.floor(p_r * (-1.0 - xmax)) - 1
    k1r = math.ceil(p_r * (1.0 - xmin)) + 1
    k0c = math.floor(p_c * (-1.0 - ymax)) - 1
    k1c = math.ceil(p_c * (1.0 - ymin)) + 1
    k0r = max(k0r, -(ny // 2))
    k1r = min(k1r, ny - 1 - ny // 2)
    k0c = max(k0c, -(nx // 2))
    k1c = min(k1c, nx - 1 - nx // 2)
    band_r = min(int(math.floor(2 * p_r)) + 1, k1r - k0r + 1)
    band_c = min(int(math.floor(2 * p_c)) + 1, k1c - k0c + 1)
    return k0r, k1r - k0r + 1, k0c, k1c - k0c + 1, band_r, band_c
