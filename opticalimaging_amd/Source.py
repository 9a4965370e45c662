"""Illumination parameter object, attribute-compatible with the reference's
``Source`` (Source.py:15-45).

The source grid is host-side parameter data (at most PntNum**2 points).  Point
selection follows the reference's float32 torch arithmetic exactly, because
the lit set is decided by float32 comparisons: on the 21 x 21 annulus
0.5 <= r <= 0.8, eight grid points have r = 0.5 exactly in real arithmetic and
are lit or dark depending on float32 rounding (Source.py:65,234-237).

Extension: ``PolarizationType`` (None by default).  The reference's live
``Calc_PolarizationMap`` returns the constant ``PolarizationVector``
(Source.py:222-226) and keeps the per-type formulas only as a string
(Source.py:184-221); setting ``PolarizationType`` selects those formulas.
"""
from __future__ import annotations

import math
import sys

import torch
import torch.nn.functional as F

__all__ = ["Source", "SourceData", "PolarizationParameters", "conv2",
           "CalculateAnnularSourceMatrix", "CalculateQuasarSourceMatrix",
           "CalculateDipoleSourceMatrix", "CalculateMultiCircSourceMatrix",
           "ConvertSourceMatrix2SourceData"]

_POLARIZATIONS = ("x_pol", "y_pol", "c_pol", "d_pol", "r_pol", "t_pol", "line_pol")


def conv2(matr, ker, mode="same"):
    """2-D correlation with zero padding, 'same' output (Source.py:6-13)."""
    mx, my = matr.shape
    kx, ky = ker.shape
    out = F.conv2d(matr.reshape(1, 1, mx, my), ker.reshape(1, 1, kx, ky), padding=mode)
    return out.reshape(mx, my)


class SourceData:
    """Source point container (Source.py:299-303)."""

    def __init__(self, x, y):
        self.X = torch.zeros((x, y))
        self.Y = torch.zeros((x, y))
        self.Value = torch.zeros((x, y))


class PolarizationParameters:
    """Source.py:305-308."""

    def __init__(self):
        self.Degree = 1
        self.Angle = torch.tensor(0)


def _grid(cx, cy):
    s = SourceData(len(cx), len(cy))
    s.X, s.Y = torch.meshgrid(cx, cy, indexing="ij")
    s.Value = torch.zeros((len(cx), len(cy)), device=cx.device)
    return s


def CalculateAnnularSourceMatrix(SigmaOut, SigmaIn, Source_Coordinate_X, Source_Coordinate_Y):
    """Annulus SigmaIn <= r <= SigmaOut, weight 1/((So^2-Si^2) pi) (Source.py:229-238)."""
    s = _grid(Source_Coordinate_X, Source_Coordinate_Y)
    r = torch.sqrt(s.X.pow(2) + s.Y.pow(2))
    s.Value[(r <= SigmaOut) & (r >= SigmaIn)] = 1.0 / ((SigmaOut ** 2 - SigmaIn ** 2) * math.pi)
    return s


def CalculateQuasarSourceMatrix(SigmaOut, SigmaIn, openAngle, Source_Coordinate_X,
                                Source_Coordinate_Y):
    """Four poles centred on the axes (0, 90, 180, 270 degrees), half-opening
    openAngle/2, weight 0.5/((So^2-Si^2) openAngle) (Source.py:240-256)."""
    s = _grid(Source_Coordinate_X, Source_Coordinate_Y)
    r = torch.sqrt(s.X.pow(2) + s.Y.pow(2))
    th = torch.atan2(s.Y, s.X)
    half = openAngle / 2
    ath = torch.abs(th)
    on_x = (ath <= half) | ((ath >= math.pi - half) & (ath <= math.pi))
    on_y = (torch.abs(th - math.pi / 2) <= half) | (torch.abs(th + math.pi / 2) <= half)
    s.Value[(r <= SigmaOut) & (r >= SigmaIn) & (on_x | on_y)] = \
        0.5 / ((SigmaOut ** 2 - SigmaIn ** 2) * openAngle)
    return s


def CalculateDipoleSourceMatrix(SigmaOut, SigmaIn, openAngle, rotateAngle, Source_Coordinate_X,
                                Source_Coordinate_Y):
    """Two poles about rotateAngle: |cos(theta - rot)| >= cos(openAngle/2)
    (Source.py:258-270)."""
    s = _grid(Source_Coordinate_X, Source_Coordinate_Y)
    r = torch.sqrt(s.X.pow(2) + s.Y.pow(2))
    th = torch.atan2(s.Y, s.X)
    pole = torch.abs(torch.cos(th - rotateAngle)) >= math.cos(openAngle / 2)
    s.Value[(r <= SigmaOut) & (r >= SigmaIn) & pole] = 1.0 / ((SigmaOut ** 2 - SigmaIn ** 2) * openAngle)
    return s


def CalculateMultiCircSourceMatrix(SigmaCenter, SigmaRadius, PoleNumber, RotateAngle,
                                   Source_Coordinate_X, Source_Coordinate_Y):
    """PoleNumber discs of radius SigmaRadius on a circle of radius SigmaCenter
    (Source.py:273-286)."""
    s = _grid(Source_Coordinate_X, Source_Coordinate_Y)
    step = 2 * math.pi / PoleNumber
    for i in range(PoleNumber):
        xc = SigmaCenter * math.cos(RotateAngle + i * step)
        yc = SigmaCenter * math.sin(RotateAngle + i * step)
        inside = ((s.X - xc).pow(2) + (s.Y - yc).pow(2)) <= SigmaRadius ** 2
        s.Value[inside] = 1.0 / PoleNumber / (SigmaRadius ** 2)
    return s


def ConvertSourceMatrix2SourceData(s):
    """Zero weights outside the unit circle and flatten to (P*P, 1); like the
    reference (Source.py:288-297) the flattened arrays are also stored on the
    ``SourceData`` class."""
    r2 = s.X ** 2 + s.Y ** 2
    n = s.X.shape[0]
    s.Value[r2 > 1] = 0
    SourceData.X = torch.reshape(s.X, (n * n, 1))
    SourceData.Y = torch.reshape(s.Y, (n * n, 1))
    SourceData.Value = torch.reshape(s.Value, (n * n, 1))
    return SourceData


def _check_rotate(angle):
    if angle > math.pi / 2 or angle < -math.pi / 2:
        raise ValueError("error: roate angle must be in the range of [-pi/2,pi/2] ")


class Source:
    """Illumination parameters (Source.py:15-45); same attribute names and
    defaults."""

    def __init__(self):
        self.PntNum = 81
        self.Wavelength = 365
        self.Shape = "annular"
        self.SigmaOut = 0.7
        self.SigmaIn = 0.1
        self.SigmaCenter = 0.5
        self.SigmaRadius = 0.1
        self.OpenAngle = math.pi / 6
        self.RotateAngle = 0
        self.PoleNumber = 2
        self.Source_Mask = None
        self.SPointX = torch.zeros(self.PntNum)
        self.SPointY = torch.zeros(self.PntNum)
        self.SPointValue = torch.zeros(self.PntNum)
        self.PSFEnable = False
        self.PSFSigma = 0.02
        self.PolarizationParameters = PolarizationParameters()
        self.source_data = SourceData(self.PntNum, self.PntNum)
        self.PolarizationVector = [1.0, 0.0]
        self.PolarizationType = None

    def Calc_SourceSimple(self, device=None):
        """Lit points only: weights below 1e-5 are dropped (Source.py:47-54).
        Sets and returns ``self.source_data`` with 1-D X, Y, Value (on
        `device` when given, see Calc_SourceAll)."""
        self.source_data = self.Calc_SourceAll(device=device)
        low = self.source_data.Value < 1e-5
        self.source_data.X = self.source_data.X[~low]
        self.source_data.Y = self.source_data.Y[~low]
        self.source_data.Value = self.source_data.Value[~low]
        return self.source_data

    def Calc_SourceValid(self):
        """Source.py:56-62."""
        data = self.Calc_SourceAll()
        self.source_data.Value = torch.where((data.X.pow(2) + data.Y.pow(2)) > 1, 0, data.Value)
        return self.source_data

    def Calc_SourceAll(self, device=None):
        """P x P source: shape -> Source_Mask -> optional PSF blur -> centre
        pixel zeroed -> flattened (Source.py:64-180).

        With `device` (e.g. the HIP device) the shape rules, mask, blur and
        flattening run there on the float32 coordinate grid built on the host
        (torch.linspace, as the reference builds it); the elementwise float32
        arithmetic and comparisons are the same operations, so the lit set and
        weights equal the host's bit for bit except under the PSF blur, whose
        device convolution sums in another order (float32 rounding)."""
        coord = torch.linspace(-1, 1, self.PntNum)
        if device is not None:
            coord = coord.to(device)
        shape = self.Shape.lower()
        if shape == "pixel":
            n2 = self.PntNum ** 2
            if self.SPointX.numel() != n2:
                raise ValueError("Source Matrix Size X and difined PntNum are not matched ")
            if self.SPointY.numel() != n2:
                raise ValueError("Source Matrix Size Y and difined PntNum are not matched ")
            if self.SPointValue.numel() != n2:
                raise ValueError("Source Matrix Size Value and difined PntNum are not matched ")
            self.source_data.X = self.SPointX.to(coord.device)
            self.source_data.Y = self.SPointY.to(coord.device)
            self.source_data.Value = self.SPointValue.to(coord.device)
        elif shape == "annular":
            self.source_data = CalculateAnnularSourceMatrix(self.SigmaOut, self.SigmaIn, coord, coord)
        elif shape == "quasar":
            _check_rotate(self.RotateAngle)
            if self.OpenAngle <= 0 or self.OpenAngle >= math.pi / 2:
                raise ValueError("error: open angle must be in the range of [0,pi/2] ")
            self.source_data = CalculateQuasarSourceMatrix(self.SigmaOut, self.SigmaIn,
                                                           self.OpenAngle, coord, coord)
        elif shape == "dipolecirc":
            _check_rotate(self.RotateAngle)
            if self.OpenAngle < 0 or self.OpenAngle > math.pi / 2:
                raise ValueError("error: open angle must be in the range of [0,pi/2] ")
            self.source_data = CalculateDipoleSourceMatrix(self.SigmaOut, self.SigmaIn,
                                                           self.OpenAngle, self.RotateAngle,
                                                           coord, coord)
        elif shape == "multipole":
            _check_rotate(self.RotateAngle)
            self.source_data = CalculateMultiCircSourceMatrix(self.SigmaCenter, self.SigmaRadius,
                                                              self.PoleNumber, self.RotateAngle,
                                                              coord, coord)
        else:
            raise ValueError("unsupported illumination")

        sx, sy = self.source_data.Value.size()
        if self.Source_Mask is not None:
            self.source_data.Value = torch.mul(self.source_data.Value,
                                               torch.as_tensor(self.Source_Mask).to(coord.device))
        if sx == sy:
            if self.PSFEnable:
                self.source_data.Value = conv2(self.source_data.Value,
                                               self._psf_kernel().to(coord.device), mode="same")
            self.source_data.Value[int((sx - 1) / 2), int((sy - 1) / 2)] = 0
            self.source_data = ConvertSourceMatrix2SourceData(self.source_data)
        return self.source_data

    def _psf_kernel(self):
        """Gaussian blur exp(-r^2/sigma^2)/(sqrt(2 pi) sigma), trimmed where it
        falls below 1e-6 and normalised (Source.py:154-173)."""
        size = round(self.PntNum / 10) * 2 + 1
        edge = 1 / (self.PntNum - 1) * (size - 1)
        axis = torch.linspace(-edge, edge, size)
        kx, ky = torch.meshgrid(axis, axis, indexing="ij")
        ker = 1 / math.sqrt(2 * math.pi) / self.PSFSigma * torch.exp(
            -(kx.pow(2) + ky.pow(2)) / self.PSFSigma ** 2)
        ker = ker[~torch.all(ker < 1e-6, 0)]
        kt = torch.transpose(ker, 0, 1)
        ker = kt[~torch.all(kt < 1e-6, 1)]
        return ker / torch.sum(ker)

    def Calc_PolarizationMap(self, theta, rho):
        """Per-point polarisation (PolarizedX, PolarizedY) as complex64.
        Default: the constant PolarizationVector (Source.py:222-226).  With
        ``PolarizationType`` set, the per-type formulas of Source.py:184-221."""
        ptype = getattr(self, "PolarizationType", None)
        if ptype is None:
            px = self.PolarizationVector[0] * torch.ones(theta.size())
            py = self.PolarizationVector[1] * torch.ones(theta.size())
            return px.to(torch.complex64), py.to(torch.complex64)
        theta = torch.as_tensor(theta)
        rho = torch.as_tensor(rho)
        ones = torch.ones(theta.size())
        if ptype == "x_pol":
            px, py = ones, torch.zeros(theta.size())
        elif ptype == "y_pol":
            px, py = torch.zeros(theta.size()), ones
        elif ptype == "c_pol":
            px, py = math.sqrt(.5) * ones, math.sqrt(.5) * 1j * ones
        elif ptype == "d_pol":
            px, py = math.sqrt(.5) * ones, math.sqrt(.5) * ones
        elif ptype == "r_pol":
            px, py = torch.cos(theta), torch.sin(theta)
        elif ptype == "t_pol":
            px, py = torch.sin(theta), -1 * torch.cos(theta)
        elif ptype == "line_pol":
            ang = torch.as_tensor(self.PolarizationParameters.Angle, dtype=torch.float32)
            px, py = torch.sin(ang) * ones, torch.cos(ang) * ones
        else:
            raise ValueError("unsupported polarization type")
        px = px.to(torch.complex64).clone()
        py = py.to(torch.complex64).clone()
        tiny = rho < sys.float_info.epsilon
        px[tiny] = 0
        py[tiny] = 0
        return px, py

    def polarization_f64(self, X, Y):
        """(PX, PY) per source point as complex128 tensors, from float64 source
        coordinates (theta = atan2(Y, X), rho = |(X, Y)|)."""
        X = torch.as_tensor(X, dtype=torch.float64)
        Y = torch.as_tensor(Y, dtype=torch.float64)
        ptype = getattr(self, "PolarizationType", None)
        if ptype is None:
            one = torch.ones_like(X, dtype=torch.complex128)
            return complex(self.PolarizationVector[0]) * one, complex(self.PolarizationVector[1]) * one
        if ptype not in _POLARIZATIONS:
            raise ValueError("unsupported polarization type")
        th = torch.atan2(Y, X)
        rho = torch.sqrt(X * X + Y * Y)
        one = torch.ones_like(X)
        if ptype == "x_pol":
            px, py = one, 0 * one
        elif ptype == "y_pol":
            px, py = 0 * one, one
        elif ptype == "c_pol":
            px, py = math.sqrt(.5) * one, 1j * math.sqrt(.5) * one
        elif ptype == "d_pol":
            px, py = math.sqrt(.5) * one, math.sqrt(.5) * one
        elif ptype == "r_pol":
            px, py = torch.cos(th), torch.sin(th)
        elif ptype == "t_pol":
            px, py = torch.sin(th), -torch.cos(th)
        else:
            ang = float(torch.as_tensor(self.PolarizationParameters.Angle))
            px, py = math.sin(ang) * one, math.cos(ang) * one
        px = torch.as_tensor(px).to(torch.complex128).clone()
        py = torch.as_tensor(py).to(torch.complex128).clone()
        tiny = rho < sys.float_info.epsilon
        px[tiny] = 0
        py[tiny] = 0
        return px, py
