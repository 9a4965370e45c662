"""Projection-lens parameter object, attribute-compatible with the reference's
``Projection`` (Projection.py:4-10).

``CalculateAberrationFast`` evaluates the 37-term Fringe Zernike wavefront on
the HIP device (``oi_zernike``); ``RotateAngle`` rotates the coefficient
vector (Projection.py:130-161) on the host.
"""
from __future__ import annotations

import ctypes
import math
import sys

import torch

from . import _native as nat

# Fringe Z1..Z37: azimuthal order m and kind (+1 cos, -1 sin, 0 symmetric).
_FRINGE = [(0, 0), (1, 1), (1, -1), (0, 0), (2, 1), (2, -1), (1, 1), (1, -1), (0, 0), (3, 1),
           (3, -1), (2, 1), (2, -1), (1, 1), (1, -1), (0, 0), (4, 1), (4, -1), (3, 1), (3, -1),
           (2, 1), (2, -1), (1, 1), (1, -1), (0, 0), (5, 1), (5, -1), (4, 1), (4, -1), (3, 1),
           (3, -1), (2, 1), (2, -1), (1, 1), (1, -1), (0, 0), (0, 0)]


def _partner(i: int) -> int:
    """Index of the cos/sin partner of Fringe term i (itself when symmetric)."""
    m, kind = _FRINGE[i]
    if kind == 0:
        return i
    return i + 1 if kind == 1 else i - 1


class Projection:
    """Lens parameters (Projection.py:5-10); same attribute names and defaults."""

    def __init__(self):
        self.Aberration_Zernike = torch.zeros(37)
        self.Magnification = 100
        self.NA = 0.9
        self.IndexImage = 1.0
        self.FocusRange = torch.tensor([0])

    def zernike_coefficients(self, Orientation=0.0):
        """The 37 coefficients (float64 list) after rotation by Orientation."""
        if abs(float(Orientation)) > sys.float_info.epsilon:
            c = Projection.RotateAngle(self.Aberration_Zernike, Orientation)
        else:
            c = self.Aberration_Zernike
        c = torch.as_tensor(c, dtype=torch.float64).reshape(-1).cpu()
        if c.numel() != 37:
            raise ValueError("Aberration_Zernike must hold 37 Fringe coefficients")
        return [float(v) for v in c.tolist()]

    def CalculateAberrationFast(self, rho, theta, Orientation):
        """Wavefront W(rho, theta) in waves (Projection.py:13-128).  Computed on
        the HIP device; returned on rho's device."""
        dev = nat.device()
        rho_t = torch.as_tensor(rho)
        out_dev = rho_t.device
        r = rho_t.to(device=dev, dtype=torch.float32).contiguous()
        t = torch.as_tensor(theta).to(device=dev, dtype=torch.float32)
        r, t = torch.broadcast_tensors(r, t)
        r = r.contiguous()
        t = t.contiguous()
        out = torch.empty_like(r)
        coeffs = (ctypes.c_double * 37)(*self.zernike_coefficients(Orientation))
        nat.check(nat.load().oi_zernike(nat.ptr(r), nat.ptr(t), r.numel(), coeffs, nat.ptr(out),
                                        nat.stream()))
        return out.to(out_dev)

    @staticmethod
    def RotateAngle(c0, theta):
        """Rotate Fringe coefficients by theta (Projection.py:130-161): a term
        of azimuthal order m keeps c cos(m theta) and hands -kind c sin(m theta)
        to its cos/sin partner."""
        c0 = torch.as_tensor(c0, dtype=torch.float32).reshape(-1)
        th = float(theta)
        c1 = torch.zeros(37)
        for i in range(37):
            m, kind = _FRINGE[i]
            c1[i] = c1[i] + c0[i] * math.cos(m * th)
            j = _partner(i)
            c1[j] = c1[j] - kind * c0[i] * math.sin(m * th)
        return c1
