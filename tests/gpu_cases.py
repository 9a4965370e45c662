"""Shared small cases for the GPU parity tests (oracle results are computed on
the CPU in float64)."""
import math

import numpy as np
import torch

from opticalimaging_amd import Numerics, Projection, Scatter, Source
from opticalimaging_amd import workloads


def rel_err(a, b):
    """max |a - b| / max |b| (peak-normalised)."""
    a = np.asarray(a)
    b = np.asarray(b)
    a = a.astype(np.complex128) if np.iscomplexobj(a) else a.astype(np.float64)
    b = b.astype(np.complex128) if np.iscomplexobj(b) else b.astype(np.float64)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def annular(P=21, lam=193, so=0.8, si=0.5):
    s = Source()
    s.PntNum = P
    s.Wavelength = lam
    s.SigmaOut = so
    s.SigmaIn = si
    return s


def quasar(P=41, lam=193, so=0.9, si=0.6, pol="t_pol"):
    s = Source()
    s.PntNum = P
    s.Wavelength = lam
    s.Shape = "quasar"
    s.SigmaOut = so
    s.SigmaIn = si
    s.OpenAngle = math.pi / 6
    s.PolarizationType = pol
    return s


def projector(NA=0.85, n=1.0, M=100, foci=(0,), zern=None):
    p = Projection()
    p.NA = NA
    p.IndexImage = n
    p.Magnification = M
    p.FocusRange = torch.tensor(list(foci), dtype=torch.float32) if len(foci) > 1 or foci[0] != 0 \
        else torch.tensor([0])
    if zern is not None:
        p.Aberration_Zernike = torch.tensor(zern, dtype=torch.float32)
    return p


def numerics(mode="scalar", order=24, kind="order", precision="split_bf16", hw=None):
    n = Numerics()
    n.ImageCalculationMode = mode
    n.Hopkins_SettingType = kind
    n.Hopkins_Order = order
    n.TCC_Precision = precision
    n.Hopkins_GridHalfWidth = hw
    return n


def scatter(field, px, py):
    sc = Scatter()
    sc.Period_X = px
    sc.Period_Y = py
    sc.ScatterField = field
    return sc


def binary_field(ny, nx, seed, frac=0.5):
    rng = np.random.default_rng(seed)
    m = (rng.random((ny, nx)) < frac).astype(np.float32)
    return workloads.field_from_mask(m)


def complex_field(ny, nx, seed, chans=3):
    rng = np.random.default_rng(seed)
    f = np.zeros((ny, nx, 3), dtype=np.complex64)
    for c in range(chans):
        f[..., c] = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))) * (0.5 ** c)
    return torch.as_tensor(f)
