"""Polarisation transfer (characteristic) matrix on the HIP device
(CalculateCharacteristicMatrix.py:3-16)."""
from __future__ import annotations

import torch

from . import _native as nat


def CalculateCharacteristicMatrix(f_calc, g_calc, NA, indexImage):
    """Returns (Mxx, Myx, Mxy, Myy, Mxz, Myz), complex64, shaped like f_calc
    and on its device: alpha = (NA/n) f, beta = (NA/n) g,
    gamma = sqrt(1 - alpha^2 - beta^2), Mxx = 1 - alpha^2/(1+gamma),
    Myx = Mxy = -alpha beta/(1+gamma), Myy = 1 - beta^2/(1+gamma),
    Mxz = -alpha, Myz = -beta."""
    dev = nat.device()
    f = torch.as_tensor(f_calc)
    out_dev = f.device
    ff = f.to(device=dev, dtype=torch.float32)
    gg = torch.as_tensor(g_calc).to(device=dev, dtype=torch.float32)
    ff, gg = torch.broadcast_tensors(ff, gg)
    shape = ff.shape
    ff = ff.contiguous().reshape(-1)
    gg = gg.contiguous().reshape(-1)
    out = torch.empty((6, ff.numel()), dtype=torch.complex64, device=dev)
    nat.check(nat.load().oi_char_matrix(nat.ptr(ff), nat.ptr(gg), ff.numel(), float(NA),
                                        float(indexImage), nat.ptr(out), nat.stream()))
    out = out.reshape((6,) + tuple(shape)).to(out_dev)
    return tuple(out[i] for i in range(6))
