"""Transmission cross coefficients on the HIP device (CalculateTCCMatrix.py).

The shifted-pupil matrix G (M x Nsrc per polarisation component) is built by
``oi_pupil_kernels`` in float64 and the Hermitian contraction
TCC = n G diag(V) G^H / sum(V) runs on MFMA (``oi_tcc_build``), in 3-term
split-bf16 with fp32 accumulation unless ``numerics.TCC_Precision`` says
'fp32' or 'fp64'.

The returned TCC tensor lives on the device and carries its low-rank factor
(G, the source weights and the scale; ``factor_of``), valid while the tensor
is unmodified.  ``DecomposeTCC_SOCS`` uses it to decompose the TCC on the
factor's range (float64 Rayleigh-Ritz on an n x n problem) instead of
solving the M x M eigenproblem.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _native as nat
from . import engine, optics

__all__ = ["CalculateTCCMatrix", "GetTCCMatrix", "TCCFactor", "tcc_factor", "tcc_contract",
           "cartesian_to_polar", "factor_of"]

_PRECISIONS = {"split_bf16": nat.OI_TCC_SPLIT_BF16, "fp32": nat.OI_TCC_FP32, "fp64": nat.OI_TCC_FP64}


def cartesian_to_polar(x, y):
    """rho, theta (CalculateTCCMatrix.py:4-7)."""
    return torch.sqrt(x ** 2 + y ** 2), torch.atan2(y, x)


@dataclass
class TCCFactor:
    """TCC = alpha * sum_e G_e diag(V) G_e^H with G: c64 [ncomp][nsrc][M]."""
    G: torch.Tensor
    V: torch.Tensor       # float64 [nsrc], device
    alpha: float
    lenf: int
    leng: int
    pitch: tuple = (0.0, 0.0)   # (p_0, p_1) = xypitch * NA / lambda


def factor_of(tcc):
    """The TCCFactor attached by CalculateTCCMatrix, or None when there is
    none or the TCC was modified after it was built."""
    return engine.attached(tcc, "_oi_factor")


def precision_code(numerics=None) -> int:
    name = getattr(numerics, "TCC_Precision", "split_bf16") if numerics is not None else "split_bf16"
    if name not in _PRECISIONS:
        raise ValueError(f"TCC_Precision must be one of {sorted(_PRECISIONS)}")
    return _PRECISIONS[name]


def tcc_contract(G: torch.Tensor, V: torch.Tensor, alpha: float, precision: int,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """tcc[m][m'] = alpha sum_e sum_s V_s G[e][s][m] conj(G[e][s][m'])."""
    dev = nat.device()
    G = G.to(device=dev, dtype=torch.complex64).contiguous()
    ncomp, nsrc, M = G.shape
    w = V.to(device=dev, dtype=torch.float32).contiguous()
    lib = nat.load()
    nbytes = int(lib.oi_tcc_work_bytes(M, nsrc, ncomp, precision))
    work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
    if out is None:
        out = torch.empty((M, M), dtype=torch.complex64, device=dev)
    nat.check(lib.oi_tcc_build(nat.ptr(G), nat.ptr(w), M, nsrc, ncomp, float(alpha), precision,
                               nat.ptr(work), nat.ptr(out), nat.stream()))
    return out


def tcc_factor(source, xypitch, projector, focus, numerics) -> TCCFactor:
    """The low-rank factor of CalculateTCCMatrix's TCC (shifted pupils G,
    source weights, scale) without the M x M contraction."""
    vector = optics.imaging_mode(numerics)
    lam = float(source.Wavelength)
    na = float(projector.NA)
    nf, ng = optics.tcc_halfwidths(xypitch, na, lam, numerics)
    p0 = xypitch[0] / (lam / na)
    p1 = xypitch[1] / (lam / na)
    src = optics.source_set(source)
    opt = optics.optics_record(source, projector, float(focus), p0, p1, vector)
    ones = torch.ones(src.n, dtype=torch.float64, device=src.V.device)
    lenf, leng = 2 * nf + 1, 2 * ng + 1
    G = optics.pupil_kernels(src, opt, ones, -nf, lenf, -ng, leng)
    ncomp = G.shape[0]
    G = G.reshape(ncomp, src.n, lenf * leng)
    alpha = float(projector.IndexImage) / src.weight
    return TCCFactor(G=G, V=src.V, alpha=alpha, lenf=lenf, leng=leng, pitch=(p0, p1))


def CalculateTCCMatrix(source, xypitch, projector, focus, numerics):
    """Returns (TCC [M, M] complex64 on the device, [len(g), len(f)]).

    Grid: f = k / p_0, g = k / p_1 with p_i = xypitch[i] NA / lambda,
    k in [-N, N], N = ceil(2 p) (CalculateTCCMatrix.py:12-22) or
    numerics.Hopkins_GridHalfWidth.  Row m = a len(g) + b holds (f[a], g[b])
    (CalculateTCCMatrix.py:23); f pairs with the source X coordinate."""
    factor = tcc_factor(source, xypitch, projector, focus, numerics)
    tcc = tcc_contract(factor.G, factor.V, factor.alpha, precision_code(numerics))
    # valid while the TCC is unmodified (engine.attached checks its version)
    engine.attach(tcc, "_oi_factor", factor)
    return tcc, [factor.leng, factor.lenf]


def GetTCCMatrix(sourceData, shiftedPupil, numerics=None):
    """(G diag(V) G^H) / sum(V) for a shifted-pupil matrix G (M x Nsrc)
    (CalculateTCCMatrix.py:58-76)."""
    dev = nat.device()
    V = torch.as_tensor(sourceData.Value).reshape(-1).to(dev, torch.float64)
    G = torch.as_tensor(shiftedPupil).to(dev, torch.complex64)
    if G.dim() != 2 or G.shape[1] != V.numel():
        raise ValueError("shiftedPupil must be (M, Nsrc) with Nsrc = len(sourceData.Value)")
    Gt = G.transpose(0, 1).contiguous().unsqueeze(0)
    return tcc_contract(Gt, V, 1.0 / float(V.sum()), precision_code(numerics))
