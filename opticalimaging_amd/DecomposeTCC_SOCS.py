"""SOCS decomposition of the TCC (DecomposeTCC_SOCS.py:3-37).

Eigenpairs are taken in float64.  When the TCC comes from this package's
``CalculateTCCMatrix`` it carries its factor TCC = B B^H (B = G_e diag(sqrt(a V)),
M x n with n = ncomp*Nsrc << M); the non-zero eigenpairs then come from the
n x n Gram matrix H = B^H B (``oi_tcc_gram``, float64): H q = l q gives the TCC
eigenvector u = B q / sqrt(l), so each SOCS kernel u sqrt(l) = B q
(``oi_kernels_from_gram``).  Otherwise the dense M x M TCC is decomposed.
The float64 Hermitian eigensolver is this library's parallel Jacobi
(``oi_eigh_hermitian``) on the device.

On the factor route K stops at the TCC's rank (further kernels are exactly
zero).  The returned dense kernels (Ny, Nx, K) complex64 follow the reference layout
(embedding at rows (Ny-len f+1)//2, columns (Nx-len g+1)//2, then fftshift);
the tensor carries the device-resident windowed kernel set as ``_oi_kernels``.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _native as nat
from . import engine

__all__ = ["DecomposeTCC_SOCS", "socs_count", "eigen_desc", "socs_kernel_set"]


def socs_count(s: torch.Tensor, M: int, numerics) -> int:
    """Number of kernels.  'order': min(Hopkins_Order, M - 1)
    (DecomposeTCC_SOCS.py:8-9).  'threshold': the reference's loop raises (it
    indexes the 2-D torch.diag(S)); its intent is kept: the number of leading
    values before the first whose running sum exceeds Hopkins_Threshold of the
    total (DecomposeTCC_SOCS.py:10-20)."""
    kind = str(numerics.Hopkins_SettingType).lower()
    if kind == "order":
        return int(min(numerics.Hopkins_Order, M - 1))
    if kind == "threshold":
        s = s.detach().to("cpu", torch.float64)
        total = float(s.sum())
        run = torch.cumsum(s[:max(M - 1, 0)], 0)
        over = torch.nonzero(run > numerics.Hopkins_Threshold * total)
        return int(over[0, 0]) if over.numel() else M - 1
    raise ValueError("Error: Hopkins_SettingType should be order or threshold!")


def eigh(H: torch.Tensor, max_sweeps: int = 60, tol: float = 1e-15):
    """Eigenvalues and eigenvectors of a Hermitian matrix in float64 on the
    device (``oi_eigh_hermitian``, parallel cyclic Jacobi).  H is not modified."""
    dev = nat.device()
    A = H.to(device=dev, dtype=torch.complex128).contiguous().clone()
    n = int(A.shape[0])
    lib = nat.load()
    w = torch.empty(n, dtype=torch.float64, device=dev)
    V = torch.empty((n, n), dtype=torch.complex128, device=dev)
    work = torch.empty(int(lib.oi_eigh_work_bytes(n)), dtype=torch.uint8, device=dev)
    sweeps = ctypes.c_int(0)
    nat.check(lib.oi_eigh_hermitian(nat.ptr(A), n, nat.ptr(w), nat.ptr(V), max_sweeps, tol,
                                    nat.ptr(work), ctypes.byref(sweeps), nat.stream()))
    return w, V


def eigen_desc(H: torch.Tensor):
    """|eigenvalues| descending and matching eigenvectors of a Hermitian
    matrix (float64) -- the singular values / left vectors torch.svd gives for
    a PSD TCC (DecomposeTCC_SOCS.py:7)."""
    lam, Q = eigh(H)
    s = lam.abs()
    order = torch.argsort(s, descending=True, stable=True)
    return s[order], Q[:, order]


def _window_kernels(tcc, M, numerics):
    """(K, M) c64 kernels u_k sqrt(s_k) in TCC-grid order, K, and the pupil
    band when the kernels hold every non-zero eigenpair (else None)."""
    factor = getattr(tcc, "_oi_factor", None)
    if factor is not None and factor.G.shape[-1] == M:
        return _factor_kernels(factor, M, numerics)
    T = tcc.to(torch.complex128)
    s, U = eigen_desc(T)
    K = socs_count(s, M, numerics)
    win = (U[:, :K] * torch.sqrt(s[:K])[None, :]).transpose(0, 1).to(torch.complex64).contiguous()
    return win, K, None


def _factor_kernels(factor, M, numerics):
    """Kernels from the TCC factor through the n x n Gram matrix (float64)."""
    lib = nat.load()
    G = factor.G
    dev = G.device
    ncomp, nsrc, _ = G.shape
    n = ncomp * nsrc
    w = (factor.V * factor.alpha).to(dev, torch.float64).contiguous()
    H = torch.empty((n, n), dtype=torch.complex128, device=dev)
    nat.check(lib.oi_tcc_gram(nat.ptr(G), nat.ptr(w), M, nsrc, ncomp, nat.ptr(H), nat.stream()))
    s, Q = eigen_desc(H)
    s_all = torch.zeros(M, dtype=torch.float64, device=dev)
    r = min(n, M)
    s_all[:r] = s[:r]
    K = socs_count(s_all, M, numerics)
    keff = min(K, r)
    win = torch.zeros((max(keff, 1), M), dtype=torch.complex64, device=dev)
    if keff > 0:
        Qk = Q[:, :keff].contiguous()
        nat.check(lib.oi_kernels_from_gram(nat.ptr(G), nat.ptr(w), M, nsrc, ncomp, nat.ptr(Qk),
                                           keff, nat.ptr(win), nat.stream()))
    band = None
    if K >= n and factor.pitch[0] > 0:
        # every eigenpair kept: sum_k phi_k phi_k^H is the TCC, whose image
        # band is one pupil diameter, 2 floor(2 p) per axis
        band = (int(math.floor(2 * factor.pitch[0])) + 1, int(math.floor(2 * factor.pitch[1])) + 1)
    # Kernels past the TCC's rank are exactly zero and change no image; they
    # are not materialised (with K = M - 1 on a large grid the reference's
    # dense (Ny, Nx, K) array would not fit in memory).
    K = max(keff, 1)
    return win[:K], K, band


def DecomposeTCC_SOCS(TCCMatrix_Stacked, FG_ValidSize, Nfg, numerics):
    """Returns the SOCS kernels (Ny, Nx, K) complex64 on the device."""
    dev = nat.device()
    tcc = TCCMatrix_Stacked
    if not isinstance(tcc, torch.Tensor):
        tcc = torch.as_tensor(tcc)
    factor = getattr(tcc, "_oi_factor", None)
    tcc_dev = tcc if tcc.is_cuda else tcc.to(dev)
    if factor is not None:
        tcc_dev._oi_factor = factor
    nx, ny = int(Nfg[0]), int(Nfg[1])
    leng, lenf = int(FG_ValidSize[0]), int(FG_ValidSize[1])
    M = lenf * leng
    if tcc_dev.shape != (M, M):
        raise ValueError("TCC size does not match FG_ValidSize")
    if lenf > ny or leng > nx:
        # the reference fails here with torch's RuntimeError (DecomposeTCC_SOCS.py:28-30)
        raise RuntimeError("TCC grid larger than the mask grid")
    win, K, band = _window_kernels(tcc_dev, M, numerics)
    win = win.reshape(K, lenf, leng)
    r0 = (ny - lenf + 1) // 2
    c0 = (nx - leng + 1) // 2
    temp = torch.zeros((K, ny, nx), dtype=torch.complex64, device=dev)
    temp[:, r0:r0 + lenf, c0:c0 + leng] = win
    dense = torch.fft.fftshift(temp, dim=(1, 2)).permute(1, 2, 0).contiguous()
    ks = _kernel_set(win, r0, c0, ny, nx, dense)
    ks.band = band
    dense._oi_kernels = ks
    return dense


def socs_kernel_set(source, xypitch, projector, focus, numerics, Nfg) -> engine.KernelSet:
    """Device-resident kernel set equal to
    DecomposeTCC_SOCS(*CalculateTCCMatrix(source, xypitch, projector, focus,
    numerics), Nfg, numerics)._oi_kernels, without forming the M x M TCC or
    the dense (Ny, Nx, K) kernels (eigenpairs from the Gram matrix of the
    TCC's factor).  For serving and optimisation loops on large grids."""
    from .CalculateTCCMatrix import tcc_factor
    factor = tcc_factor(source, xypitch, projector, focus, numerics)
    nx, ny = int(Nfg[0]), int(Nfg[1])
    lenf, leng = factor.lenf, factor.leng
    if lenf > ny or leng > nx:
        raise RuntimeError("TCC grid larger than the mask grid")
    win, K, band = _factor_kernels(factor, lenf * leng, numerics)
    win = win.reshape(K, lenf, leng)
    r0 = (ny - lenf + 1) // 2
    c0 = (nx - leng + 1) // 2
    dense = None
    if ny % 2 or nx % 2:
        temp = torch.zeros((K, ny, nx), dtype=torch.complex64, device=win.device)
        temp[:, r0:r0 + lenf, c0:c0 + leng] = win
        dense = torch.fft.fftshift(temp, dim=(1, 2)).permute(1, 2, 0).contiguous()
    ks = _kernel_set(win, r0, c0, ny, nx, dense)
    ks.band = band
    return ks


def _kernel_set(win, r0, c0, ny, nx, dense):
    """Windowed kernel set cropped to the exact non-zero support.  The kernel
    paired with spectrum index kappa is dense[kappa mod N]
    (CalculateAerialImage_SOCS.py:6); for even N that is window entry
    kappa + (len-1)/2.  Odd grids go through the dense array (the reference's
    fftshift/ifftshift pair shifts them by one bin)."""
    if ny % 2 or nx % 2:
        return engine.KernelSet.from_dense(dense)
    K, lenf, leng = win.shape
    nzr = torch.nonzero((win != 0).any(dim=2).any(dim=0)).reshape(-1)
    nzc = torch.nonzero((win != 0).any(dim=1).any(dim=0)).reshape(-1)
    if nzr.numel() == 0:
        return engine.KernelSet(torch.zeros((K, 1, 1), dtype=torch.complex64, device=win.device),
                                0, 0, ny, nx)
    a0, a1 = int(nzr.min()), int(nzr.max())
    b0, b1 = int(nzc.min()), int(nzc.max())
    k0r = r0 + a0 - ny // 2
    k0c = c0 + b0 - nx // 2
    return engine.KernelSet(win[:, a0:a1 + 1, b0:b1 + 1].contiguous(), k0r, k0c, ny, nx)
