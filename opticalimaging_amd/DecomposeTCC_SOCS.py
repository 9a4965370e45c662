"""SOCS decomposition of the TCC (DecomposeTCC_SOCS.py:3-37).

Eigenpairs are taken in float64 on the device (this library's parallel
Jacobi, ``oi_eigh_hermitian``).  Three routes, by what the TCC carries and
``numerics.TCC_Decomposition``:

* ``'projected'`` (default when the TCC comes from this package's
  ``CalculateTCCMatrix`` and is unmodified).  The TCC is B B^H up to the
  rounding of its contraction (B = [G_e diag(sqrt(alpha V))]_e, M x n,
  n = ncomp*Nsrc << M).  An orthonormal basis U of range(B) comes from the
  n x n float64 Gram matrix H = B^H B = W S^2 W^H (``oi_tcc_gram``):
  U = B W S^-1 (``oi_kernels_from_gram``).  The TCC *as computed* (split-bf16
  by default) is then decomposed on that range, Rayleigh-Ritz in float64:
  U^H TCC U = Z L Z^H, kernels U Z L^(1/2).  The contraction's rounding
  reaches the kernels exactly as far as it lies inside the TCC's true range.
* ``'factor'``: kernels straight from the Gram eigenpairs, B q_k (the TCC's
  values are not read; exact to float64 rounding of G).
* ``'dense'`` (and any TCC without a valid factor, e.g. the reference's own
  complex64 TCC, a copy, or one modified in place after it was built):
  Jacobi on the M x M TCC when M <= DENSE_JACOBI_MAX, else a float64
  subspace iteration (random range finder, power steps, Rayleigh-Ritz) on
  a block that grows until it holds the TCC's trace to 1e-6.

On the factor routes K stops at the TCC's rank (further kernels are exactly
zero).  The returned dense kernels (Ny, Nx, K) complex64 follow the reference layout
(embedding at rows (Ny-len f+1)//2, columns (Nx-len g+1)//2, then fftshift);
the tensor carries the device-resident windowed kernel set (``kernels_of``).
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _native as nat
from . import engine

__all__ = ["DecomposeTCC_SOCS", "socs_count", "eigen_desc", "socs_kernel_set", "kernels_of"]

DENSE_JACOBI_MAX = 1024     # larger dense TCCs go through the subspace iteration
_ROUTES = ("projected", "factor", "dense")


def kernels_of(kernels):
    """The device-resident KernelSet attached to DecomposeTCC_SOCS output, or
    None when there is none or the tensor was modified after it was made."""
    return engine.attached(kernels, "_oi_kernels")


def socs_count(s: torch.Tensor, M: int, numerics, total: float | None = None) -> int:
    """Number of kernels.  'order': min(Hopkins_Order, M - 1)
    (DecomposeTCC_SOCS.py:8-9).  'threshold': the reference's loop raises (it
    indexes the 2-D torch.diag(S)); its intent is kept: the number of leading
    values before the first whose running sum exceeds Hopkins_Threshold of the
    total (DecomposeTCC_SOCS.py:10-20).  `total` is the sum of all M singular
    values (the TCC's trace for a PSD TCC); by default the sum of `s`."""
    kind = str(numerics.Hopkins_SettingType).lower()
    if kind == "order":
        return int(min(numerics.Hopkins_Order, M - 1))
    if kind == "threshold":
        s = s.detach().to("cpu", torch.float64)
        total = float(s.sum()) if total is None else float(total)
        run = torch.cumsum(s[:max(M - 1, 0)], 0)
        over = torch.nonzero(run > numerics.Hopkins_Threshold * total)
        return int(over[0, 0]) if over.numel() else M - 1
    raise ValueError("Error: Hopkins_SettingType should be order or threshold!")


def eigh(H: torch.Tensor, max_sweeps: int = 60, tol: float = 1e-15):
    """Eigenvalues and eigenvectors of a Hermitian matrix in float64 on the
    device (``oi_eigh_hermitian``, parallel cyclic Jacobi).  H is not modified.
    Raises RuntimeError when the solver does not converge or meets a
    non-finite value."""
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


def _route(numerics) -> str:
    r = getattr(numerics, "TCC_Decomposition", "projected")
    if r not in _ROUTES:
        raise ValueError(f"TCC_Decomposition must be one of {_ROUTES}")
    return r


def _window_kernels(tcc, M, numerics):
    """(K, M) c64 kernels u_k sqrt(s_k) in TCC-grid order, K, and the pupil
    band when the kernels hold every non-zero eigenpair (else None)."""
    route = _route(numerics)
    factor = engine.attached(tcc, "_oi_factor")
    if factor is not None and factor.G.shape[-1] == M and route != "dense":
        if route == "factor":
            return _factor_kernels(factor, M, numerics)
        return _projected_kernels(tcc, factor, M, numerics)
    if M <= DENSE_JACOBI_MAX:
        s, U = eigen_desc(tcc.to(torch.complex128))
    else:
        s, U = _subspace_eigs(tcc, M, numerics)
    K = socs_count(s, M, numerics, total=float(torch.diagonal(tcc).real.double().sum()))
    K = min(K, int(s.numel()))
    win = (U[:, :K] * torch.sqrt(s[:K])[None, :]).transpose(0, 1).to(torch.complex64).contiguous()
    return win, K, None


def _gram(factor, M):
    """Float64 Gram matrix H = B^H B of the factor and the weights it uses."""
    lib = nat.load()
    G = factor.G
    dev = G.device
    ncomp, nsrc, _ = G.shape
    n = ncomp * nsrc
    w = (factor.V * factor.alpha).to(dev, torch.float64).contiguous()
    H = torch.empty((n, n), dtype=torch.complex128, device=dev)
    nat.check(lib.oi_tcc_gram(nat.ptr(G), nat.ptr(w), M, nsrc, ncomp, nat.ptr(H), nat.stream()))
    return H, w


def _from_gram(factor, w, Q, M, k):
    """out[j][m] = (B Q)[m][j] for the first k columns of Q, c64 [k][M]."""
    G = factor.G
    ncomp, nsrc, _ = G.shape
    out = torch.zeros((max(k, 1), M), dtype=torch.complex64, device=G.device)
    if k > 0:
        Qk = Q[:, :k].contiguous()
        nat.check(nat.load().oi_kernels_from_gram(nat.ptr(G), nat.ptr(w), M, nsrc, ncomp,
                                                  nat.ptr(Qk), k, nat.ptr(out), nat.stream()))
    return out


def _full_band(factor, K, n):
    if K >= n and factor.pitch[0] > 0:
        # every eigenpair kept: sum_k phi_k phi_k^H is the TCC, whose image
        # band is one pupil diameter, 2 floor(2 p) per axis
        return (int(math.floor(2 * factor.pitch[0])) + 1, int(math.floor(2 * factor.pitch[1])) + 1)
    return None


def _factor_kernels(factor, M, numerics):
    """Kernels from the TCC factor through the n x n Gram matrix (float64)."""
    G = factor.G
    dev = G.device
    n = G.shape[0] * G.shape[1]
    H, w = _gram(factor, M)
    s, Q = eigen_desc(H)
    s_all = torch.zeros(M, dtype=torch.float64, device=dev)
    r = min(n, M)
    s_all[:r] = s[:r]
    K = socs_count(s_all, M, numerics)
    keff = min(K, r)
    win = _from_gram(factor, w, Q, M, keff)
    # Kernels past the TCC's rank are exactly zero and change no image; they
    # are not materialised (with K = M - 1 on a large grid the reference's
    # dense (Ny, Nx, K) array would not fit in memory).
    K = max(keff, 1)
    return win[:K], K, _full_band(factor, K, n)


def _projected_kernels(tcc, factor, M, numerics):
    """Rayleigh-Ritz of the computed TCC on the factor's range (float64)."""
    G = factor.G
    dev = G.device
    n = G.shape[0] * G.shape[1]
    H, w = _gram(factor, M)
    s, Wv = eigen_desc(H)
    smax = float(s[0]) if s.numel() else 0.0
    r = int((s > smax * 1e-13).sum()) if smax > 0 else 0
    if r == 0:
        return torch.zeros((1, M), dtype=torch.complex64, device=dev), 1, None
    # orthonormal basis of range(B): U = B W S^-1 (columns), float64 from c64 G
    Ub = _from_gram(factor, w, Wv[:, :r] / torch.sqrt(s[:r])[None, :], M, r)     # c64 [r][M]
    U = Ub.to(torch.complex128).transpose(0, 1)                                   # (M, r)
    T = tcc.to(torch.complex128)
    A = U.conj().transpose(0, 1) @ (T @ U)
    A = 0.5 * (A + A.conj().transpose(0, 1))
    lam, Z = eigen_desc(A)
    s_all = torch.zeros(M, dtype=torch.float64, device=dev)
    s_all[:r] = lam[:r]
    K = socs_count(s_all, M, numerics, total=float(torch.diagonal(tcc).real.double().sum()))
    K = max(min(K, r), 1)
    win = ((U @ Z[:, :K]) * torch.sqrt(lam[:K])[None, :]).transpose(0, 1)
    # With every Ritz pair kept, sum_k phi_k phi_k^H = P TCC P (P the projector
    # on range(B)); it differs from the exact, pupil-band-limited TCC only by
    # the contraction's rounding, so the image keeps the pupil band.
    return win.to(torch.complex64).contiguous(), K, _full_band(factor, K, r)


def _orthonormal(Y: torch.Tensor, rel: float = 1e-13):
    """Orthonormal basis of range(Y) (float64 columns): Cholesky-QR through
    the library's eigensolver, twice for orthogonality at rounding level."""
    for _ in range(2):
        g = Y.conj().transpose(0, 1) @ Y
        g = 0.5 * (g + g.conj().transpose(0, 1))
        lam, W = eigen_desc(g)
        lmax = float(lam[0]) if lam.numel() else 0.0
        keep = int((lam > lmax * rel).sum()) if lmax > 0 else 0
        if keep == 0:
            return Y[:, :0]
        Y = (Y @ W[:, :keep]) / torch.sqrt(lam[:keep])[None, :]
    return Y


def _subspace_eigs(tcc, M, numerics, seed: int = 1234):
    """Leading eigenpairs of a dense Hermitian PSD TCC without an O(M^3)
    solve: block Y = T Omega (complex Gaussian, seeded), two power steps,
    Rayleigh-Ritz in float64; the block doubles until the Ritz values hold
    the TCC's trace to 1e-6 (or reaches M).  Returns (values descending,
    vectors (M, p))."""
    dev = tcc.device
    T = tcc.to(torch.complex128)
    trace = float(torch.diagonal(T).real.sum())
    kind = str(numerics.Hopkins_SettingType).lower()
    want = int(numerics.Hopkins_Order) if kind == "order" else 64
    p = min(M, max(64, 2 * min(want, M) + 16))
    gen = torch.Generator(device="cpu").manual_seed(seed)
    while True:
        om = torch.complex(torch.randn(M, p, generator=gen, dtype=torch.float64),
                           torch.randn(M, p, generator=gen, dtype=torch.float64)).to(dev)
        Y = _orthonormal(T @ om)
        for _ in range(2):
            Y = _orthonormal(T @ Y)
        A = Y.conj().transpose(0, 1) @ (T @ Y)
        A = 0.5 * (A + A.conj().transpose(0, 1))
        lam, Z = eigen_desc(A)
        held = float(lam.clamp_min(0).sum())
        if p >= M or Y.shape[1] < p or trace <= 0 or held >= (1 - 1e-6) * trace:
            return lam, Y @ Z
        p = min(M, 2 * p)


def DecomposeTCC_SOCS(TCCMatrix_Stacked, FG_ValidSize, Nfg, numerics):
    """Returns the SOCS kernels (Ny, Nx, K) complex64 on the device."""
    dev = nat.device()
    tcc = TCCMatrix_Stacked
    if not isinstance(tcc, torch.Tensor):
        tcc = torch.as_tensor(tcc)
    # a TCC copied to the device carries no factor: it takes the dense route
    tcc_dev = tcc if tcc.is_cuda else tcc.to(dev)
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
    engine.attach(dense, "_oi_kernels", ks)
    return dense


def socs_kernel_set(source, xypitch, projector, focus, numerics, Nfg) -> engine.KernelSet:
    """Device-resident kernel set equal to
    DecomposeTCC_SOCS(*CalculateTCCMatrix(source, xypitch, projector, focus,
    numerics), Nfg, numerics) (its attached kernel set) on the 'factor'
    route, without forming the M x M TCC or the dense (Ny, Nx, K) kernels
    (eigenpairs from the Gram matrix of the TCC's factor).  For serving and
    optimisation loops on large grids."""
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
