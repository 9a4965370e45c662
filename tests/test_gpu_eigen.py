"""The float64 Hermitian eigensolver (oi_eigh_hermitian, parallel cyclic
Jacobi) and the SOCS decomposition built on it (DecomposeTCC_SOCS.py:3-37).

Eigenvalues are compared with numpy's LAPACK eigvalsh; eigenvectors through
the residual |A V - V diag(w)| and orthonormality, which hold for any basis
of a degenerate eigenspace.  The decomposition is checked by reconstructing
the float64 oracle TCC from the kernels (sum_k phi_k phi_k^H = TCC at full
rank)."""
import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import annular, numerics, projector, quasar
from opticalimaging_amd.DecomposeTCC_SOCS import kernels_of

pytestmark = pytest.mark.gpu


def _check(A, tol=1e-11):
    from opticalimaging_amd.DecomposeTCC_SOCS import eigh
    w, V = eigh(torch.as_tensor(A))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, float(np.max(np.abs(ref))))
    ev_err = float(np.max(np.abs(np.sort(w) - np.sort(ref)))) / scale
    res = float(np.max(np.abs(A @ V - V * w[None, :]))) / scale
    orth = float(np.max(np.abs(V.conj().T @ V - np.eye(len(w)))))
    print(f"n={len(w)}: eigenvalues {ev_err:.1e}, residual {res:.1e}, orthonormality {orth:.1e}")
    assert ev_err < tol and res < tol and orth < tol


@pytest.mark.parametrize("n", [1, 2, 17, 120, 240])
def test_random_hermitian(gpu, n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    _check((X + X.conj().T) / 2)


def test_degenerate_spectrum(gpu):
    """Repeated eigenvalues (as the 4-fold symmetric sources give) and a
    rank-deficient PSD matrix."""
    rng = np.random.default_rng(5)
    n = 64
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.repeat(np.array([3.0, 1.5, 1.5, 0.25, 0.0, 0.0, 0.0, 0.0]), n // 8)
    _check((Q * lam[None, :]) @ Q.conj().T)


def test_real_symmetric_keeps_real_vectors(gpu):
    """A real symmetric input gives real eigenvectors (the rotations' phases
    are +-1), which the real-mask kernel pairing relies on."""
    from opticalimaging_amd.DecomposeTCC_SOCS import eigh
    rng = np.random.default_rng(2)
    X = rng.standard_normal((48, 48))
    A = (X + X.T) / 2
    _, V = eigh(torch.as_tensor(A.astype(np.complex128)))
    assert float(V.imag.abs().max()) == 0.0


@pytest.mark.parametrize("src_kind", ["annular", "quasar"])
def test_decomposition_reconstructs_tcc(gpu, src_kind):
    """Full-rank SOCS kernels of the device route (fp64 TCC contraction,
    Gram matrix, Jacobi) rebuild the float64 oracle TCC."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS
    src = annular() if src_kind == "annular" else quasar(P=21, pol="x_pol")
    proj = projector()
    num = numerics(order=10 ** 6, precision="fp64")
    period = 1000.0
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    kern = DecomposeTCC_SOCS(tcc, fg, [64, 64], num)
    ks = kernels_of(kern)
    ref, fg_o = O.tcc_matrix(src, [period, period], proj, 0.0, num)
    lenf, leng = fg_o[1], fg_o[0]
    # kernels on the TCC grid: the dense (Ny, Nx, K) array, un-fftshifted, cropped
    dense = np.fft.ifftshift(kern.cpu().numpy().astype(np.complex128), axes=(0, 1))
    r0, c0 = (64 - lenf + 1) // 2, (64 - leng + 1) // 2
    phi = dense[r0:r0 + lenf, c0:c0 + leng, :].reshape(lenf * leng, -1)
    rec = phi @ phi.conj().T
    err = float(np.max(np.abs(rec - ref)) / np.max(np.abs(ref)))
    print(f"{src_kind}: K = {ks.K}, reconstruction of the float64 TCC {err:.2e}")
    assert err < 1e-6
