"""The identity behind imaging real masks with paired kernels
(engine.real_mask_kernels), checked in float64 on the CPU with the oracle's
SOCS image (CalculateAerialImage_SOCS.py:3-11).

The pairing is restated here in numpy from its definition: J phi(kappa) =
conj phi(-kappa), the J-real parts u = (phi + J phi)/2, w = (phi - J phi)/(2i),
the eigenvectors of their real Gram matrix, and pairs chi_a + i chi_b.  The
tests show that it leaves every real-mask image unchanged, halves the kernel
count for in-focus aberration-free optics and odd aberrations, gains nothing
with even aberrations, and does not hold for complex masks (so imaging keeps
the original kernels there)."""
import numpy as np
import pytest

import oracle as O
from gpu_cases import annular, numerics, projector, rel_err


def _J(a):
    """(J a)[p] = conj a[-p mod N] on the standard FFT layout (DC at index 0)."""
    return np.conj(np.roll(np.flip(a, (0, 1)), (1, 1), (0, 1)))


def _pack(kern, rel_tol=1e-10):
    ny, nx, K = kern.shape
    u = [(kern[..., k] + _J(kern[..., k])) / 2 for k in range(K)]
    w = [(kern[..., k] - _J(kern[..., k])) / 2j for k in range(K)]
    X = np.stack(u + w).reshape(2 * K, -1)
    G = X.conj() @ X.T
    assert np.abs(G.imag).max() <= 1e-12 * np.abs(G).max()
    lam, Q = np.linalg.eigh(G.real)
    order = np.argsort(-lam)
    lam, Q = np.clip(lam[order], 0, None), Q[:, order]
    tail = np.cumsum(lam[::-1])[::-1]
    keep = int((tail > rel_tol * lam.sum()).sum())
    chi = Q[:, :keep].T @ X
    if keep % 2:
        chi = np.concatenate([chi, np.zeros_like(chi[:1])])
    psi = chi[0::2] + 1j * chi[1::2]
    return np.moveaxis(psi.reshape(-1, ny, nx), 0, -1), keep


def _kernels(n, period, order, focus=0.0):
    src, proj, num = annular(), projector(), numerics(order=order)
    tcc, fg = O.tcc_matrix(src, [period, period], proj, focus, num)
    return O.decompose_tcc_socs(tcc, fg, [n, n], num), proj, src


def _image(field2d, kern, period, proj, src):
    f3 = np.zeros(field2d.shape + (3,), dtype=np.complex128)
    f3[..., 0] = field2d
    return O.socs_image(f3, kern, period, period, proj.NA, src.Wavelength)


@pytest.mark.parametrize("order", [10 ** 6, 24, 7])
def test_pairing_keeps_real_mask_images(order):
    n, period = 64, 1000.0
    kern, proj, src = _kernels(n, period, order)
    packed, keep = _pack(kern)
    K = kern.shape[2]
    assert packed.shape[2] == (keep + 1) // 2 < K
    rng = np.random.default_rng(order % 97)
    for mask in [(rng.random((n, n)) < 0.4).astype(float), rng.random((n, n))]:
        ref = _image(mask, kern, period, proj, src)
        got = _image(mask, packed, period, proj, src)
        err = rel_err(got, ref)
        print(f"order {order}: K {K} -> {packed.shape[2]} kernels, real-mask image err {err:.1e}")
        assert err < 1e-12


def test_full_rank_set_halves_in_focus():
    """The full TCC commutes with J for this symmetric source and real pupil:
    rank(G) = rank(TCC) and half as many kernels remain."""
    kern, _, _ = _kernels(64, 1000.0, 10 ** 6)
    energy = (np.abs(kern) ** 2).sum(axis=(0, 1))
    rank = int((energy > 1e-10 * energy.sum()).sum())      # eigenvalues of the TCC above round-off
    packed, keep = _pack(kern)
    assert keep == rank and packed.shape[2] == (rank + 1) // 2


@pytest.mark.parametrize("term,expect_half", [(3, False), (4, False), (8, False), (6, True)])
def test_aberrations(term, expect_half):
    """An aberration whose pupil phase is even under f -> -f breaks the
    symmetry (2K J-real components, nothing gained); an odd one keeps it."""
    z = [0.0] * 37
    z[term] = 0.05
    src, proj, num = annular(), projector(zern=z), numerics(order=12)
    tcc, fg = O.tcc_matrix(src, [1000.0, 1000.0], proj, 0.0, num)
    kern = O.decompose_tcc_socs(tcc, fg, [64, 64], num)
    packed, keep = _pack(kern)
    assert keep == (12 if expect_half else 24)
    mask = (np.random.default_rng(term).random((64, 64)) < 0.4).astype(float)
    assert rel_err(_image(mask, packed, 1000.0, proj, src), _image(mask, kern, 1000.0, proj, src)) < 1e-12


def test_pairing_does_not_hold_for_complex_masks():
    n, period = 64, 1000.0
    kern, proj, src = _kernels(n, period, 10 ** 6)
    packed, _ = _pack(kern)
    rng = np.random.default_rng(5)
    field = rng.random((n, n)) * np.exp(1j * np.pi * (rng.random((n, n)) < 0.5) * 0.5)
    ref = _image(field, kern, period, proj, src)
    got = _image(field, packed, period, proj, src)
    assert rel_err(got, ref) > 1e-3
