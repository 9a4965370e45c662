"""Pins of the float64 oracle (CPU only).

The reference ships no tests or golden vectors for the imaging path, so every
oracle function is checked here against what the mathematics fixes: closed
forms, brute-force sums on tiny inputs, invariants and identities.  Each check
is chosen so that a plausible slip (a dropped term, a wrong sign or index, a
transposed axis) fails it.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import annular, numerics, projector, quasar, scatter


# --------------------------------------------------------------------- source
def test_annular_source_counts_and_weights():
    X, Y, V = O.source_points(annular())
    assert len(V) == 120
    w = np.float32(1.0 / ((0.8 ** 2 - 0.5 ** 2) * math.pi))
    assert np.all(V == w)
    assert abs(V.sum() - 97.9415) < 1e-3
    r = np.hypot(X, Y)
    assert r.min() >= 0.5 - 1e-6 and r.max() <= 0.8 + 1e-6


def test_quasar_source_counts_and_weights():
    X, Y, V = O.source_points(quasar(P=41))
    assert len(V) == 196
    assert abs(V.sum() - 415.925) < 1e-2
    th = np.degrees(np.arctan2(Y, X)) % 90.0
    # poles centred on the axes, half opening 15 degrees
    assert np.all((th <= 15 + 1e-4) | (th >= 75 - 1e-4))


@pytest.mark.parametrize("maker", [annular, lambda: quasar(P=41)])
def test_source_fourfold_symmetry(maker):
    X, Y, V = O.source_points(maker())
    pts = {(round(x, 5), round(y, 5)) for x, y in zip(X, Y)}
    assert pts == {(round(-y, 5), round(x, 5)) for x, y in pts}


def test_pixel_source_passthrough_and_centre_zeroed():
    s = annular(P=5)
    s.Shape = "pixel"
    g = torch.linspace(-1, 1, 5)
    s.SPointX, s.SPointY = torch.meshgrid(g, g, indexing="ij")
    s.SPointValue = torch.arange(25, dtype=torch.float32).reshape(5, 5) / 10 + 0.5
    X, Y, V = O.source_points(s)
    # centre pixel zeroed (Source.py:178), points with r^2 > 1 zeroed (Source.py:291):
    # on the 5 x 5 grid 13 points have r <= 1, one of them the centre
    assert len(V) == 12
    assert not np.any((X == 0) & (Y == 0))
    assert np.all(X ** 2 + Y ** 2 <= 1 + 1e-6)


def test_pixel_source_size_mismatch_raises():
    s = annular(P=5)
    s.Shape = "pixel"
    with pytest.raises(ValueError):
        O.source_points(s)


def test_multipole_and_dipole():
    s = annular(P=31)
    s.Shape = "multipole"
    s.PoleNumber = 4
    s.SigmaCenter, s.SigmaRadius = 0.6, 0.15
    X, Y, V = O.source_points(s)
    assert np.allclose(V, np.float32(1 / 4 / 0.15 ** 2))
    # every lit point lies within SigmaRadius of a pole centre
    poles = [(0.6 * math.cos(i * math.pi / 2), 0.6 * math.sin(i * math.pi / 2)) for i in range(4)]
    d = np.min([np.hypot(X - a, Y - b) for a, b in poles], axis=0)
    assert d.max() <= 0.15 + 1e-6
    s.Shape = "dipolecirc"
    s.SigmaIn, s.SigmaOut, s.OpenAngle, s.RotateAngle = 0.4, 0.9, math.pi / 3, 0.0
    X, Y, V = O.source_points(s)
    assert np.all(np.abs(np.cos(np.arctan2(Y, X))) >= math.cos(math.pi / 6) - 1e-6)


def test_psf_blur_keeps_mass_in_interior():
    s = annular(P=41, so=0.5, si=0.2)
    X0, Y0, V0 = O.source_points(s)
    s.PSFEnable = True
    s.PSFSigma = 0.05
    X1, Y1, V1 = O.source_points(s)
    # normalised blur: mass is conserved up to the zeroed centre pixel and the 1e-5 cut
    assert abs(V1.sum() - V0.sum()) / V0.sum() < 0.02
    assert len(V1) > len(V0)


# -------------------------------------------------------------------- zernike
# Textbook Fringe polynomials written out explicitly (rho power coefficients),
# independent of the oracle's factorial-sum formulation.
TEXTBOOK = {
    3: (0, {2: 2, 0: -1}), 8: (0, {4: 6, 2: -6, 0: 1}), 6: (1, {3: 3, 1: -2}),
    11: (2, {4: 4, 2: -3}), 13: (1, {5: 10, 3: -12, 1: 3}), 15: (0, {6: 20, 4: -30, 2: 12, 0: -1}),
    18: (3, {5: 5, 3: -4}), 20: (2, {6: 15, 4: -20, 2: 6}), 22: (1, {7: 35, 5: -60, 3: 30, 1: -4}),
    24: (0, {8: 70, 6: -140, 4: 90, 2: -20, 0: 1}), 27: (4, {6: 6, 4: -5}),
    29: (3, {7: 21, 5: -30, 3: 10}), 31: (2, {8: 56, 6: -105, 4: 60, 2: -10}),
    33: (1, {9: 126, 7: -280, 5: 210, 3: -60, 1: 5}),
    35: (0, {10: 252, 8: -630, 6: 560, 4: -210, 2: 30, 0: -1}),
    36: (0, {12: 924, 10: -2772, 8: 3150, 6: -1680, 4: 420, 2: -42, 0: 1}),
}


@pytest.mark.parametrize("idx", sorted(TEXTBOOK))
def test_zernike_textbook_polynomials(idx):
    m, poly = TEXTBOOK[idx]
    rng = np.random.default_rng(idx)
    rho = rng.random(50)
    th = rng.uniform(-np.pi, np.pi, 50)
    c = np.zeros(37)
    c[idx] = 1.0
    expect = sum(a * rho ** p for p, a in poly.items()) * (np.cos(m * th) if m else 1.0)
    assert np.allclose(O.zernike_fringe(rho, th, c), expect, atol=1e-12)


def test_zernike_orthogonality_and_unit_edge():
    # orthogonality on the unit disk pins the (n, m, kind) table; Gauss-Legendre
    # in rho and a uniform theta grid integrate these polynomials exactly
    x, wx = np.polynomial.legendre.leggauss(24)
    r = 0.5 * (x + 1)
    nt = 64
    t = np.arange(nt) * 2 * np.pi / nt
    R, T = np.meshgrid(r, t, indexing="ij")
    Z = []
    for i in range(37):
        c = np.zeros(37)
        c[i] = 1
        Z.append(O.zernike_fringe(R, T, c))
    w = (0.5 * wx * r)[:, None] * (2 * np.pi / nt)
    G = np.array([[np.sum(a * b * w) for b in Z] for a in Z])
    off = G - np.diag(np.diag(G))
    assert np.max(np.abs(off)) < 1e-12
    # diagonal: integral of Z^2 over the disk = pi / (n + 1), halved for m > 0
    expect = [math.pi / (n + 1) / (2 if m else 1) for n, m, k in O.FRINGE_NM]
    assert np.allclose(np.diag(G), expect, rtol=1e-12)
    # R_n^m(1) = 1: |Z_i(1, theta)| reaches 1
    for i, (n, m, kind) in enumerate(O.FRINGE_NM):
        c = np.zeros(37)
        c[i] = 1
        th0 = 0.0 if kind >= 0 else np.pi / (2 * max(m, 1))
        assert abs(O.zernike_fringe(np.array([1.0]), np.array([th0]), c)[0] - 1.0) < 1e-12


# ------------------------------------------------------ characteristic matrix
def test_characteristic_matrix_transverse_unit():
    rng = np.random.default_rng(1)
    f, g = rng.uniform(-0.7, 0.7, 200), rng.uniform(-0.7, 0.7, 200)
    NA, n = 1.35, 1.44
    mxx, myx, mxy, myy, mxz, myz = O.characteristic_matrix(f, g, NA, n)
    a, b = NA / n * f, NA / n * g
    gam = np.sqrt(1 - a ** 2 - b ** 2)
    for px, py in [(1, 0), (0, 1), (0.6, 0.8j)]:
        ex, ey, ez = px * mxx + py * myx, px * mxy + py * myy, px * mxz + py * myz
        assert np.max(np.abs(a * ex + b * ey + gam * ez)) < 1e-12
    assert np.allclose(np.abs(mxx) ** 2 + np.abs(mxy) ** 2 + np.abs(mxz) ** 2, 1, atol=1e-12)
    # normal incidence: identity in x, y, no z component
    m0 = O.characteristic_matrix(np.zeros(1), np.zeros(1), NA, n)
    assert np.allclose([v[0] for v in m0], [1, 0, 0, 1, 0, 0])


# ------------------------------------------------------------------ spectrum
@pytest.mark.parametrize("ny,nx", [(6, 5), (8, 8), (7, 4)])
def test_spectrum_bruteforce(ny, nx):
    rng = np.random.default_rng(ny * 10 + nx)
    h = rng.standard_normal((ny, nx, 3)) + 1j * rng.standard_normal((ny, nx, 3))
    px, py = 700.0 * 0.85 / 193, 900.0 * 0.85 / 193
    hk, f, g, _ = O.spectrum(h, 700.0, 900.0, 0.85, 193.0)
    # Hk at centred index (kr, kc) = px py / (ny nx) * sum_m h[m] exp(+2 pi i k.(m - N//2)/N)
    for jr in range(ny):
        for jc in range(nx):
            kr, kc = jr - ny // 2, jc - nx // 2
            m_r = np.arange(ny) - ny // 2
            m_c = np.arange(nx) - nx // 2
            # ifftshift moves sample m to position m - N//2
            ph = np.exp(2j * np.pi * (kr * m_r[:, None] / ny + kc * m_c[None, :] / nx))
            h_sh = np.roll(h, (-(ny // 2), -(nx // 2)), axis=(0, 1))
            pos_r = (np.arange(ny) - ny // 2) % ny
            pos_c = (np.arange(nx) - nx // 2) % nx
            val = px * py / (ny * nx) * np.einsum("ab,abc->c", ph, h_sh[np.ix_(pos_r, pos_c)])
            assert np.allclose(hk[jr, jc], val, atol=1e-12)
    assert np.allclose(f, np.arange(-(nx // 2), (nx + 1) // 2) / px)
    assert np.allclose(g, np.arange(-(ny // 2), (ny + 1) // 2) / py)


def test_spectrum_dc_and_pupil_crop():
    rng = np.random.default_rng(3)
    h = rng.random((64, 48, 3))
    hk, f, g, pupil = O.spectrum(h, 1000.0, 1200.0, 0.85, 193.0)
    px, py = 1000 * 0.85 / 193, 1200 * 0.85 / 193
    assert np.allclose(hk[32, 24], h.mean(axis=(0, 1)) * px * py)
    ox, oy = int(px) + 3, int(py) + 3
    assert pupil.shape == (2 * oy, 2 * ox, 3)
    assert np.isnan(pupil[0, 0, 0]) and not np.isnan(pupil[oy, ox, 0])


# ------------------------------------------------------------------- Abbe
def _direct_abbe(mask, X, Y, V, p_x, p_y, NA, M, n):
    """Independent statement: I[j] = (n/W)/(NyNx)^2 sum_s V_s
    |sum_kappa P_s(kappa) Mhat(kappa) exp(-2 pi i kappa.j / N)|^2,
    Mhat(kappa) = sum_m h[m] exp(+2 pi i kappa.m / N), brute force."""
    ny, nx = mask.shape
    kr = np.arange(-(ny // 2), ny - ny // 2)
    kc = np.arange(-(nx // 2), nx - nx // 2)
    mr, mc = np.arange(ny), np.arange(nx)
    Fr = np.exp(2j * np.pi * np.outer(kr, mr) / ny)
    Fc = np.exp(2j * np.pi * np.outer(kc, mc) / nx)
    mhat = Fr @ mask @ Fc.T
    I = np.zeros((ny, nx))
    for x, y, v in zip(X, Y, V):
        f = kr[:, None] / p_y + x
        g = kc[None, :] / p_x + y
        r2 = f ** 2 + g ** 2
        with np.errstate(invalid="ignore"):
            ob = ((1 - (NA / M) ** 2 * r2) / (1 - NA ** 2 * r2)) ** 0.25
        P = np.where(r2 <= 1, ob, 0)
        E = Fr.conj().T @ (P * mhat) @ Fc.conj()
        I += v * np.abs(E) ** 2
    return n / V.sum() / (ny * nx) ** 2 * I


def test_abbe_bruteforce_small():
    rng = np.random.default_rng(0)
    mask = (rng.random((12, 10)) < 0.5).astype(float)
    src = annular(P=9)
    proj = projector()
    f = np.zeros((12, 10, 3))
    f[..., 0] = mask
    sc = scatter(f, 900.0, 700.0)
    I, *_ = O.abbe_image(src, sc, proj, numerics())
    X, Y, V = O.source_points(src)
    ref = _direct_abbe(mask, X, Y, V, 900 * 0.85 / 193, 700 * 0.85 / 193, 0.85, 100, 1.0)
    assert np.max(np.abs(I[0] - ref)) < 1e-12 * np.max(ref) + 1e-15


def test_abbe_clear_field_closed_form():
    src, proj = annular(), projector()
    f = np.zeros((32, 32, 3))
    f[..., 0] = 1
    I, *_ = O.abbe_image(src, scatter(f, 1000.0, 1000.0), proj, numerics())
    assert np.allclose(I, 1.21476847, atol=1e-7)
    assert abs(O.clear_field_intensity(src, proj) - 1.21476847) < 1e-7
    # per source point: a clear mask passes only the DC order, so each
    # unweighted term is the constant squared obliquity at rho = |s|
    J, V, n = O.abbe_source_terms(src, scatter(f, 1000.0, 1000.0), proj, numerics())
    X, Y, _ = O.source_points(src)
    r2 = X ** 2 + Y ** 2
    ob2 = np.sqrt((1 - (0.85 / 100) ** 2 * r2) / (1 - 0.85 ** 2 * r2))
    assert np.allclose(J[0], ob2[:, None, None], atol=1e-12)
    assert np.allclose(n / V.sum() * np.einsum("s,sij->ij", V, J[0]), I[0], atol=1e-12)


def test_abbe_dark_and_shift_equivariance():
    src, proj = annular(P=11), projector()
    f = np.zeros((32, 24, 3))
    I, *_ = O.abbe_image(src, scatter(f, 1000.0, 800.0), proj, numerics())
    assert np.all(I == 0)
    rng = np.random.default_rng(5)
    f[..., 0] = rng.random((32, 24)) < 0.4
    I1, *_ = O.abbe_image(src, scatter(f, 1000.0, 800.0), proj, numerics())
    I2, *_ = O.abbe_image(src, scatter(np.roll(f, (3, -5), axis=(0, 1)), 1000.0, 800.0), proj, numerics())
    assert np.allclose(np.roll(I1, (3, -5), axis=(1, 2)), I2, atol=1e-12)


def test_abbe_mirror_symmetry():
    # symmetric source and a mask symmetric under column reversal about the grid
    # centre give an image with the same symmetry
    src, proj = annular(P=11), projector()
    m = np.zeros((32, 32))
    m[10:20, 8:13] = 1
    m = np.maximum(m, m[:, (-np.arange(32)) % 32])
    f = np.zeros((32, 32, 3))
    f[..., 0] = m
    I, *_ = O.abbe_image(src, scatter(f, 1000.0, 1000.0), proj, numerics())
    assert np.allclose(I[0], I[0][:, (-np.arange(32)) % 32], atol=1e-12)


# ------------------------------------------------------------------------ TCC
def test_tcc_definition_hermitian_psd_trace():
    src, proj, num = annular(P=11), projector(), numerics()
    tcc, fg = O.tcc_matrix(src, [900.0, 700.0], proj, 0.0, num)
    X, Y, V = O.source_points(src)
    f, g = O.tcc_grid([900.0, 700.0], 0.85, 193.0)
    assert fg == [len(g), len(f)]
    # brute-force definition with explicit loops
    pts = [(a, b) for a in f for b in g]
    G = np.zeros((len(pts), len(V)), dtype=complex)
    for mi, (a, b) in enumerate(pts):
        for j in range(len(V)):
            r2 = (a + X[j]) ** 2 + (b + Y[j]) ** 2
            if r2 <= 1:
                G[mi, j] = ((1 - (0.85 / 100) ** 2 * r2) / (1 - 0.85 ** 2 * r2)) ** 0.25
    ref = sum(V[j] * np.outer(G[:, j], G[:, j].conj()) for j in range(len(V))) / V.sum()
    assert np.allclose(tcc, ref, atol=1e-13)
    assert np.allclose(tcc, tcc.conj().T)
    assert np.linalg.eigvalsh(tcc).min() > -1e-12
    assert abs(np.trace(tcc).real - np.sum(V[None, :] * np.abs(G) ** 2) / V.sum()) < 1e-12
    c = len(pts) // 2
    assert abs(tcc[c, c].real - O.clear_field_intensity(src, proj)) < 1e-12


def test_tcc_grid_rule_float32():
    f, g = O.tcc_grid([1000.0, 1000.0], 0.85, 193.0)
    assert len(f) == 19  # ceil(2 * 4.404) = 9
    f, g = O.tcc_grid([4096.0, 4096.0], 0.85, 193.0)
    assert len(f) == 75
    f, g = O.tcc_grid([4096.0, 4096.0], 0.85, 193.0, halfwidth=50)
    assert len(f) == 101


# -------------------------------------------------------------- decomposition
def test_decompose_full_rank_reconstructs_tcc():
    src, proj = annular(P=9), projector()
    num = numerics(order=10 ** 6)
    tcc, fg = O.tcc_matrix(src, [600.0, 600.0], proj, 0.0, num)
    s, U = O.eig_desc(tcc)
    K = O.socs_count(s, tcc.shape[0], num)
    rec = (U[:, :K] * s[:K]) @ U[:, :K].conj().T
    assert np.allclose(rec, tcc, atol=1e-12)
    assert abs(s.sum() - np.trace(tcc).real) < 1e-12


def test_lowrank_route_equals_dense():
    src, proj = annular(P=11), projector()
    num = numerics(order=12)
    Gs, V, f, g = O.shifted_pupil_matrix(src, [800.0, 800.0], proj, 0.0, num)
    tcc, fg = O.tcc_matrix(src, [800.0, 800.0], proj, 0.0, num)
    k_dense = O.decompose_tcc_socs(tcc, fg, [48, 48], num)
    k_low = O.decompose_lowrank(Gs, V, 1.0, [len(g), len(f)], [48, 48], num)
    # kernels agree up to per-kernel phase: compare projectors sum_k phi phi^H
    Pd = np.einsum("abk,cdk->abcd", k_dense, k_dense.conj())
    Pl = np.einsum("abk,cdk->abcd", k_low, k_low.conj())
    assert np.max(np.abs(Pd - Pl)) < 1e-12 * np.max(np.abs(Pd)) + 1e-14


def test_socs_count_threshold_semantics():
    s = np.array([5.0, 3.0, 1.0, 0.5, 0.5])
    num = numerics(kind="threshold")
    num.Hopkins_Threshold = 0.8     # 0.8 * 10 = 8: running sums 5, 8, 9 -> first > 8 at i = 2
    assert O.socs_count(s, 5, num) == 2
    num.Hopkins_Threshold = 0.95    # 9.5: sums 5, 8, 9, 9.5 (not > 9.5) -> none -> M - 1
    assert O.socs_count(s, 5, num) == 4
    with pytest.raises(ValueError):
        O.socs_count(s, 5, numerics(kind="bogus"))


# ----------------------------------------------------------------- SOCS image
def test_full_order_socs_equals_abbe_even_grid():
    src, proj = annular(P=9), projector()
    num = numerics(order=10 ** 6)
    rng = np.random.default_rng(2)
    f = np.zeros((24, 32, 3))
    f[..., 0] = rng.random((24, 32)) < 0.5
    sc = scatter(f, 1100.0, 900.0)
    hop, _ = O.hopkins_image(src, sc, proj, num)
    ab, *_ = O.abbe_image(src, sc, proj, num)
    assert np.max(np.abs(hop - ab)) < 1e-12 * np.max(ab)


def test_socs_clear_field_full_order():
    src, proj = annular(P=9), projector()
    num = numerics(order=10 ** 6)
    f = np.zeros((16, 16, 3))
    f[..., 0] = 1
    hop, _ = O.hopkins_image(src, scatter(f, 700.0, 700.0), proj, num)
    assert np.allclose(hop, O.clear_field_intensity(src, proj), atol=1e-12)


# --------------------------------------------------------------- gradient pins
def _random_kernels(ny, nx, K, w, seed):
    """Dense (ny, nx, K) kernels supported on a w x w window around DC, in the
    fftshifted embedding of DecomposeTCC_SOCS.py:27-31."""
    rng = np.random.default_rng(seed)
    k = np.zeros((ny, nx, K), dtype=np.complex128)
    r0, c0 = ny // 2 - w // 2, nx // 2 - w // 2
    k[r0:r0 + w, c0:c0 + w] = rng.standard_normal((w, w, K)) + 1j * rng.standard_normal((w, w, K))
    return np.fft.fftshift(k, axes=(0, 1))


def test_torch_socs_restatement_equals_numpy_oracle():
    from oracle.socs_torch import socs_image_t
    ny, nx = 24, 32
    kern = _random_kernels(ny, nx, 3, 7, seed=4)
    mask = np.random.default_rng(5).random((ny, nx))
    f3 = np.zeros((ny, nx, 3))
    f3[..., 0] = mask
    ref = O.socs_image(f3, kern, 900.0, 700.0, 0.8, 193.0)
    got = socs_image_t(torch.as_tensor(mask), torch.as_tensor(kern), 900.0, 700.0, 0.8, 193.0)
    assert np.max(np.abs(got.numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_autograd_gradient_matches_central_differences():
    """sum(g * I) is quadratic in the mask, so a central difference is exact up
    to rounding: the float64 autograd gradient must match it to ~1e-9."""
    from oracle.socs_torch import socs_mask_gradient
    ny, nx = 16, 20
    kern = _random_kernels(ny, nx, 4, 5, seed=6)
    rng = np.random.default_rng(7)
    mask = rng.random((ny, nx))
    g = rng.standard_normal((ny, nx))
    args = (1000.0, 800.0, 0.85, 193.0)
    grad = socs_mask_gradient(mask, g, kern, *args)

    def loss(m):
        f3 = np.zeros((ny, nx, 3))
        f3[..., 0] = m
        return float(np.sum(g * O.socs_image(f3, kern, *args)))

    eps = 1e-3
    for (y, x) in [(0, 0), (3, 7), (15, 19), (8, 1)]:
        mp, mm = mask.copy(), mask.copy()
        mp[y, x] += eps
        mm[y, x] -= eps
        fd = (loss(mp) - loss(mm)) / (2 * eps)
        assert abs(fd - grad[y, x]) <= 1e-9 * np.max(np.abs(grad)), (y, x, fd, grad[y, x])


# ------------------------------------------------------- Floquet-mode SOCS pins
def _floquet_kernel(n0, n1, K, entries):
    """Kernels in DecomposeTCC_SOCS's layout (transposed, fftshifted) holding
    entries {(ox, oy, k): value} at Floquet orders (ox, oy)."""
    kc = np.zeros((n0, n1, K), dtype=np.complex128)
    for (ox, oy, k), v in entries.items():
        kc[ox + n0 // 2, oy + n1 // 2, k] = v
    return np.transpose(np.fft.fftshift(kc, axes=(0, 1)), (1, 0, 2))


def test_calculate_socs_single_order_constant():
    Ox, Oy = 2, 3
    n0, n1 = 2 * Ox + 1, 2 * Oy + 1
    fm = np.zeros((n0, n1, 3), dtype=np.complex128)
    fm[Ox + 1, Oy - 2, 1] = 0.5 - 1.5j
    kern = _floquet_kernel(n0, n1, 2, {(1, -2, 1): 2.0 + 1.0j})
    I = O.calculate_socs(fm, kern, 9, 7, Ox, Oy)
    expect = abs((0.5 - 1.5j) * (2.0 + 1.0j)) ** 2
    assert I.shape == (9, 7)
    assert np.allclose(I, expect, rtol=1e-13)


def test_calculate_socs_two_order_fringe_closed_form():
    """I = |a1|^2 + |a2|^2 + 2 |a1 a2| cos(2 pi (d_x (ix + 1/2)/Gx + d_y (iy + 1/2)/Gy) + arg a1 - arg a2),
    d = o1 - o2: pins the phase sign, which array axis carries order_x, and
    the half-pixel sample offset."""
    Ox, Oy, Gx, Gy = 3, 2, 11, 8
    n0, n1 = 2 * Ox + 1, 2 * Oy + 1
    o1, o2 = (2, -1), (-1, 1)
    f1, f2 = 0.7 + 0.2j, -0.3 + 0.9j
    p1, p2 = 1.1 - 0.4j, 0.6 + 0.5j
    fm = np.zeros((n0, n1, 3), dtype=np.complex128)
    fm[o1[0] + Ox, o1[1] + Oy, 2] = f1
    fm[o2[0] + Ox, o2[1] + Oy, 2] = f2
    kern = _floquet_kernel(n0, n1, 1, {(o1[0], o1[1], 0): p1, (o2[0], o2[1], 0): p2})
    I = O.calculate_socs(fm, kern, Gx, Gy, Ox, Oy)
    a1, a2 = f1 * p1, f2 * p2
    ix = (np.arange(Gx)[:, None] + 0.5) / Gx
    iy = (np.arange(Gy)[None, :] + 0.5) / Gy
    ph = 2 * np.pi * ((o1[0] - o2[0]) * ix + (o1[1] - o2[1]) * iy) + np.angle(a1) - np.angle(a2)
    expect = abs(a1) ** 2 + abs(a2) ** 2 + 2 * abs(a1 * a2) * np.cos(ph)
    assert np.max(np.abs(I - expect)) < 1e-12


def test_calculate_socs_sums_channels_and_kernels():
    rng = np.random.default_rng(21)
    Ox, Oy = 2, 2
    n = 2 * Ox + 1
    fm = rng.standard_normal((n, n, 3)) + 1j * rng.standard_normal((n, n, 3))
    kern = rng.standard_normal((n, n, 3)) + 1j * rng.standard_normal((n, n, 3))
    total = O.calculate_socs(fm, kern, 6, 10, Ox, Oy)
    parts = 0
    for c in range(3):
        for k in range(3):
            f1 = np.zeros_like(fm)
            f1[..., c] = fm[..., c]
            parts = parts + O.calculate_socs(f1, kern[..., k:k + 1], 6, 10, Ox, Oy)
    assert np.max(np.abs(total - parts)) < 1e-11 * np.max(total)
