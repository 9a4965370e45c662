"""HIP path vs the float64 oracle on the same seeded inputs.

Tolerances (peak-normalised max error, max|I - I_ref| / max|I_ref|):
  * fp32 imaging kernels (spectrum, SOCS image, Abbe): 1e-5.
  * TCC contraction: fp64 2e-7 (c64 output rounding), fp32 1e-5, split-bf16
    3-term 1e-4 (dropped lo*lo term ~2^-16 relative per product).
  * Images through the split-bf16 TCC + float64 eigen route: 1e-4.
"""
import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import (annular, binary_field, complex_field, numerics, projector, quasar,
                       rel_err, scatter)

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


# ------------------------------------------------------------------ spectrum
@pytest.mark.parametrize("ny,nx,px,py", [(48, 32, 1300.0, 900.0), (128, 128, 1000.0, 1000.0),
                                         (45, 75, 700.0, 1100.0), (200, 100, 3000.0, 1500.0)])
def test_spectrum_full(gpu, ny, nx, px, py):
    src, proj = annular(), projector()
    f = complex_field(ny, nx, seed=ny * nx)
    sc = scatter(f, px, py)
    hk, fx, gy, pupil = sc.CalculateSpectrum(proj, src)
    hk_o, f_o, g_o, pupil_o = O.spectrum(_np(f), px, py, proj.NA, src.Wavelength)
    print(f"spectrum {ny}x{nx} err", rel_err(_np(hk), hk_o))
    assert rel_err(_np(hk), hk_o) < 1e-5
    assert np.allclose(_np(fx), f_o, rtol=1e-6) and np.allclose(_np(gy), g_o, rtol=1e-6)
    pn, po = _np(pupil), pupil_o
    assert pn.shape == po.shape
    assert np.array_equal(np.isnan(pn), np.isnan(po))
    ok = ~np.isnan(po)
    if ok.any():
        assert np.max(np.abs(pn[ok] - po[ok])) <= 1e-5 * np.max(np.abs(hk_o))


# ----------------------------------------------------------------- SOCS image
SOCS_CASES = [
    # ny, nx, Period_X, Period_Y, order
    (128, 128, 1000.0, 1000.0, 24),      # config-1 optics, coarse grid + interpolation
    (96, 160, 1200.0, 800.0, 24),        # non-square grid and periods
    (32, 32, 1000.0, 1000.0, 30),        # coarse grid = image grid (direct)
    (45, 75, 700.0, 1100.0, 12),         # odd sizes (reference's one-bin kernel shift)
]


@pytest.mark.parametrize("ny,nx,pxx,pyy,order", SOCS_CASES)
def test_socs_image_dense_kernels(gpu, ny, nx, pxx, pyy, order):
    """Oracle kernels (float64 -> c64) fed to the HIP SOCS image."""
    src, proj, num = annular(), projector(), numerics(order=order)
    field = binary_field(ny, nx, seed=7 + ny)
    sc = scatter(field, pxx, pyy)
    tcc, fg = O.tcc_matrix(src, [pyy, pxx], proj, 0.0, num)
    kern = O.decompose_tcc_socs(tcc, fg, [nx, ny], num)
    ref = O.socs_image(_np(field), kern, pxx, pyy, proj.NA, src.Wavelength)
    from opticalimaging_amd import CalculateAerialImage_SOCS
    img, pupil = CalculateAerialImage_SOCS(sc, torch.as_tensor(kern.astype(np.complex64)), src, proj)
    err = rel_err(_np(img), ref)
    print(f"socs {ny}x{nx} err {err:.3e}")
    assert img.shape == (ny, nx) and img.dtype == torch.float32
    assert err < 1e-5


def test_socs_complex_three_channel(gpu):
    src, proj, num = annular(), projector(), numerics(order=16)
    field = complex_field(64, 64, seed=5)
    sc = scatter(field, 1000.0, 1000.0)
    tcc, fg = O.tcc_matrix(src, [1000.0, 1000.0], proj, 0.0, num)
    kern = O.decompose_tcc_socs(tcc, fg, [64, 64], num)
    ref = O.socs_image(_np(field), kern, 1000.0, 1000.0, proj.NA, src.Wavelength)
    from opticalimaging_amd import CalculateAerialImage_SOCS
    img, _ = CalculateAerialImage_SOCS(sc, torch.as_tensor(kern.astype(np.complex64)), src, proj)
    err = rel_err(_np(img), ref)
    print("socs 3-channel err", err)
    assert err < 1e-5


# ---------------------------------------------------------------------- TCC
@pytest.mark.parametrize("precision,tol", [("fp64", 2e-7), ("fp32", 1e-5), ("split_bf16", 1e-4)])
def test_tcc_scalar(gpu, precision, tol):
    src, proj, num = annular(), projector(), numerics(precision=precision)
    from opticalimaging_amd import CalculateTCCMatrix
    tcc, fg = CalculateTCCMatrix(src, [1000.0, 1000.0], proj, 0.0, num)
    ref, fg_o = O.tcc_matrix(src, [1000.0, 1000.0], proj, 0.0, num)
    assert list(fg) == list(fg_o)
    err = rel_err(_np(tcc), ref)
    print(f"tcc scalar {precision} err {err:.3e}")
    assert err < tol


def test_tcc_vector_aberrated(gpu):
    zern = np.random.default_rng(11).normal(0, 0.01, 37)
    src = quasar(P=21)
    proj = projector(NA=1.35, n=1.44, M=1, foci=(30.0,), zern=zern)
    num = numerics(mode="vector", precision="fp32")
    from opticalimaging_amd import CalculateTCCMatrix
    tcc, fg = CalculateTCCMatrix(src, [600.0, 500.0], proj, 30.0, num)
    ref, fg_o = O.tcc_matrix(src, [600.0, 500.0], proj, 30.0, num)
    assert list(fg) == list(fg_o)
    err = rel_err(_np(tcc), ref)
    print("tcc vector err", err)
    assert err < 1e-5


# ------------------------------------------------------------------- Abbe
def test_abbe_scalar_config1(gpu):
    from opticalimaging_amd import workloads, Scatter, CalculateAbbeImage
    mask, src, proj, num, period = workloads.scalar_abbe_128()
    field = workloads.field_from_mask(mask)
    sc = scatter(field, period, period)
    img, pupil = CalculateAbbeImage(src, sc, proj, num)
    ref, ix, iy, iz, pupil_o = O.abbe_image(src, sc, proj, num)
    err = rel_err(_np(img.Intensity), ref)
    print("abbe config-1 err", err)
    assert img.Intensity.shape == (1, 128, 128)
    assert err < 1e-5
    assert np.allclose(_np(img.ImageX), ix, atol=1e-3)


def test_abbe_vector_small(gpu):
    zern = np.random.default_rng(3).normal(0, 0.01, 37)
    src = quasar(P=21)
    proj = projector(NA=1.35, n=1.44, M=1, foci=(-60.0, 0.0, 60.0), zern=zern)
    num = numerics(mode="vector")
    rng = np.random.default_rng(9)
    m = np.where(rng.random((64, 64)) < 0.4, -np.sqrt(0.06), 1.0).astype(np.float32)
    from opticalimaging_amd import workloads, CalculateAbbeImage
    sc = scatter(workloads.field_from_mask(m), 256.0, 256.0)
    img, _ = CalculateAbbeImage(src, sc, proj, num)
    ref, *_ = O.abbe_image(src, sc, proj, num)
    err = rel_err(_np(img.Intensity), ref)
    print("abbe vector err", err)
    assert err < 1e-5


def test_abbe_pupil_single_real_channel_not_first(gpu):
    """One real live channel at index 1 (channels 0 and 2 zero) travels as a
    float32 mask; the PupilImage must carry it in channel 1, NaN outside the
    pupil, as Scatter.py:25-35 builds it."""
    from opticalimaging_amd import workloads, CalculateAbbeImage
    mask, src, proj, num, period = workloads.scalar_abbe_128()
    field = np.zeros(mask.shape + (3,), dtype=np.complex64)
    field[:, :, 1] = mask
    sc = scatter(field, period, period)
    img, pupil = CalculateAbbeImage(src, sc, proj, num)
    ref, *_, pupil_o = O.abbe_image(src, sc, proj, num)
    assert rel_err(_np(img.Intensity), ref) < 1e-5
    pn = _np(pupil)
    assert pn.shape == pupil_o.shape
    assert np.array_equal(np.isnan(pn), np.isnan(pupil_o))
    fin = ~np.isnan(pupil_o)
    scale = np.abs(pupil_o[fin]).max()
    assert np.abs(pn[fin] - pupil_o[fin]).max() < 1e-5 * scale
    assert np.abs(pn[:, :, 1][fin[:, :, 1]]).max() > 0.1 * scale
    assert np.all(pn[:, :, 0][fin[:, :, 0]] == 0) and np.all(pn[:, :, 2][fin[:, :, 2]] == 0)


def test_abbe_complex_field_nonsquare(gpu):
    src, proj, num = annular(P=11), projector(), numerics()
    field = complex_field(48, 32, seed=21)
    sc = scatter(field, 1300.0, 900.0)
    from opticalimaging_amd import CalculateAbbeImage
    img, pupil = CalculateAbbeImage(src, sc, proj, num)
    ref, *_, pupil_o = O.abbe_image(src, sc, proj, num)
    err = rel_err(_np(img.Intensity), ref)
    print("abbe complex nonsquare err", err)
    assert err < 1e-5


# ---------------------------------------------------------------- Hopkins
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-5), ("split_bf16", 1e-4)])
def test_hopkins_config1(gpu, precision, tol):
    from opticalimaging_amd import workloads, CalculateHopkinsImage
    mask, src, proj, num, period = workloads.scalar_abbe_128()
    num.Hopkins_Order = 24
    num.TCC_Precision = precision
    sc = scatter(workloads.field_from_mask(mask), period, period)
    img, _ = CalculateHopkinsImage(src, sc, proj, num)
    ref, _ = O.hopkins_image(src, sc, proj, num, lowrank=True)
    err = rel_err(_np(img.Intensity), ref)
    print(f"hopkins {precision} err {err:.3e}")
    assert err < tol


def test_full_order_socs_equals_abbe(gpu):
    """Full-order SOCS reproduces Abbe (exact identity for even grids)."""
    from opticalimaging_amd import workloads, CalculateHopkinsImage, CalculateAbbeImage
    mask, src, proj, num, period = workloads.scalar_abbe_128()
    num.Hopkins_Order = 10 ** 6
    num.TCC_Precision = "fp64"
    sc = scatter(workloads.field_from_mask(mask), period, period)
    hop, _ = CalculateHopkinsImage(src, sc, proj, num)
    abbe, _ = CalculateAbbeImage(src, sc, proj, num)
    err = rel_err(_np(hop.Intensity), _np(abbe.Intensity))
    print("full-order socs vs abbe", err)
    assert err < 1e-5


def test_batched_socs_matches_single(gpu):
    from opticalimaging_amd import engine
    src, proj, num = annular(), projector(), numerics(order=24)
    tcc, fg = O.tcc_matrix(src, [1000.0, 1000.0], proj, 0.0, num)
    kern = torch.as_tensor(O.decompose_tcc_socs(tcc, fg, [128, 128], num).astype(np.complex64))
    ks = engine.KernelSet.from_dense(kern)
    rng = np.random.default_rng(4)
    masks = torch.as_tensor((rng.random((5, 128, 128)) < 0.5).astype(np.float32)).cuda()
    plan = engine.Plan(128, 128, ks.k0r, ks.wr, ks.k0c, ks.wc, nchan=1, max_batch=5)
    batch = engine.socs_image(masks, ks, plan=plan)
    for b in range(5):
        one = engine.socs_image(masks[b:b + 1].contiguous(), ks, plan=plan)
        assert torch.equal(one[0], batch[b])


# ------------------------------------------------ standalone optics operators
def test_zernike_device_matches_oracle(gpu):
    """Projection.CalculateAberrationFast (oi_zernike) against the float64
    Fringe Zernike sum on random pupil points, all 37 terms; a rotated call
    equals the unrotated wavefront at theta + Orientation."""
    from opticalimaging_amd import Projection
    rng = np.random.default_rng(7)
    p = Projection()
    c = rng.normal(0.0, 0.05, 37)
    p.Aberration_Zernike = torch.tensor(c, dtype=torch.float32)
    rho = np.sqrt(rng.random(4000)).astype(np.float32)
    th = (rng.random(4000) * 2 * np.pi - np.pi).astype(np.float32)
    w = _np(p.CalculateAberrationFast(torch.tensor(rho).cuda(), torch.tensor(th).cuda(), 0.0))
    ref = O.zernike_fringe(rho.astype(np.float64), th.astype(np.float64),
                           np.asarray(p.Aberration_Zernike, dtype=np.float64))
    err = float(np.max(np.abs(w - ref)) / np.max(np.abs(ref)))
    print(f"zernike device vs oracle: {err:.2e}")
    assert err < 1e-5
    phi = 0.37
    wr = _np(p.CalculateAberrationFast(torch.tensor(rho).cuda(), torch.tensor(th).cuda(), phi))
    ref_r = O.zernike_fringe(rho.astype(np.float64), th.astype(np.float64) + phi,
                             np.asarray(p.Aberration_Zernike, dtype=np.float64))
    err_r = float(np.max(np.abs(wr - ref_r)) / np.max(np.abs(ref_r)))
    print(f"rotated zernike device vs oracle at theta + phi: {err_r:.2e}")
    assert err_r < 1e-5


def test_characteristic_matrix_device_matches_oracle(gpu):
    """CalculateCharacteristicMatrix (oi_char_matrix) against the complex128
    closed form, inside and outside the pupil (complex gamma)."""
    from opticalimaging_amd import CalculateCharacteristicMatrix
    rng = np.random.default_rng(3)
    f = (rng.random(3000) * 2.4 - 1.2).astype(np.float32)
    g = (rng.random(3000) * 2.4 - 1.2).astype(np.float32)
    got = CalculateCharacteristicMatrix(torch.tensor(f).cuda(), torch.tensor(g).cuda(), 1.35, 1.44)
    ref = O.characteristic_matrix(f.astype(np.float64), g.astype(np.float64), 1.35, 1.44)
    for name, a, b in zip(("Mxx", "Myx", "Mxy", "Myy", "Mxz", "Myz"), got, ref):
        err = float(np.max(np.abs(_np(a) - b)) / np.max(np.abs(b)))
        print(f"{name}: {err:.2e}")
        assert err < 1e-5
