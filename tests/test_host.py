"""CPU tests of the host side: the C ABI library loads and exports every
declared entry point, the parameter objects match the reference's attributes
and the oracle's source selection, and host-side geometry logic."""
import math
import os
import re

import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import annular, quasar

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "opticalimaging_amd.h")).read()
    return sorted(set(re.findall(r"\b(oi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from opticalimaging_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_native.SIGNATURES) == set(names)
    assert _native.load().oi_version() >= 10000


def test_error_string_and_status_mapping():
    from opticalimaging_amd import _native
    lib = _native.load()
    g = _native.Geometry(ny=0, nx=0)
    import ctypes
    h = ctypes.c_void_p()
    rc = lib.oi_plan_create(ctypes.byref(g), ctypes.byref(h))
    assert rc == _native.OI_ERR_INVALID
    with pytest.raises(ValueError, match="sizes must be positive"):
        _native.check(rc)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-device path")
def test_imaging_refuses_without_device():
    from opticalimaging_amd import CalculateAbbeImage, Scatter, workloads
    mask, src, proj, num, period = workloads.scalar_abbe_128()
    sc = Scatter()
    sc.ScatterField = workloads.field_from_mask(mask)
    with pytest.raises(RuntimeError, match="requires a HIP"):
        CalculateAbbeImage(src, sc, proj, num)


def test_parameter_defaults_match_reference():
    from opticalimaging_amd import Numerics, Projection, Scatter, Source
    s, p, sc, n = Source(), Projection(), Scatter(), Numerics()
    assert (s.PntNum, s.Wavelength, s.Shape, s.SigmaOut, s.SigmaIn) == (81, 365, "annular", 0.7, 0.1)
    assert (s.SigmaCenter, s.SigmaRadius, s.PoleNumber, s.PSFEnable, s.PSFSigma) == (0.5, 0.1, 2, False, 0.02)
    assert s.OpenAngle == math.pi / 6 and s.RotateAngle == 0 and s.PolarizationVector == [1.0, 0.0]
    assert p.Magnification == 100 and p.NA == 0.9 and p.IndexImage == 1.0
    assert p.Aberration_Zernike.shape == (37,) and p.FocusRange.dtype == torch.int64
    assert (sc.Period_X, sc.Period_Y) == (1000, 1000)
    assert (n.ImageCalculationMode, n.ImageCalculationMethod, n.Hopkins_SettingType) == \
        ("scalar", "hopkins", "order")
    assert (n.Hopkins_Order, n.Hopkins_Threshold) == (100, 0.95)


@pytest.mark.parametrize("shape", ["annular", "quasar", "dipolecirc", "multipole", "psf"])
def test_product_source_matches_oracle(shape):
    s = quasar(P=41) if shape == "quasar" else annular(P=33)
    if shape == "dipolecirc":
        s.Shape, s.OpenAngle, s.RotateAngle = "dipolecirc", math.pi / 4, 0.3
    if shape == "multipole":
        s.Shape, s.PoleNumber, s.RotateAngle = "multipole", 3, 0.2
    if shape == "psf":
        s.PSFEnable = True
    X, Y, V = O.source_points(s)
    d = s.Calc_SourceSimple()
    assert np.array_equal(d.X.double().numpy(), X)
    assert np.array_equal(d.Y.double().numpy(), Y)
    assert np.array_equal(d.Value.double().numpy(), V)
    assert s.source_data is d
    # as in the reference, the flattened arrays live on the SourceData class and
    # Calc_SourceSimple filters them in place (Source.py:51-53,293-295)
    from opticalimaging_amd.Source import SourceData
    assert d is SourceData and SourceData.Value.dim() == 1


def test_source_validation_errors():
    s = annular()
    s.Shape = "quasar"
    s.OpenAngle = 2.0
    with pytest.raises(ValueError):
        s.Calc_SourceSimple()
    s.Shape = "bogus"
    with pytest.raises(ValueError):
        s.Calc_SourceSimple()


def test_polarization_types():
    s = annular()
    th = torch.tensor([0.0, 0.5, 2.0])
    rho = torch.tensor([0.0, 0.6, 0.7])
    px, py = s.Calc_PolarizationMap(th, rho)
    assert torch.all(px == 1) and torch.all(py == 0)
    s.PolarizationType = "t_pol"
    px, py = s.Calc_PolarizationMap(th, rho)
    assert px[0] == 0 and py[0] == 0                     # zeroed at rho < eps
    assert torch.allclose(px[1:].real, torch.sin(th[1:])) and torch.allclose(py[1:].real, -torch.cos(th[1:]))
    ox, oy = O.polarization(s, th.numpy(), rho.numpy())
    ex, ey = s.polarization_f64(rho * torch.cos(th), rho * torch.sin(th))
    assert np.allclose(ex.numpy()[1:], ox[1:]) and np.allclose(ey.numpy()[1:], oy[1:])


def test_rotate_angle_rotates_wavefront():
    """W(rho, phi; RotateAngle(c, t)) = W(rho, phi + t; c)."""
    from opticalimaging_amd import Projection
    rng = np.random.default_rng(4)
    c = rng.normal(0, 0.1, 37).astype(np.float32)
    t = 0.37
    c2 = Projection.RotateAngle(torch.tensor(c), t).numpy()
    rho = rng.random(40)
    phi = rng.uniform(-np.pi, np.pi, 40)
    a = O.zernike_fringe(rho, phi, c2.astype(np.float64))
    b = O.zernike_fringe(rho, phi + t, c.astype(np.float64))
    assert np.allclose(a, b, atol=1e-5)


def test_tcc_halfwidths_and_abbe_window():
    from opticalimaging_amd import optics
    assert optics.tcc_halfwidths([1000.0, 1000.0], 0.85, 193.0) == (9, 9)
    assert optics.tcc_halfwidths([4096.0, 2048.0], 0.85, 193.0) == (37, 19)

    class N:
        Hopkins_GridHalfWidth = 50
    assert optics.tcc_halfwidths([4096.0, 4096.0], 0.85, 193.0, N) == (50, 50)

    X, Y, V = O.source_points(annular())
    src = optics.SourceSet(torch.tensor(X), torch.tensor(Y), torch.tensor(V), None, None)
    p = 4096 * 0.85 / 193
    k0r, wr, k0c, wc, br, bc = optics.abbe_window(src, p, p, 1024, 1024)
    # every shifted pupil |kappa/p + s| <= 1 lies inside the window, each within `band` indices
    for x in X:
        lo, hi = math.ceil(p * (-1 - x)), math.floor(p * (1 - x))
        assert k0r <= lo and hi <= k0r + wr - 1
        assert hi - lo + 1 <= br


def test_workloads_deterministic():
    from opticalimaging_amd import workloads
    a, per = workloads.socs_1024_masks(batch=2, n=256)
    b, _ = workloads.socs_1024_masks(batch=2, n=256)
    assert workloads.digest(a) == workloads.digest(b)
    assert per == 1024.0 and a.mean() >= 0.35
    m1, *_ = workloads.scalar_abbe_128(seed=1)
    assert m1.shape == (128, 128) and set(np.unique(m1)) <= {0.0, 1.0}
    m2, src, proj, num, period = workloads.vector_abbe_512(n=128)
    assert set(np.unique(m2)) <= {1.0, np.float32(workloads.ATT_PSM_AMPLITUDE)}
    assert num.ImageCalculationMode == "vector" and proj.Magnification == 1


def test_band_step_and_bit_decode_reject_arguments_before_any_device_call():
    """The argument checks of the band-form ILT entry points, the bit decode
    and the multi-focus pupils run on the host, ahead of any HIP call: a null
    pointer, a gradient buffer that is the logits buffer itself, a misaligned
    bit buffer, an empty focus list, zero pitches."""
    import ctypes
    from opticalimaging_amd import _native
    lib = _native.load()
    buf = (ctypes.c_float * 64)()
    a = ctypes.cast(buf, ctypes.c_void_p)
    b = ctypes.c_void_p(ctypes.addressof(buf) + 128)
    null = ctypes.c_void_p()
    assert lib.oi_ilt_target_band_bytes(null, 4) == 0
    assert lib.oi_ilt_target_band(null, a, 1, a, a, null) == _native.OI_ERR_INVALID
    assert lib.oi_socs_band_residual_adjoint(null, a, 1, a, 1, a, null, a, a, null) == \
        _native.OI_ERR_INVALID
    assert lib.oi_ilt_band_step(null, a, 1, a, 1, a, 0.1, null, a, b, null) == _native.OI_ERR_INVALID
    # a plan pointer the call never reads: the alias check comes first
    rc = lib.oi_ilt_band_step(a, a, 1, a, 1, a, 0.1, null, b, a, null)
    assert rc == _native.OI_ERR_INVALID
    with pytest.raises(ValueError, match="grad_theta must not be theta"):
        _native.check(rc)
    levels = (ctypes.c_float * 2)(0.0, 1.0)
    assert lib.oi_mask_decode_bits(null, 32, levels, a, null) == _native.OI_ERR_INVALID
    rc = lib.oi_mask_decode_bits(ctypes.c_void_p(ctypes.addressof(buf) + 2), 32, levels, a, null)
    assert rc == _native.OI_ERR_INVALID
    with pytest.raises(ValueError, match="4-byte"):
        _native.check(rc)
    opt = _native.Optics()
    zs = (ctypes.c_double * 2)(0.0, 50.0)
    assert lib.oi_pupil_kernels_foci(ctypes.byref(opt), None, 2, a, a, a, null, null, 4, 0, 3, 0, 3, a,
                                     null) == _native.OI_ERR_INVALID
    assert lib.oi_pupil_kernels_foci(ctypes.byref(opt), zs, 0, a, a, a, null, null, 4, 0, 3, 0, 3, a,
                                     null) == _native.OI_ERR_INVALID
    rc = lib.oi_pupil_kernels_foci(ctypes.byref(opt), zs, 2, a, a, a, null, null, 4, 0, 3, 0, 3, a, null)
    assert rc == _native.OI_ERR_INVALID                          # zero pitches and wavelength
    with pytest.raises(ValueError, match="pitches and wavelength"):
        _native.check(rc)
    al = ctypes.c_void_p((ctypes.addressof(buf) + 15) & ~15)
    assert lib.oi_mask_decode_bits(al, 0, levels, al, null) == _native.OI_OK   # nothing to decode
