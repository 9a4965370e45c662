"""Golden vectors written by the reference implementation itself
(tests/golden/make_golden.py, reference_small.npz): 48 x 32 and 48 x 48 grids
with Period_X != Period_Y, Zernike aberrations, defocus, scalar mode on a real
mask and vector mode (quasar, immersion, PolarizationVector [0.6, 0.8]) on a
complex 3-channel field.  They pin the axis/period/source pairing, the vector
path and the PupilImage crop, which no closed form fixes.

The reference computes in complex64, so agreement is to its float32 rounding:
tolerance 2e-5 of the peak (observed 1e-6 .. 7e-6)."""
import math
import os

import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import rel_err

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                            "reference_small.npz"))
TOL = 2e-5


def setup(mode, shape="annular", M=100, NA=0.85, n=1.0, foci=(0,), order=12):
    from opticalimaging_amd import Numerics, Projection, Source
    s = Source()
    s.PntNum = 11
    s.Wavelength = 193
    s.Shape = shape
    s.SigmaOut, s.SigmaIn = 0.8, 0.5
    if shape == "quasar":
        s.SigmaOut, s.SigmaIn, s.OpenAngle = 0.9, 0.6, math.pi / 6
    s.PolarizationVector = [0.6, 0.8]
    p = Projection()
    p.NA, p.IndexImage, p.Magnification = NA, n, M
    p.Aberration_Zernike = torch.tensor(GOLD["zern"])
    p.FocusRange = torch.tensor(list(foci), dtype=torch.float32)
    num = Numerics()
    num.ImageCalculationMode = mode
    num.Hopkins_SettingType = "order"
    num.Hopkins_Order = order
    return s, p, num


def scalar_case():
    from opticalimaging_amd import Scatter
    s, p, num = setup("scalar", foci=(0, 40))
    sc = Scatter()
    sc.Period_X, sc.Period_Y = GOLD["per"]
    f = np.zeros((48, 32, 3), dtype=np.float32)
    f[..., 0] = GOLD["mask"]
    sc.ScatterField = torch.tensor(f)
    return s, p, num, sc


def vector_case():
    from opticalimaging_amd import Scatter
    s, p, num = setup("vector", shape="quasar", M=1, NA=1.2, n=1.44, foci=(-30,))
    sc = Scatter()
    sc.Period_X, sc.Period_Y = GOLD["per"]
    sc.ScatterField = torch.tensor(GOLD["cfield"])
    return s, p, num, sc


@pytest.mark.parametrize("case,key", [(scalar_case, "scalar"), (vector_case, "vector")])
def test_oracle_matches_reference(case, key):
    s, p, num, sc = case()
    A, *_, pupil = O.abbe_image(s, sc, p, num)
    assert rel_err(A, GOLD["abbe_" + key]) < TOL
    assert np.array_equal(np.isnan(pupil), np.isnan(GOLD["pupil_" + key]))
    ok = ~np.isnan(pupil)
    assert np.max(np.abs(pupil[ok] - GOLD["pupil_" + key][ok])) < TOL * np.max(np.abs(pupil[ok]))
    H, _ = O.hopkins_image(s, sc, p, num)
    assert rel_err(H, GOLD["hopkins_" + key]) < TOL


def test_oracle_tcc_matches_reference():
    s, p, num, sc = scalar_case()
    T, fg = O.tcc_matrix(s, [sc.Period_Y, sc.Period_X], p, 40.0, num)
    assert list(fg) == list(GOLD["tcc_scalar_fg"])
    assert rel_err(T, GOLD["tcc_scalar"]) < TOL


@pytest.mark.gpu
@pytest.mark.parametrize("case,key", [(scalar_case, "scalar"), (vector_case, "vector")])
def test_hip_matches_reference(gpu, case, key):
    from opticalimaging_amd import CalculateAbbeImage, CalculateHopkinsImage
    s, p, num, sc = case()
    num.TCC_Precision = "fp32"
    a, pupil = CalculateAbbeImage(s, sc, p, num)
    err_a = rel_err(a.Intensity.numpy(), GOLD["abbe_" + key])
    pn = pupil.numpy()
    assert np.array_equal(np.isnan(pn), np.isnan(GOLD["pupil_" + key]))
    ok = ~np.isnan(pn)
    assert np.max(np.abs(pn[ok] - GOLD["pupil_" + key][ok])) < TOL * np.max(np.abs(pn[ok]))
    # a second call takes the pupil's per-grid tables (and the source set) from the caches
    a2, pupil2 = CalculateAbbeImage(s, sc, p, num)
    assert np.array_equal(pupil2.numpy(), pn, equal_nan=True)
    assert torch.equal(a2.Intensity, a.Intensity)
    h, _ = CalculateHopkinsImage(s, sc, p, num)
    err_h = rel_err(h.Intensity.numpy(), GOLD["hopkins_" + key])
    print(f"golden {key}: abbe {err_a:.2e} hopkins {err_h:.2e}")
    assert err_a < TOL and err_h < TOL
    assert a.Intensity.device.type == "cpu"   # results follow the ScatterField's device


@pytest.mark.gpu
def test_hip_tcc_matches_reference(gpu):
    from opticalimaging_amd import CalculateTCCMatrix
    s, p, num, sc = scalar_case()
    num.TCC_Precision = "fp32"
    T, fg = CalculateTCCMatrix(s, [sc.Period_Y, sc.Period_X], p, 40.0, num)
    assert list(fg) == list(GOLD["tcc_scalar_fg"])
    assert rel_err(T.cpu().numpy(), GOLD["tcc_scalar"]) < TOL
