"""Writes golden vectors from the reference implementation (run in the build
container, where the reference sources are mounted read-only).

Purpose: pin conventions that no closed form fixes -- the row/column <->
Period_Y/Period_X <-> source X/Y pairing on non-square grids and periods, the
vector-mode polarisation path and the PupilImage crop (CalculateAerialImage.py,
CalculateTCCMatrix.py, DecomposeTCC_SOCS.py, CalculateAerialImage_SOCS.py,
Scatter.py).  The reference computes in complex64, so these vectors pin
conventions at ~1e-6, not float64 accuracy.

Only the imaging modules are imported (no RCWA, no plotting).  Output:
tests/golden/reference_small.npz (inputs and outputs, a few hundred KB).
"""
import math
import os
import sys

import numpy as np
import torch

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_small.npz")


def main():
    sys.path.insert(0, REF)
    from Numerics import Numerics
    from Source import Source
    from Scatter import Scatter
    from Projection import Projection
    from CalculateAerialImage import CalculateAbbeImage, CalculateHopkinsImage
    from CalculateTCCMatrix import CalculateTCCMatrix

    torch.manual_seed(0)
    rng = np.random.default_rng(2024)
    ny, nx = 48, 32
    per_x, per_y = 1300.0, 900.0
    mask = (rng.random((ny, nx)) < 0.45).astype(np.float32)
    # the vector case needs a wider grid: its TCC grid at NA 1.2 spans 35 columns
    cfield = (rng.standard_normal((ny, ny, 3)) + 1j * rng.standard_normal((ny, ny, 3))).astype(np.complex64)
    cfield[..., 1:] *= 0.3
    zern = (rng.normal(0, 0.02, 37)).astype(np.float32)
    out = {"mask": mask, "cfield": cfield, "zern": zern,
           "per": np.array([per_x, per_y])}

    def setup(mode, shape="annular", M=100, NA=0.85, n=1.0, foci=(0,)):
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
        p.Aberration_Zernike = torch.tensor(zern)
        p.FocusRange = torch.tensor(list(foci), dtype=torch.float32)
        num = Numerics()
        num.ImageCalculationMode = mode
        num.Hopkins_SettingType = "order"
        num.Hopkins_Order = 12
        return s, p, num

    def scatter(field):
        sc = Scatter()
        sc.Period_X, sc.Period_Y = per_x, per_y
        sc.ScatterField = torch.tensor(field)
        return sc

    # scalar, real mask in channel 0
    f0 = np.zeros((ny, nx, 3), dtype=np.float32)
    f0[..., 0] = mask
    s, p, num = setup("scalar", foci=(0, 40))
    img, pupil = CalculateAbbeImage(s, scatter(f0), p, num)
    out["abbe_scalar"] = img.Intensity.numpy()
    out["pupil_scalar"] = pupil.numpy()
    img, _ = CalculateHopkinsImage(s, scatter(f0), p, num)
    out["hopkins_scalar"] = img.Intensity.numpy()
    tcc, fg = CalculateTCCMatrix(s, [per_y, per_x], p, 40.0, num)
    out["tcc_scalar"] = tcc.numpy()
    out["tcc_scalar_fg"] = np.array(fg)
    # vector, complex 3-channel field, immersion (M = 1 keeps the obliquity finite)
    s, p, num = setup("vector", shape="quasar", M=1, NA=1.2, n=1.44, foci=(-30,))
    img, pupil = CalculateAbbeImage(s, scatter(cfield), p, num)
    out["abbe_vector"] = img.Intensity.numpy()
    out["pupil_vector"] = pupil.numpy()
    img, _ = CalculateHopkinsImage(s, scatter(cfield), p, num)
    out["hopkins_vector"] = img.Intensity.numpy()
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
