"""Floquet-mode SOCS (CalculateSOCS, CalculateOptimized) against the float64
direct-sum oracle: max |I - I_ref| / max |I_ref| <= 1e-5."""
import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import annular, numerics, projector, rel_err

pytestmark = pytest.mark.gpu


def _numerics(gx, gy, ox, oy, order=12):
    n = numerics(order=order)
    n.ScatterGrid_X, n.ScatterGrid_Y = gx, gy
    n.ScatterOrder_X, n.ScatterOrder_Y = ox, oy
    return n


@pytest.mark.parametrize("gx,gy,ox,oy,K", [(256, 128, 5, 5, 7), (100, 100, 15, 15, 24), (60, 45, 7, 4, 3)])
def test_calculate_socs_random(gpu, gx, gy, ox, oy, K):
    from opticalimaging_amd import CalculateSOCS
    rng = np.random.default_rng(gx + gy)
    n0, n1 = 2 * ox + 1, 2 * oy + 1
    fm = rng.standard_normal((n0, n1, 3)) + 1j * rng.standard_normal((n0, n1, 3))
    fm[..., 1] = 0                                  # a dead channel is skipped
    kern = rng.standard_normal((n1, n0, K)) + 1j * rng.standard_normal((n1, n0, K))
    num = _numerics(gx, gy, ox, oy)
    img = CalculateSOCS(torch.as_tensor(fm.astype(np.complex64)),
                        torch.as_tensor(kern.astype(np.complex64)), num)
    ref = O.calculate_socs(fm, kern, gx, gy, ox, oy)
    err = rel_err(img.numpy(), ref)
    print(f"CalculateSOCS {gx}x{gy} orders {ox},{oy} K={K}: {err:.2e}")
    assert img.shape == (gx, gy) and err < 1e-5


def test_calculate_optimized_two_foci(gpu):
    from opticalimaging_amd import CalculateOptimized, Scatter
    src, proj = annular(), projector(foci=(-40.0, 40.0), M=1)
    num = _numerics(128, 96, 9, 9, order=16)
    rng = np.random.default_rng(4)
    fm = (rng.standard_normal((19, 19, 3)) + 1j * rng.standard_normal((19, 19, 3))) * 0.1
    sc = Scatter()
    sc.Period_X, sc.Period_Y = 1000.0, 900.0
    sc.FloquetMode = torch.as_tensor(fm.astype(np.complex64))
    d, empty = CalculateOptimized(src, sc, proj, num)
    ref = O.optimized_image(src, sc, proj, num)
    assert tuple(d.Intensity.shape) == (2, 128, 96) and empty.numel() == 0
    for i in range(2):
        err = rel_err(d.Intensity[i].numpy(), ref[i])
        print(f"CalculateOptimized focus {i}: {err:.2e}")
        assert err < 1e-5
