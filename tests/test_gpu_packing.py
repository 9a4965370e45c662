"""Paired kernels for real masks (engine.real_mask_kernels) on the device:
the same images and mask gradients as the original kernel set, against the
float64 oracle, with half the kernels for in-focus aberration-free optics."""
import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import annular, numerics, projector, rel_err

pytestmark = pytest.mark.gpu


def _masks(n, b, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((b, n, n)) < 0.4).astype(np.float32)


def test_full_rank_set_is_halved_and_images_match(gpu):
    """Kernel set of the package's own decomposition (every eigenpair of the
    TCC) at 256^2: 2K J-real components of rank K, K/2 paired kernels; images
    with and without pairing agree with each other and with the oracle."""
    from opticalimaging_amd import engine
    from opticalimaging_amd.DecomposeTCC_SOCS import socs_kernel_set
    n, period = 256, 1024.0
    src, proj, num = annular(), projector(), numerics(order=10 ** 6)
    ks = socs_kernel_set(src, [period, period], proj, 0.0, num, [n, n])
    pk = ks.for_real_masks()
    print(f"K {ks.K} -> {pk.K} paired kernels, window {ks.wr}x{ks.wc}")
    assert pk.K == (ks.K + 1) // 2 and pk.band == ks.band
    assert (pk.k0r, pk.k0c, pk.wr, pk.wc) == (ks.k0r, ks.k0c, ks.wr, ks.wc)
    masks = torch.as_tensor(_masks(n, 3, seed=4)).cuda()
    a = engine.socs_image(masks, ks).cpu().numpy()
    b = engine.socs_image(masks, ks, pack_real=False).cpu().numpy()
    dense = ks.dense().cpu().numpy().astype(np.complex128)
    for i in range(3):
        f3 = np.zeros((n, n, 3))
        f3[..., 0] = masks[i].cpu().numpy()
        ref = O.socs_image(f3, dense, period, period, proj.NA, src.Wavelength)
        e_pk, e_un, e_ab = rel_err(a[i], ref), rel_err(b[i], ref), rel_err(a[i], b[i])
        print(f"mask {i}: paired {e_pk:.2e}, unpaired {e_un:.2e}, between {e_ab:.2e}")
        assert e_pk < 1e-5 and e_un < 1e-5 and e_ab < 2e-6


def test_truncated_set_and_adjoint(gpu):
    """K = 24 from the dense oracle decomposition at 128^2: fewer kernels, the
    same images and the same mask gradient (oi_socs_adjoint) within 1e-5."""
    from opticalimaging_amd import engine
    n, period = 128, 1000.0
    src, proj, num = annular(), projector(), numerics(order=24)
    tcc, fg = O.tcc_matrix(src, [period, period], proj, 0.0, num)
    kern = O.decompose_tcc_socs(tcc, fg, [n, n], num)
    ks = engine.KernelSet.from_dense(torch.as_tensor(kern.astype(np.complex64)))
    pk = ks.for_real_masks()
    print(f"K {ks.K} -> {pk.K}")
    assert pk.K < ks.K
    rng = np.random.default_rng(11)
    m = torch.as_tensor(rng.random((2, n, n)).astype(np.float32)).cuda()
    g = torch.as_tensor(rng.standard_normal((2, n, n)).astype(np.float32)).cuda()
    a = engine.socs_image(m, ks).cpu().numpy()
    b = engine.socs_image(m, ks, pack_real=False).cpu().numpy()
    for i in range(2):
        f3 = np.zeros((n, n, 3))
        f3[..., 0] = m[i].cpu().numpy()
        ref = O.socs_image(f3, kern, period, period, proj.NA, src.Wavelength)
        assert rel_err(a[i], ref) < 1e-5 and rel_err(b[i], ref) < 1e-5
    ga = engine.socs_adjoint(m, g, ks).cpu().numpy()
    gb = engine.socs_adjoint(m, g, ks, pack_real=False).cpu().numpy()
    err = rel_err(ga, gb)
    print(f"adjoint paired vs unpaired {err:.2e}")
    assert err < 1e-5


def test_even_aberration_keeps_the_set(gpu):
    """An even pupil phase (Zernike term 4) breaks the symmetry: the set is
    returned unchanged and imaging uses it as before."""
    from opticalimaging_amd import engine
    n, period = 128, 1000.0
    z = [0.0] * 37
    z[4] = 0.05
    src, proj, num = annular(), projector(zern=z), numerics(order=12)
    tcc, fg = O.tcc_matrix(src, [period, period], proj, 0.0, num)
    kern = O.decompose_tcc_socs(tcc, fg, [n, n], num)
    ks = engine.KernelSet.from_dense(torch.as_tensor(kern.astype(np.complex64)))
    assert ks.for_real_masks() is ks
    m = torch.as_tensor(_masks(n, 1, seed=2)).cuda()
    img = engine.socs_image(m, ks).cpu().numpy()[0]
    f3 = np.zeros((n, n, 3))
    f3[..., 0] = m[0].cpu().numpy()
    assert rel_err(img, O.socs_image(f3, kern, period, period, proj.NA, src.Wavelength)) < 1e-5
