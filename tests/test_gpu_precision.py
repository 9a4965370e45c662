"""Precision pattern and decomposition routes on the device.

The TCC contraction runs in 3-term split-bf16 on MFMA with fp32 accumulation
(CalculateTCCMatrix.py:58-76 restated as oi_tcc_build); the eigenproblem that
turns it into SOCS kernels (DecomposeTCC_SOCS.py:7-35) runs in float64.  These
tests pin, against the float64 oracle:
  * the contraction's error at the benchmark grid (M = 101^2 = 10201);
  * images through kernels decomposed from the split-bf16 TCC itself, on the
    dense route (no factor) and on the default projected route;
  * the threshold truncation (Hopkins_Threshold 0.95);
  * that an in-place edit of a TCC or of dense kernels is never hidden by an
    attachment made before the edit, and that a new tensor at a freed
    tensor's address is not mistaken for it;
  * GetTCCMatrix (CalculateTCCMatrix.py:58-76).
"""
import time

import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import annular, numerics, projector, rel_err, scatter

pytestmark = pytest.mark.gpu


def _config1(order=24, precision="split_bf16"):
    from opticalimaging_amd import workloads
    mask, src, proj, num, period = workloads.scalar_abbe_128()
    num.Hopkins_Order = order
    num.TCC_Precision = precision
    field = workloads.field_from_mask(mask)
    return field, src, proj, num, period


def _image(field, kern, src, proj, period):
    from opticalimaging_amd import CalculateAerialImage_SOCS
    sc = scatter(field.cuda(), period, period)
    img, _ = CalculateAerialImage_SOCS(sc, kern, src, proj)
    return img.double().cpu().numpy()


def test_tcc_split_bf16_error_at_m10201(gpu):
    """Split-bf16 / fp32 / fp64 contraction of the 101^2 x 120 benchmark TCC
    against the float64 oracle product n G diag(V) G^H / sum V."""
    from opticalimaging_amd import CalculateTCCMatrix, workloads
    src, proj, num = workloads.socs_1024_optics(order=24, grid_halfwidth=50)
    Gs, V, f, g = O.shifted_pupil_matrix(src, [4096.0, 4096.0], proj, 0.0, num, halfwidth=50)
    ref = float(proj.IndexImage) * sum(O.get_tcc_matrix(V, G) for G in Gs)
    assert ref.shape == (10201, 10201)
    peak = np.max(np.abs(ref))
    for precision, tol in (("split_bf16", 1e-4), ("fp32", 1e-5), ("fp64", 1e-6)):
        num.TCC_Precision = precision
        tcc, fg = CalculateTCCMatrix(src, [4096.0, 4096.0], proj, 0.0, num)
        err = float(np.max(np.abs(tcc.cpu().numpy().astype(np.complex128) - ref)) / peak)
        print(f"TCC M=10201 {precision}: max|T - T64| / max|T64| = {err:.2e}", flush=True)
        assert err < tol
        del tcc


@pytest.mark.parametrize("halfwidth", [4, 12, 20, 31])
def test_tcc_split_bf16_tile_orders(gpu, halfwidth):
    """Grids whose tile counts (M = (2h + 1)^2, ceil(M / 64) = 2, 10, 27, 63
    tiles per side) leave partial supertiles at the edge of the triangle: every
    tile of the split-bf16 contraction is written, the result is Hermitian
    (exactly across tiles, where one tile is written as the other's conjugate
    transpose; to rounding inside the diagonal tiles) and within the 3-term
    bound of the float64 product."""
    from opticalimaging_amd import CalculateTCCMatrix, workloads
    src, proj, num = workloads.socs_1024_optics(order=24, grid_halfwidth=halfwidth)
    Gs, V, f, g = O.shifted_pupil_matrix(src, [4096.0, 4096.0], proj, 0.0, num, halfwidth=halfwidth)
    ref = float(proj.IndexImage) * sum(O.get_tcc_matrix(V, G) for G in Gs)
    num.TCC_Precision = "split_bf16"
    tcc, fg = CalculateTCCMatrix(src, [4096.0, 4096.0], proj, 0.0, num)
    t = tcc.cpu().numpy()
    M = (2 * halfwidth + 1) ** 2
    assert t.shape == (M, M) and np.all(np.isfinite(t))
    peak = np.max(np.abs(t))
    assert np.max(np.abs(t - t.conj().T)) <= 1e-6 * peak
    tb = 64                                   # the contraction's workgroup tile
    off = np.ones((M, M), dtype=bool)
    for i in range(0, M, tb):
        off[i:i + tb, i:i + tb] = False
    assert np.array_equal(t[off], t.conj().T[off])
    err = float(np.max(np.abs(t.astype(np.complex128) - ref)) / np.max(np.abs(ref)))
    print(f"TCC M={M} split-bf16: {err:.2e}")
    assert err < 1e-4


def test_bf16_tcc_through_dense_jacobi(gpu):
    """The split-bf16 TCC itself decomposed (no factor attached: the dense
    float64 Jacobi route at M = 361) gives images within 1e-4 of the float64
    SOCS oracle at the same truncation."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS
    from opticalimaging_amd.CalculateTCCMatrix import factor_of
    field, src, proj, num, period = _config1()
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    plain = tcc.clone()
    assert factor_of(plain) is None and factor_of(tcc) is not None
    kern = DecomposeTCC_SOCS(plain, fg, [128, 128], num)
    img = _image(field, kern, src, proj, period)
    sc = scatter(field.numpy(), period, period)
    ref, _ = O.hopkins_image(src, sc, proj, num)
    err = rel_err(img, ref[0])
    print(f"bf16 TCC -> dense Jacobi -> SOCS K=24 vs float64 SOCS: {err:.2e}", flush=True)
    assert err < 1e-4


@pytest.mark.parametrize("route,tol", [("projected", 1e-4), ("factor", 1e-5)])
def test_factor_routes(gpu, route, tol):
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS
    field, src, proj, num, period = _config1()
    num.TCC_Decomposition = route
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    kern = DecomposeTCC_SOCS(tcc, fg, [128, 128], num)
    img = _image(field, kern, src, proj, period)
    sc = scatter(field.numpy(), period, period)
    ref, _ = O.hopkins_image(src, sc, proj, num, lowrank=True)
    err = rel_err(img, ref[0])
    print(f"route {route}: SOCS K=24 vs float64 SOCS {err:.2e}", flush=True)
    assert err < tol


def test_projected_full_rank_benchmark_optics(gpu):
    """Benchmark optics and grid (M = 10201, every eigenpair kept) on the
    default projected route from the split-bf16 TCC: the 1024^2 image of one
    benchmark mask is within 1e-4 of the float64 Abbe image (full-order SOCS
    equals Abbe), with the pupil-band coarse grid."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS, engine, workloads
    from opticalimaging_amd.DecomposeTCC_SOCS import kernels_of
    masks, period = workloads.socs_1024_masks(batch=1)
    src, proj, num = workloads.socs_1024_optics(order=10 ** 6, grid_halfwidth=50)
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    ks = kernels_of(DecomposeTCC_SOCS(tcc, fg, [1024, 1024], num))
    assert ks.band is not None and ks.K == 120
    img = engine.socs_image(torch.as_tensor(masks).cuda(), ks)[0].double().cpu().numpy()
    sc = scatter(workloads.field_from_mask(masks[0]).numpy(), period, period)
    num.ImageCalculationMethod = "abbe"
    ref, *_ = O.abbe_image(src, sc, proj, num)
    err = rel_err(img, ref[0])
    print(f"bf16 TCC -> projected full rank (K={ks.K}) 1024^2 vs float64 Abbe: {err:.2e}", flush=True)
    assert err < 1e-4


def test_dense_tcc_subspace_route_m10201(gpu):
    """A benchmark-size TCC without its factor (tcc * 1) goes through the
    float64 subspace iteration; its K = 24 image agrees with the factor
    route's and with the float64 SOCS oracle."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS, engine, workloads
    from opticalimaging_amd.DecomposeTCC_SOCS import kernels_of
    masks, period = workloads.socs_1024_masks(batch=1)
    src, proj, num = workloads.socs_1024_optics(order=24, grid_halfwidth=50)
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    num.TCC_Decomposition = "factor"
    ks_f = kernels_of(DecomposeTCC_SOCS(tcc, fg, [1024, 1024], num))
    num.TCC_Decomposition = "projected"
    plain = tcc * 1
    torch.cuda.synchronize()
    t = time.perf_counter()
    ks_d = kernels_of(DecomposeTCC_SOCS(plain, fg, [1024, 1024], num))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    m = torch.as_tensor(masks).cuda()
    a = engine.socs_image(m, ks_f)[0].double().cpu().numpy()
    b = engine.socs_image(m, ks_d)[0].double().cpu().numpy()
    sc = scatter(workloads.field_from_mask(masks[0]).numpy(), period, period)
    ref, _ = O.hopkins_image(src, sc, proj, num, halfwidth=50, lowrank=True)
    e_fd = rel_err(b, a)
    e_ref = rel_err(b, ref[0])
    print(f"dense M=10201 subspace route: {dt:.2f} s, K={ks_d.K}, window {ks_d.wr}x{ks_d.wc}; "
          f"vs factor route {e_fd:.2e}, vs float64 SOCS {e_ref:.2e}", flush=True)
    assert ks_d.K == 24 and e_fd < 1e-4 and e_ref < 1e-4


def test_threshold_095(gpu):
    """Hopkins_SettingType 'threshold' at 0.95 (DecomposeTCC_SOCS.py:10-20,
    the reference's evident intent): the kernel count equals the float64
    oracle's and the image matches the float64 SOCS at that count."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS
    from opticalimaging_amd.DecomposeTCC_SOCS import kernels_of
    for route in ("factor", "projected"):
        field, src, proj, num, period = _config1(precision="fp64")
        num.Hopkins_SettingType = "threshold"
        num.Hopkins_Threshold = 0.95
        num.TCC_Decomposition = route
        tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
        kern = DecomposeTCC_SOCS(tcc, fg, [128, 128], num)
        t64, fg64 = O.tcc_matrix(src, [period, period], proj, 0.0, num)
        s, _ = O.eig_desc(t64)
        k_ref = O.socs_count(s, t64.shape[0], num)
        img = _image(field, kern, src, proj, period)
        sc = scatter(field.numpy(), period, period)
        ref, _ = O.hopkins_image(src, sc, proj, num)
        err = rel_err(img, ref[0])
        print(f"threshold 0.95 ({route}): K = {kernels_of(kern).K} (oracle {k_ref}), "
              f"vs float64 SOCS {err:.2e}", flush=True)
        assert kernels_of(kern).K == k_ref == kern.shape[2]
        assert err < 1e-5


def test_in_place_tcc_edit_is_not_hidden(gpu):
    """tcc.mul_(2) after CalculateTCCMatrix: the attached factor no longer
    describes the tensor, so DecomposeTCC_SOCS decomposes what the tensor
    holds, and the image doubles (the image is linear in the TCC)."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS
    from opticalimaging_amd.CalculateTCCMatrix import factor_of
    field, src, proj, num, period = _config1(precision="fp32")
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    base = _image(field, DecomposeTCC_SOCS(tcc, fg, [128, 128], num), src, proj, period)
    tcc.mul_(2)
    assert factor_of(tcc) is None
    twice = _image(field, DecomposeTCC_SOCS(tcc, fg, [128, 128], num), src, proj, period)
    err = rel_err(twice, 2 * base)
    print(f"after tcc.mul_(2): image / (2 x before) error {err:.2e}", flush=True)
    assert err < 1e-5


def test_in_place_kernel_edit_is_not_hidden(gpu):
    """Zeroing the tail of DecomposeTCC_SOCS's dense kernels in place: the
    image is that of the kept kernels, not of the attached full set."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS
    field, src, proj, num, period = _config1(precision="fp32")
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    kern = DecomposeTCC_SOCS(tcc, fg, [128, 128], num)
    full = _image(field, kern, src, proj, period)
    kern[..., 12:] = 0
    cut = _image(field, kern, src, proj, period)
    num.Hopkins_Order = 12
    ref12 = _image(field, DecomposeTCC_SOCS(tcc, fg, [128, 128], num), src, proj, period)
    print(f"kernels 12.. zeroed in place: vs 12-kernel image {rel_err(cut, ref12):.2e}, "
          f"vs 24-kernel image {rel_err(cut, full):.2e}", flush=True)
    assert rel_err(cut, ref12) < 1e-5 and rel_err(cut, full) > 1e-3


def test_dense_kernel_cache_is_not_address_keyed(gpu):
    """Reference-format kernels (no attachment): a second tensor of the same
    shape, allocated where a freed one was, and a numpy array edited in place
    are each imaged with their own values."""
    field, src, proj, num, period = _config1()
    t64, fg = O.tcc_matrix(src, [period, period], proj, 0.0, num)
    k64 = O.decompose_tcc_socs(t64, fg, [128, 128], num).astype(np.complex64)
    a = torch.as_tensor(k64).cuda()
    img_a = _image(field, a, src, proj, period)
    del a
    b = torch.as_tensor(0.5 * k64).cuda()          # same shape, likely the freed block
    img_b = _image(field, b, src, proj, period)
    assert rel_err(img_b, 0.25 * img_a) < 1e-6
    kn = k64.copy()
    img_n = _image(field, kn, src, proj, period)
    kn *= 0.5
    img_n2 = _image(field, kn, src, proj, period)
    assert rel_err(img_n2, 0.25 * img_n) < 1e-6
    b.mul_(2)                                        # in-place edit of a cached tensor
    assert rel_err(_image(field, b, src, proj, period), img_a) < 1e-6


def test_get_tcc_matrix(gpu):
    """GetTCCMatrix(sourceData, shiftedPupil) (CalculateTCCMatrix.py:58-76)
    against the float64 oracle, on the config-1 shifted-pupil matrix."""
    from opticalimaging_amd import GetTCCMatrix
    field, src, proj, num, period = _config1()
    Gs, V, f, g = O.shifted_pupil_matrix(src, [period, period], proj, 0.0, num)
    ref = O.get_tcc_matrix(V, Gs[0])
    sd = src.Calc_SourceSimple()
    G = torch.as_tensor(Gs[0].astype(np.complex64))
    for precision, tol in (("fp64", 2e-7), ("fp32", 1e-5), ("split_bf16", 1e-4)):
        num.TCC_Precision = precision
        got = GetTCCMatrix(sd, G, num)
        err = rel_err(got.cpu().numpy(), ref)
        print(f"GetTCCMatrix {precision}: {err:.2e}", flush=True)
        assert got.shape == ref.shape and err < tol
    with pytest.raises(ValueError):
        GetTCCMatrix(sd, G[:, :-1])


def test_eigh_reports_failures(gpu):
    """oi_eigh_hermitian returns an error (RuntimeError) instead of
    unconverged or non-finite eigenpairs."""
    from opticalimaging_amd.DecomposeTCC_SOCS import eigh
    rng = np.random.default_rng(0)
    a = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    h = torch.as_tensor(a + a.conj().T).cuda()
    w, v = eigh(h)
    assert np.allclose(np.sort(w.cpu().numpy()), np.linalg.eigvalsh(a + a.conj().T), atol=1e-10)
    with pytest.raises(RuntimeError, match="no convergence"):
        eigh(h, max_sweeps=0)
    bad = h.clone()
    bad[3, 5] = float("nan")
    with pytest.raises(RuntimeError, match="not finite"):
        eigh(bad)
