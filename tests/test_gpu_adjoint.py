"""Hand-written SOCS adjoint (oi_socs_adjoint) and the ILT autograd path.

Checks, for real masks:
  * the HIP mask gradient of sum(g * I) against float64 torch autograd through
    the SOCS restatement (oracle/socs_torch.py), itself pinned against central
    differences of the numpy oracle; generic, direct, compile-time-size and
    full-rank band plans; batches of two;
  * at 1024^2 (full size), the exact dot-product identity of a quadratic map:
    sum g (I(h + d) - I(h - d)) / 2 = sum grad . d;
  * the ILT objective ||I(sigmoid(theta)) - T||^2 at the mask-optimisation
    workload (16 x 1024^2, 24 kernels from the vector quasar optics):
    gradient against float64 autograd, against central differences
    (eps = 1e-3) at 64 random pixels, and 50 descent steps that lower the loss.
Tolerance: max |grad - ref| / max |ref| <= 1e-5 (float32 path vs float64).
"""
import numpy as np
import pytest
import torch

import oracle as O
from oracle.socs_torch import socs_mask_gradient
from gpu_cases import annular, numerics, projector, rel_err
from opticalimaging_amd.DecomposeTCC_SOCS import kernels_of

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _masks(ny, nx, b, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((b, ny, nx)) < 0.4).astype(np.float32)


def _oracle_grads(masks, gI, kern, pxx, pyy, NA, lam):
    return np.stack([socs_mask_gradient(masks[b].astype(np.float64), gI[b].astype(np.float64), kern,
                                        pxx, pyy, NA, lam) for b in range(masks.shape[0])])


CASES = [
    (128, 128, 1000.0, 1000.0, 24),   # compile-time-size path
    (128, 256, 2000.0, 1000.0, 20),   # compile-time-size path, non-square
    (96, 160, 800.0, 1200.0, 16),     # generic kernels, non-square
    (32, 32, 1000.0, 1000.0, 30),     # coarse grid = image grid (direct)
    (60, 90, 900.0, 1300.0, 10),      # generic, 2-3-5 sizes
]


@pytest.mark.parametrize("ny,nx,pxx,pyy,order", CASES)
def test_adjoint_matches_float64_autograd(gpu, ny, nx, pxx, pyy, order):
    from opticalimaging_amd import engine
    src, proj, num = annular(), projector(), numerics(order=order)
    tcc, fg = O.tcc_matrix(src, [pyy, pxx], proj, 0.0, num)
    kern = O.decompose_tcc_socs(tcc, fg, [nx, ny], num)
    ks = engine.KernelSet.from_dense(torch.as_tensor(kern.astype(np.complex64)))
    masks = _masks(ny, nx, 2, seed=ny + nx)
    gI = np.random.default_rng(3).standard_normal((2, ny, nx)).astype(np.float32)
    plan = engine.plan_for(ks, 1, 2)
    grad = engine.socs_adjoint(torch.as_tensor(masks).cuda(), torch.as_tensor(gI).cuda(), ks,
                               plan=plan).cpu().numpy()
    ref = _oracle_grads(masks, gI, kern, pxx, pyy, proj.NA, src.Wavelength)
    for b in range(2):
        err = rel_err(grad[b], ref[b])
        print(f"adjoint {ny}x{nx} K={ks.K} plan {plan.info()['lr']}x{plan.info()['lc']} "
              f"path {plan.info()['path']} direct {plan.info()['direct']} img {b}: {err:.2e}")
        assert err < TOL


def test_adjoint_full_rank_band_plan(gpu):
    """Every eigenpair kept: the adjoint runs on the narrower TCC-band coarse grid."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS, engine
    n, period = 256, 1024.0
    src, proj = annular(), projector()
    num = numerics(order=10 ** 6, precision="fp32")
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    kern = DecomposeTCC_SOCS(tcc, fg, [n, n], num)
    ks = kernels_of(kern)
    plan = engine.plan_for(ks, 1, 1)
    assert plan.info()["lr"] < 2 * ks.wr - 1, plan.info()
    masks = _masks(n, n, 1, seed=11)
    gI = np.random.default_rng(12).standard_normal((1, n, n)).astype(np.float32)
    grad = engine.socs_adjoint(torch.as_tensor(masks).cuda(), torch.as_tensor(gI).cuda(), ks,
                               plan=plan).cpu().numpy()
    dense = ks.dense().cpu().numpy().astype(np.complex128)
    ref = _oracle_grads(masks, gI, dense, period, period, proj.NA, src.Wavelength)
    err = rel_err(grad[0], ref[0])
    print(f"adjoint full-rank band coarse {plan.info()['lr']} (window {ks.wr}): {err:.2e}")
    assert err < TOL


def test_adjoint_dot_product_1024(gpu):
    """Full size, fast path: I is quadratic in the mask, so the central
    difference of I along d equals the derivative exactly for any step.  The
    step is large (0.5 per pixel, white) because the left side is a sum over
    10^6 pixels of g * (I(h+d) - I(h-d)): its float32 rounding relative to the
    result scales as eps * I / (I(h+d) - I(h-d)), about 5e-5 at a 0.05 step."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS, engine, workloads
    masks_np, period = workloads.socs_1024_masks(batch=2)
    src, proj, num = workloads.socs_1024_optics(order=24, grid_halfwidth=50)
    tcc, fg = CalculateTCCMatrix(src, [period, period], proj, 0.0, num)
    ks = kernels_of(DecomposeTCC_SOCS(tcc, fg, [1024, 1024], num))
    h = torch.as_tensor(masks_np).cuda()
    gen = torch.Generator(device="cuda").manual_seed(5)
    d = 0.5 * torch.randn(h.shape, device="cuda", generator=gen)
    g = torch.randn(h.shape, device="cuda", generator=gen)
    ks.for_real_masks()                 # the paired set, used by the pack_real=True calls
    for pack in (True, False):
        grad = engine.socs_adjoint(h, g, ks, pack_real=pack)
        Ip = engine.socs_image(h + d, ks, pack_real=pack)
        Im = engine.socs_image(h - d, ks, pack_real=pack)
        for b in range(2):
            lhs = float((g[b].double() * (Ip[b].double() - Im[b].double())).sum() / 2)
            rhs = float((grad[b].double() * d[b].double()).sum())
            print(f"dot test 1024 paired={pack} img {b}: {lhs:.8e} vs {rhs:.8e}")
            assert abs(lhs - rhs) <= 1e-4 * abs(rhs)


# ----------------------------------------------------------------- ILT (C5)
def _c2_kernels_1024():
    """SOCS kernels (K = 24) of the vector quasar optics at 1024^2 / 4096 nm."""
    from opticalimaging_amd import CalculateTCCMatrix, DecomposeTCC_SOCS, workloads
    src, proj, num = workloads.ilt_optics()
    tcc, fg = CalculateTCCMatrix(src, [4096.0, 4096.0], proj, 0.0, num)
    kern = DecomposeTCC_SOCS(tcc, fg, [1024, 1024], num)
    del tcc
    return kernels_of(kern), src, proj


def _fields_raw(h, dense, period, NA, lam):
    """Coherent fields before the reference's final fftshift
    (CalculateAerialImage_SOCS.py:7), float64, as a complex128 torch tensor
    [K, Ny, Nx]: E_k = df dg fft2(phi_k * ifftshift(Hk))."""
    f3 = np.zeros(h.shape + (3,))
    f3[..., 0] = h
    hk, f, g, _ = O.spectrum(f3, period, period, NA, lam)
    sp = torch.as_tensor(np.fft.ifftshift(hk[:, :, 0]))
    scale = (f[1] - f[0]) * (g[1] - g[0])
    k = torch.as_tensor(dense).permute(2, 0, 1)
    return scale * torch.fft.fft2(k * sp[None])


def _fd_central(theta, IT, E, psi, b, m, eps):
    """(L(theta_m + eps) - L(theta_m - eps)) / (2 eps) for L = sum (I - T)^2,
    I = sum_k |E_k|^2 (raw layout; IT = I - T), evaluated exactly: a one-pixel mask change
    delta adds delta psi_k(j - m') to E_k (E is linear in the mask; m' is the
    pixel's index after ifftshift), and the loss difference is formed as
    sum (I+ - I-)(I+ + I- - 2 T) so that nothing cancels."""
    y, x = m
    n0, n1 = theta.shape
    sh = ((y - n0 // 2) % n0, (x - n1 // 2) % n1)
    s0 = 1 / (1 + np.exp(-theta[y, x]))
    dp = 1 / (1 + np.exp(-(theta[y, x] + eps))) - s0
    dm = 1 / (1 + np.exp(-(theta[y, x] - eps))) - s0
    a = (torch.conj(E) * torch.roll(psi, sh, dims=(1, 2))).sum(0).real
    bb = torch.roll(b, sh, dims=(0, 1))
    D = 2 * (dp - dm) * a + (dp * dp - dm * dm) * bb
    S = 2 * IT + 2 * (dp + dm) * a + (dp * dp + dm * dm) * bb
    return float((D * S).sum()) / (2 * eps)


def test_ilt_gradient_fd_autograd_and_descent(gpu):
    from opticalimaging_amd import ILT, workloads
    ks, src, proj = _c2_kernels_1024()
    theta_np, T_np = workloads.ilt_problem(batch=16, n=1024)
    theta = torch.as_tensor(theta_np).cuda()
    T = torch.as_tensor(T_np).cuda()
    t = theta.clone().requires_grad_(True)
    with torch.enable_grad():
        loss = ILT.ilt_loss(t, T, ks)
        (g,) = torch.autograd.grad(loss, t)
    g0 = g[0].double().cpu().numpy()

    # float64 autograd through the torch restatement (image 0)
    dense = ks.dense().cpu().numpy().astype(np.complex128)
    from oracle.socs_torch import socs_image_t
    th = torch.as_tensor(theta_np[0].astype(np.float64), dtype=torch.float64).requires_grad_(True)
    kt = torch.as_tensor(dense)
    with torch.enable_grad():
        I = socs_image_t(torch.sigmoid(th), kt, 4096.0, 4096.0, proj.NA, src.Wavelength)
        L = ((I - torch.as_tensor(T_np[0], dtype=torch.float64)) ** 2).sum()
        (gref,) = torch.autograd.grad(L, th)
    gref = gref.numpy()
    err = rel_err(g0, gref)
    print(f"ILT gradient vs float64 autograd: {err:.2e} (K={ks.K}, window {ks.wr})")
    assert err < TOL

    # central differences at 64 random pixels (float64 fields of image 0)
    period, NA, lam = 4096.0, proj.NA, src.Wavelength
    h = 1 / (1 + np.exp(-theta_np[0].astype(np.float64)))
    E = _fields_raw(h, dense, period, NA, lam)
    imp = np.zeros_like(h)
    imp[512, 512] = 1.0                      # index 0 after ifftshift
    psi = _fields_raw(imp, dense, period, NA, lam)
    b = (psi.real ** 2 + psi.imag ** 2).sum(0)
    Iref = torch.fft.fftshift((E.real ** 2 + E.imag ** 2).sum(0)).numpy()
    assert rel_err(Iref, I.detach().numpy()) < 1e-12
    T_raw = torch.as_tensor(np.fft.ifftshift(T_np[0].astype(np.float64)))
    IT = (E.real ** 2 + E.imag ** 2).sum(0) - T_raw
    rng = np.random.default_rng(64)
    pix = [(int(a), int(c)) for a, c in rng.integers(0, 1024, size=(64, 2))]
    eps = 1e-3
    worst = 0.0
    for m in pix:
        fd = _fd_central(theta_np[0].astype(np.float64), IT, E, psi, b, m, eps)
        worst = max(worst, abs(fd - g0[m]) / np.max(np.abs(gref)))
    print(f"ILT gradient vs central differences (64 pixels, eps 1e-3): {worst:.2e}")
    assert worst < 1e-4

    # 50 descent steps lower the loss
    theta2, losses = ILT.ilt(theta, T, ks, steps=50, lr=workloads.ILT_STEP)
    lo = losses.cpu().numpy()
    print(f"ILT 50 steps: loss {lo[0]:.4e} -> {lo[-1]:.4e}")
    assert np.all(np.isfinite(lo)) and lo[-1] < 0.5 * lo[0]


def test_adjoint_rejects_complex_field_plans(gpu):
    """The adjoint is defined for real masks: a plan carrying complex channels
    is refused with ValueError (status OI_ERR_INVALID), nothing is launched."""
    from opticalimaging_amd import engine
    import opticalimaging_amd._native as nat
    src, proj, num = annular(), projector(), numerics(order=8)
    tcc, fg = O.tcc_matrix(src, [1000.0, 1000.0], proj, 0.0, num)
    kern = O.decompose_tcc_socs(tcc, fg, [64, 64], num)
    ks = engine.KernelSet.from_dense(torch.as_tensor(kern.astype(np.complex64)))
    plan = engine.Plan(ks.ny, ks.nx, ks.k0r, ks.wr, ks.k0c, ks.wc, nchan=2, max_batch=1)
    m = torch.zeros((1, 64, 64), device="cuda")
    with pytest.raises(ValueError):
        engine.socs_adjoint(m, m.clone(), ks, plan=plan)
