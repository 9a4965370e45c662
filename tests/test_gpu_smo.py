"""Source-weight gradient of the Abbe image (opticalimaging_amd.SMO) against
float64 autograd of the oracle's I(V) = (n / sum V) sum_s V_s J_s, with J_s
the oracle's per-point intensities (CalculateAerialImage.py:46-92 restated
per source point), and against central differences of the device image.
Also: device-side source generation (Source.Calc_SourceAll on the HIP
device) equals the host's float32 selection bit for bit."""
import math

import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import annular, numerics, projector, quasar, rel_err, scatter

pytestmark = pytest.mark.gpu


def _field(n, seed):
    from opticalimaging_amd import workloads
    rng = np.random.default_rng(seed)
    m = np.where(rng.random((n, n)) < 0.4, -np.sqrt(0.06), 1.0).astype(np.float32)
    return workloads.field_from_mask(m)


CASES = {
    "scalar annular, 2 foci": (lambda: annular(P=11), lambda: projector(M=1, foci=(-40.0, 40.0)),
                               lambda: numerics(mode="scalar"), 64, 1000.0),
    "vector quasar t_pol, aberrated": (
        lambda: quasar(P=11), lambda: projector(NA=1.2, n=1.44, M=1, foci=(0.0, 50.0),
                                                zern=np.random.default_rng(3).normal(0, 0.02, 37)),
        lambda: numerics(mode="vector"), 64, 512.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_source_gradient_matches_float64_autograd(gpu, name):
    from opticalimaging_amd import SMO
    mk_src, mk_proj, mk_num, n, period = CASES[name]
    src, proj, num = mk_src(), mk_proj(), mk_num()
    field = _field(n, 7)
    J, V, n_index = O.abbe_source_terms(src, scatter(field.numpy(), period, period), proj, num)
    nf = J.shape[0]
    rng = np.random.default_rng(11)
    g = rng.standard_normal((nf, n, n))
    Vt = torch.tensor(V, dtype=torch.float64, requires_grad=True)
    Jt = torch.tensor(J)
    I_ref = n_index / Vt.sum() * torch.einsum("s,fsij->fij", Vt, Jt)
    (I_ref * torch.tensor(g)).sum().backward()
    grad_ref = Vt.grad.numpy()

    prob = SMO.abbe_source_problem(field.cuda(), src, scatter(field.cuda(), period, period), proj, num)
    assert prob.nsrc == len(V)
    Vd = prob.V0.clone().requires_grad_(True)
    I = SMO.abbe_image_weighted(Vd, prob)
    gd = torch.tensor(g, dtype=torch.float32, device="cuda")
    (I * gd).sum().backward()
    e_img = rel_err(I.detach().double().cpu().numpy(), I_ref.detach().numpy())
    e_grad = rel_err(Vd.grad.cpu().numpy(), grad_ref)
    print(f"{name}: {len(V)} points, coarse {prob.Lr}x{prob.Lc}; image vs float64 {e_img:.2e}; "
          f"dL/dV vs float64 autograd {e_grad:.2e}", flush=True)
    assert e_img < 1e-5 and e_grad < 1e-5
    # central differences of the device image along three source weights
    gsum = lambda Vx: float((prob.image(Vx).double() * gd.double()).sum())  # noqa: E731
    for s in rng.choice(len(V), 3, replace=False):
        d = torch.zeros_like(prob.V0)
        d[s] = 0.05 * float(prob.V0[s])
        fd = (gsum(prob.V0 + d) - gsum(prob.V0 - d)) / (2 * float(d[s]))
        assert abs(fd - float(Vd.grad[s])) < 2e-3 * float(Vd.grad.abs().max()), (s, fd, float(Vd.grad[s]))


def test_abbe_weighted_image_equals_drop_in(gpu):
    """At the source's own weights the weighted image is CalculateAbbeImage's."""
    from opticalimaging_amd import SMO, CalculateAbbeImage
    src, proj, num = quasar(P=11), projector(NA=1.2, n=1.44, M=1, foci=(0.0,)), numerics(mode="vector")
    field = _field(64, 5)
    sc = scatter(field.cuda(), 512.0, 512.0)
    ref, _ = CalculateAbbeImage(src, sc, proj, num)
    prob = SMO.abbe_source_problem(field.cuda(), src, sc, proj, num)
    img = prob.image(prob.V0)
    assert rel_err(img.cpu().numpy(), ref.Intensity.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("shape", ["annular", "quasar", "dipolecirc", "multipole", "pixel", "masked"])
def test_device_source_generation_is_bit_exact(gpu, shape):
    from opticalimaging_amd import Source
    s = Source()
    s.PntNum = 21 if shape != "quasar" else 41
    s.SigmaIn, s.SigmaOut = 0.5, 0.8
    if shape == "quasar":
        s.Shape, s.SigmaIn, s.SigmaOut, s.OpenAngle = "quasar", 0.6, 0.9, math.pi / 6
    elif shape == "dipolecirc":
        s.Shape, s.OpenAngle, s.RotateAngle = "dipolecirc", math.pi / 5, 0.3
    elif shape == "multipole":
        s.Shape, s.PoleNumber, s.SigmaCenter, s.SigmaRadius, s.RotateAngle = "multipole", 4, 0.6, 0.2, 0.1
    elif shape == "pixel":
        s.Shape = "pixel"
        c = torch.linspace(-1, 1, s.PntNum)
        X, Y = torch.meshgrid(c, c, indexing="ij")
        s.SPointX, s.SPointY = X.clone(), Y.clone()
        s.SPointValue = torch.rand(s.PntNum, s.PntNum, generator=torch.Generator().manual_seed(2))
    elif shape == "masked":
        s.Source_Mask = (torch.rand(s.PntNum, s.PntNum, generator=torch.Generator().manual_seed(4)) > 0.3
                         ).float()
    host = s.Calc_SourceSimple()
    hx, hy, hv = host.X.clone(), host.Y.clone(), host.Value.clone()
    dev = s.Calc_SourceSimple(device="cuda")
    assert dev.Value.is_cuda
    print(f"{shape}: {hv.numel()} lit points on the host, {dev.Value.numel()} on the device")
    assert torch.equal(dev.X.cpu(), hx) and torch.equal(dev.Y.cpu(), hy)
    assert torch.equal(dev.Value.cpu(), hv)


def test_device_source_generation_with_psf(gpu):
    """With the PSF blur the device convolution sums in another order: the
    weights agree to float32 rounding."""
    from opticalimaging_amd import Source
    s = Source()
    s.PntNum, s.SigmaIn, s.SigmaOut, s.PSFEnable, s.PSFSigma = 31, 0.4, 0.8, True, 0.05
    host = s.Calc_SourceAll()
    hv = host.Value.clone()
    dv = s.Calc_SourceAll(device="cuda").Value.cpu()
    assert float((dv - hv).abs().max()) <= 1e-6 * float(hv.abs().max())
