"""Source-weight gradients of the Abbe image, for source-mask optimisation.

The Abbe image (CalculateAerialImage.py:46-92) is an explicit function of
the lit source points' weights V:

    I(x) = (n / W) sum_s V_s J_s(x),   W = sum_s V_s,
    J_s  = sum_c sum_e |E_sce|^2  (unit-weight coherent field of point s),

so for a loss with image cotangent g,

    dL/dV_s = (n / W) <g, J_s> - <g, I> / W.

<g, J_s> is taken on a coarse grid: J_s is a trigonometric polynomial whose
band |q| <= B is one shifted pupil wide, so with gamma = the band of g
resampled to an L x L grid (L > 2B, L holding the spectrum window),

    <g, J_s>_(N grid) = (N^2 / L^2) sum_t gamma(t) J_s^L(t)

exactly.  All J_s^L come from one oi_coherent_image call on a direct L x L
plan with the roles of spectrum and kernels exchanged: the unit-weight
shifted pupils of the points are the "spectra" of a batch of Nsrc images with
one channel per polarisation component, and the field's windowed spectrum
(one kernel per field channel) is the "kernel set".  gamma comes from a
float64 FFT of g on the device.

``AbbeSourceImage`` wraps this as a torch.autograd.Function of V;
``abbe_source_problem`` sets up a fixed field and optics.  The reference has
no source gradient (it would differentiate its complex64 loop with torch
autograd); the float64 gradient of the oracle's I(V) pins this one.
"""
from __future__ import annotations

import math

import torch

from . import _native as nat
from . import engine, optics
from .Scatter import _pitches

__all__ = ["AbbeSourceProblem", "AbbeSourceImage", "abbe_source_problem", "abbe_image_weighted"]


def _smooth_at_least(n: int) -> int:
    """Smallest 2,3,5-smooth even integer >= n."""
    m = max(2, n)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1 and m % 2 == 0:
            return m
        m += 1


class AbbeSourceProblem:
    """A field and optics held fixed, the source weights V free.

    Positions, polarisation and the lit set come from ``source`` (the
    reference's Calc_SourceSimple selection); ``V0`` holds its weights."""

    def __init__(self, field, source, scatter, projector, numerics):
        self.vector = optics.imaging_mode(numerics)
        self.dev = nat.device()
        h = torch.as_tensor(field)
        self.ny, self.nx = int(h.shape[0]), int(h.shape[1])
        px, py = _pitches(scatter, projector, source)
        self.src = optics.source_set(source)
        self.V0 = self.src.V.clone()
        self.n_index = float(projector.IndexImage)
        self.ncomp = 3 if self.vector else 1
        k0r, wr, k0c, wc, br, bc = optics.abbe_window(self.src, py, px, self.ny, self.nx)
        self.win = (k0r, wr, k0c, wc)
        self.band = (br, bc)
        tens, kind, cin, live = engine.field_encoding(h)
        self.C = len(live)
        nsrc = self.src.n
        self.plan = engine.Plan(self.ny, self.nx, k0r, wr, k0c, wc, nchan=self.C, max_batch=1,
                                band_r=br, band_c=bc, max_kernels=self.ncomp * nsrc)
        lib = nat.load()
        self.spec = torch.empty((1, self.C, wr, wc), dtype=torch.complex64, device=self.dev)
        cptr, keep = nat.int_array(live)
        nat.check(lib.oi_spectrum_window(self.plan.handle, nat.ptr(tens), kind, cin, cptr, 1,
                                         nat.ptr(self.spec), nat.stream()))
        del keep
        # unit-weight shifted pupils per focus, c64 [ncomp][nsrc][wr][wc]
        ones = torch.ones(nsrc, dtype=torch.float64, device=self.dev)
        self.foci = optics.focus_values(projector)
        self.G = []
        for z in self.foci:
            opt = optics.optics_record(source, projector, z, py, px, self.vector)
            self.G.append(optics.pupil_kernels(self.src, opt, ones, k0r, wr, k0c, wc))
        # coarse grid of the per-point intensities: holds the window and
        # twice a pupil's band
        need = max(2 * max(br, bc) - 1, 2 * max(-k0r, k0r + wr - 1, -k0c, k0c + wc - 1) + 2)
        L = _smooth_at_least(need)
        self.Lr = self.ny if L >= self.ny else L
        self.Lc = self.nx if L >= self.nx else L
        self.cplan = engine.Plan(self.Lr, self.Lc, k0r, wr, k0c, wc, nchan=self.ncomp, max_batch=nsrc,
                                 band_r=wr, band_c=wc, lr=self.Lr, lc=self.Lc, max_kernels=self.C)
        if not self.cplan.geo.direct:
            raise RuntimeError("per-point intensity plan is not direct")

    @property
    def nsrc(self) -> int:
        return self.src.n

    def image(self, V: torch.Tensor) -> torch.Tensor:
        """Abbe intensities (nFocus, ny, nx) float32 for weights V (nsrc,)."""
        V = V.detach().to(self.dev, torch.float64)
        scale = torch.sqrt(self.n_index * V / V.sum()).to(torch.float32)
        lib = nat.load()
        out = torch.empty((len(self.foci), self.ny, self.nx), dtype=torch.float32, device=self.dev)
        for i, G in enumerate(self.G):
            K = (G * scale[None, :, None, None]).contiguous()
            nat.check(lib.oi_coherent_image(self.plan.handle, nat.ptr(self.spec), 1, nat.ptr(K),
                                            K.shape[0] * K.shape[1], nat.ptr(out[i]), nat.stream()))
        return out

    def point_intensities(self, focus_index: int) -> torch.Tensor:
        """J_s on the coarse Lr x Lc grid, float32 (nsrc, Lr, Lc)."""
        G = self.G[focus_index]
        pupils = G.permute(1, 0, 2, 3).contiguous()          # [nsrc][ncomp][wr][wc]
        kern = self.spec[0].contiguous()                      # [C][wr][wc]
        out = torch.empty((self.nsrc, self.Lr, self.Lc), dtype=torch.float32, device=self.dev)
        nat.check(nat.load().oi_coherent_image(self.cplan.handle, nat.ptr(pupils), self.nsrc,
                                               nat.ptr(kern), self.C, nat.ptr(out), nat.stream()))
        return out

    def _gamma(self, g: torch.Tensor) -> torch.Tensor:
        """Band |q| <= band - 1 of g (ny, nx) resampled to the Lr x Lc grid."""
        ny, nx, Lr, Lc = self.ny, self.nx, self.Lr, self.Lc
        if (Lr, Lc) == (ny, nx):
            return g.to(torch.float64)
        gh = torch.fft.fft2(g.to(torch.float64)) / (ny * nx)
        Br, Bc = self.band[0] - 1, self.band[1] - 1
        qr = torch.arange(-Br, Br + 1, device=self.dev)
        qc = torch.arange(-Bc, Bc + 1, device=self.dev)
        sub = gh[qr.remainder(ny)][:, qc.remainder(nx)]
        gl = torch.zeros((Lr, Lc), dtype=torch.complex128, device=self.dev)
        gl[qr.remainder(Lr)[:, None], qc.remainder(Lc)[None, :]] = sub
        return (torch.fft.ifft2(gl) * (Lr * Lc)).real

    def source_gradient(self, V: torch.Tensor, grad_image: torch.Tensor) -> torch.Tensor:
        """dL/dV (float64, nsrc) for image cotangents grad_image (nFocus, ny, nx)."""
        V = V.detach().to(self.dev, torch.float64)
        W = float(V.sum())
        g = grad_image.detach().to(self.dev, torch.float32)
        img = self.image(V)
        scale = (self.ny * self.nx) / (self.Lr * self.Lc)
        out = torch.zeros(self.nsrc, dtype=torch.float64, device=self.dev)
        for i in range(len(self.foci)):
            gam = self._gamma(g[i])
            J = self.point_intensities(i).to(torch.float64)
            gJ = scale * (J.reshape(self.nsrc, -1) @ gam.reshape(-1))
            gI = float((g[i].double() * img[i].double()).sum())
            out += self.n_index / W * gJ - gI / W
        return out


class AbbeSourceImage(torch.autograd.Function):
    """I(V) of an AbbeSourceProblem; backward = source_gradient."""

    @staticmethod
    def forward(ctx, V: torch.Tensor, problem: AbbeSourceProblem) -> torch.Tensor:
        ctx.problem = problem
        ctx.save_for_backward(V)
        return problem.image(V)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (V,) = ctx.saved_tensors
        g = ctx.problem.source_gradient(V, grad_out)
        return g.to(device=V.device, dtype=V.dtype), None


def abbe_source_problem(field, source, scatter, projector, numerics) -> AbbeSourceProblem:
    return AbbeSourceProblem(field, source, scatter, projector, numerics)


def abbe_image_weighted(V: torch.Tensor, problem: AbbeSourceProblem) -> torch.Tensor:
    """Differentiable Abbe intensities (nFocus, ny, nx) as a function of the
    source weights V (nsrc,)."""
    if V.dim() != 1 or V.numel() != problem.nsrc:
        raise ValueError(f"V must hold {problem.nsrc} source weights")
    return AbbeSourceImage.apply(V, problem)
