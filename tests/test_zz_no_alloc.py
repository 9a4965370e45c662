"""oi_socs_image_noalloc (the plan-free, allocation-free form of
oi_socs_image) through the C ABI: two calls in a row, the second with a
smaller input, each against oi_socs_image on a plan made for the same batch
(float32 images; tolerance 1e-6 of the image peak, the kernel groups being
the same the images agree bit for bit unless the device-built twiddle tables
differ from the host-built ones in a last bit)."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import annular, numerics, projector

pytestmark = pytest.mark.gpu


def _kernel_set(n, period, order):
    from opticalimaging_amd import engine
    src, proj, num = annular(), projector(), numerics(order=order)
    tcc, fg = O.tcc_matrix(src, [period, period], proj, 0.0, num)
    kern = O.decompose_tcc_socs(tcc, fg, [n, n], num)
    return engine.KernelSet.from_dense(torch.as_tensor(kern.astype(np.complex64)))


def _masks(n, b, seed):
    rng = np.random.default_rng(seed)
    return torch.as_tensor((rng.random((b, n, n)) < 0.4).astype(np.float32)).cuda()


def _noalloc(ks, masks):
    from opticalimaging_amd import _native as nat
    B, ny, nx = masks.shape
    g = nat.Geometry(ny=ny, nx=nx, k0r=ks.k0r, wr=ks.wr, k0c=ks.k0c, wc=ks.wc, nchan=1)
    out = torch.full((B, ny, nx), float("nan"), dtype=torch.float32, device="cuda")
    chans, keep = nat.int_array([0])
    nat.check(nat.load().oi_socs_image_noalloc(ctypes.byref(g), nat.ptr(masks), nat.OI_FIELD_REAL,
                                               1, chans, B, nat.ptr(ks.win), ks.K, nat.ptr(out),
                                               nat.stream()))
    return out


def _with_plan(ks, masks):
    from opticalimaging_amd import _native as nat, engine
    B, ny, nx = masks.shape
    plan = engine.Plan(ny, nx, ks.k0r, ks.wr, ks.k0c, ks.wc, nchan=1, max_batch=B, max_kernels=ks.K)
    out = torch.empty((B, ny, nx), dtype=torch.float32, device="cuda")
    chans, keep = nat.int_array([0])
    nat.check(nat.load().oi_socs_image(plan.handle, nat.ptr(masks), nat.OI_FIELD_REAL, 1, chans, B,
                                       nat.ptr(ks.win), ks.K, nat.ptr(out), nat.stream()))
    torch.cuda.synchronize()
    return out


def test_noalloc_matches_planned_image_twice(gpu):
    cases = [("1024^2 x 40, K=24 (passes over the scratch)", 1024, 4096.0, 24, 40, 11),
             ("256^2 x 3, K=12", 256, 1024.0, 12, 3, 12)]
    for name, n, period, order, b, seed in cases:
        ks = _kernel_set(n, period, order)
        masks = _masks(n, b, seed)
        print(f"noalloc case {name}", flush=True)
        got = _noalloc(ks, masks)
        torch.cuda.synchronize()
        ref = _with_plan(ks, masks)
        assert not torch.isnan(got).any()
        diff = float((got - ref).abs().max())
        peak = float(ref.abs().max())
        print(f"  max |noalloc - planned| = {diff:.3e} (peak {peak:.3e}), "
              f"bit-identical: {bool(torch.equal(got, ref))}", flush=True)
        assert diff <= 1e-6 * peak


def test_noalloc_rejects_an_image_beyond_its_scratch(gpu):
    from opticalimaging_amd import _native as nat
    n = 65536                     # one image's row spectra (65 x 65536 c64) exceed the scratch
    g = nat.Geometry(ny=n, nx=n, k0r=-32, wr=65, k0c=-32, wc=65, nchan=1)
    dummy = torch.zeros(16, dtype=torch.float32, device="cuda")
    chans, keep = nat.int_array([0])
    rc = nat.load().oi_socs_image_noalloc(ctypes.byref(g), nat.ptr(dummy), nat.OI_FIELD_REAL, 1,
                                          chans, 1, nat.ptr(dummy), 1, nat.ptr(dummy), nat.stream())
    assert rc == nat.OI_ERR_INVALID
    assert "bytes of scratch" in nat.load().oi_last_error().decode()
