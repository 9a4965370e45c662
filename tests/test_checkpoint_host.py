"""Kernel-set checkpoints on the host: the optics fingerprint sees every
input of the kernels and nothing a call leaves behind; save / load round
trips; a stale or foreign file is refused."""
import os

import numpy as np
import pytest
import torch

from opticalimaging_amd import checkpoint, engine, workloads


def _optics():
    return workloads.socs_1024_optics(order=24, grid_halfwidth=50)


def _key(src, proj, num, xy=(4096.0, 4096.0), focus=0.0, nfg=(1024, 1024)):
    return checkpoint.optics_fingerprint(src, proj, num, list(xy), focus, list(nfg))


def test_fingerprint_ignores_call_results():
    src, proj, num = _optics()
    k0 = _key(src, proj, num)
    src.Calc_SourceSimple()          # leaves source.source_data behind
    assert _key(src, proj, num) == k0
    src2, proj2, num2 = _optics()
    assert _key(src2, proj2, num2) == k0


@pytest.mark.parametrize("change", [
    lambda s, p, n: setattr(s, "SigmaOut", 0.81),
    lambda s, p, n: setattr(s, "Wavelength", 248),
    lambda s, p, n: setattr(p, "NA", 0.86),
    lambda s, p, n: setattr(p, "Aberration_Zernike", p.Aberration_Zernike + torch.nn.functional.one_hot(
        torch.tensor(8), p.Aberration_Zernike.numel()).to(p.Aberration_Zernike.dtype) * 1e-3),
    lambda s, p, n: setattr(n, "Hopkins_Order", 23),
    lambda s, p, n: setattr(n, "TCC_Precision", "fp64"),
])
def test_fingerprint_sees_every_parameter(change):
    src, proj, num = _optics()
    k0 = _key(src, proj, num)
    change(src, proj, num)
    assert _key(src, proj, num) != k0


def test_fingerprint_sees_period_focus_and_grid():
    src, proj, num = _optics()
    k0 = _key(src, proj, num)
    assert _key(src, proj, num, xy=(4096.0, 4000.0)) != k0
    assert _key(src, proj, num, focus=10.0) != k0
    assert _key(src, proj, num, nfg=(512, 512)) != k0


def _kernel_set():
    rng = np.random.default_rng(3)
    win = torch.as_tensor((rng.normal(size=(5, 7, 9)) + 1j * rng.normal(size=(5, 7, 9)))
                          .astype(np.complex64))
    return engine.KernelSet(win, -3, -4, 64, 96, (5, 6))


def test_save_load_round_trip(tmp_path):
    ks = _kernel_set()
    path = os.path.join(tmp_path, "k.npz")
    checkpoint.save_kernel_set(path, ks, "abc123")
    back = checkpoint.load_kernel_set(path, "abc123", device="cpu")
    assert torch.equal(back.win, ks.win)
    assert (back.k0r, back.k0c, back.ny, back.nx, back.band) == (-3, -4, 64, 96, (5, 6))
    assert sorted(os.listdir(tmp_path)) == ["k.npz"]       # no temporary file left
    ks.band = None
    checkpoint.save_kernel_set(path, ks, "abc123")
    assert checkpoint.load_kernel_set(path, device="cpu").band is None


def test_stale_or_foreign_files_are_refused(tmp_path):
    ks = _kernel_set()
    path = os.path.join(tmp_path, "k.npz")
    checkpoint.save_kernel_set(path, ks, "abc123")
    with pytest.raises(ValueError, match="fingerprint"):
        checkpoint.load_kernel_set(path, "other", device="cpu")
    with np.load(path) as z:
        parts = {k: z[k] for k in z.files}
    parts["meta"] = parts["meta"].copy()
    parts["meta"][0] = checkpoint.FORMAT + 1
    np.savez(path, **parts)
    with pytest.raises(ValueError, match="format"):
        checkpoint.load_kernel_set(path, device="cpu")
