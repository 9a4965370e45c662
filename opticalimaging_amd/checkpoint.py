"""Warm start for SOCS kernel sets.

A device-resident ``engine.KernelSet`` is written to an ``.npz`` file (plain
numpy arrays; nothing pickled) together with a fingerprint of everything its
kernels depend on: the source, projector and numerics parameters, the period,
the focus, the mask grid and the library version.  ``load_kernel_set``
returns it only when the fingerprint it is given matches the stored one, so a
changed optic never images with stale kernels.  The reference has no
counterpart: it recomputes the TCC and its decomposition on every call
(CalculateAerialImage.py:105-130).
"""
from __future__ import annotations

import hashlib
import json
import os
import tempfile

import numpy as np
import torch

from . import _native as nat
from . import engine, fingerprint

__all__ = ["optics_fingerprint", "save_kernel_set", "load_kernel_set", "cached_kernel_set"]

FORMAT = 1
# results a call leaves on the parameter objects, not inputs to the kernels
_DERIVED = {"source_data"}


def _canon(v, depth=0):
    return fingerprint.canon(v, depth)


def _params(obj, depth=0):
    return fingerprint.params(obj, depth, exclude=_DERIVED)


def optics_fingerprint(source, projector, numerics, xypitch, focus, Nfg) -> str:
    """sha256 over the canonical JSON of every input of
    DecomposeTCC_SOCS.socs_kernel_set(source, xypitch, projector, focus,
    numerics, Nfg) and the library version."""
    parts = {
        "format": FORMAT,
        "library": int(nat.load().oi_version()),
        "source": _params(source),
        "projector": _params(projector),
        "numerics": _params(numerics),
        "xypitch": [_canon(float(x)) for x in xypitch],
        "focus": _canon(float(focus)),
        "Nfg": [int(x) for x in Nfg],
    }
    blob = json.dumps(parts, sort_keys=True, separators=(",", ":")).encode()
    return hashlib.sha256(blob).hexdigest()


def save_kernel_set(path: str, ks: engine.KernelSet, fingerprint: str) -> None:
    """Write ks (windowed kernels, placement, band) and its fingerprint to
    path (.npz), atomically (a temporary file renamed over path)."""
    win = ks.win.detach().to("cpu", torch.complex64).contiguous().numpy()
    band = np.array(ks.band if ks.band is not None else (-1, -1), dtype=np.int64)
    meta = np.array([FORMAT, ks.k0r, ks.k0c, ks.ny, ks.nx], dtype=np.int64)
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=d, suffix=".npz")
    os.close(fd)
    try:
        np.savez(tmp, win=win, meta=meta, band=band,
                 fingerprint=np.frombuffer(fingerprint.encode(), dtype=np.uint8))
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)


def load_kernel_set(path: str, fingerprint: str | None = None, device=None) -> engine.KernelSet:
    """The kernel set saved at path, on `device` (default: the HIP device).
    Raises ValueError when the file's format is not this one's or when
    `fingerprint` is given and differs from the stored one."""
    with np.load(path, allow_pickle=False) as z:
        meta = z["meta"]
        if int(meta[0]) != FORMAT:
            raise ValueError(f"{path}: kernel-set format {int(meta[0])}, expected {FORMAT}")
        stored = bytes(z["fingerprint"]).decode()
        if fingerprint is not None and stored != fingerprint:
            raise ValueError(f"{path}: kernel set was built for other optics (fingerprint mismatch)")
        win = torch.from_numpy(np.ascontiguousarray(z["win"]))
        band = tuple(int(x) for x in z["band"])
    dev = nat.device() if device is None else torch.device(device)
    _, k0r, k0c, ny, nx = (int(x) for x in meta)
    return engine.KernelSet(win.to(dev), k0r, k0c, ny, nx, None if band == (-1, -1) else band)


def cached_kernel_set(cache_dir: str, source, xypitch, projector, focus, numerics, Nfg):
    """DecomposeTCC_SOCS.socs_kernel_set(...) through a cache directory: the
    file named by the optics fingerprint is loaded when present, otherwise
    the set is built and saved there.  Returns (kernel set, loaded)."""
    from .DecomposeTCC_SOCS import socs_kernel_set
    key = optics_fingerprint(source, projector, numerics, xypitch, focus, Nfg)
    path = os.path.join(cache_dir, f"socs_{key[:32]}.npz")
    if os.path.exists(path):
        return load_kernel_set(path, key), True
    ks = socs_kernel_set(source, xypitch, projector, focus, numerics, Nfg)
    save_kernel_set(path, ks, key)
    return ks, False
