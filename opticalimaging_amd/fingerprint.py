"""Canonical, hashable forms of parameter objects.

A parameter object (Source, Projection, Numerics, or any duck-typed object
the drop-in entry points accept) enters by its public, non-callable
attributes; tensors and arrays enter by shape, dtype and a digest of their
bytes.  Used to key warm-start files and in-process caches, so that a
changed parameter never reuses a result computed for another.
"""
from __future__ import annotations

import hashlib
import json

import numpy as np
import torch

__all__ = ["canon", "params", "digest"]


_DTYPE_NAMES = {}


def _dtype_name(dt) -> str:
    name = _DTYPE_NAMES.get(dt)
    if name is None:
        name = _DTYPE_NAMES[dt] = str(dt)
    return name


def canon(v, depth=0):
    """A JSON-able form of a parameter value; arrays enter by shape, dtype
    and a digest of their bytes."""
    if depth > 4:
        raise ValueError("parameter objects nest too deeply to fingerprint")
    t = type(v)
    if t is float:
        return repr(v)
    if v is None or t is int or t is str or t is bool or isinstance(v, (bool, int, str)):
        return v
    if isinstance(v, float):
        return repr(v)
    if isinstance(v, complex):
        return [repr(v.real), repr(v.imag)]
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().contiguous().numpy()
    if isinstance(v, np.ndarray):
        a = np.ascontiguousarray(v)
        return {"shape": list(a.shape), "dtype": _dtype_name(a.dtype),
                "sha256": hashlib.sha256(a.tobytes()).hexdigest()}
    if isinstance(v, np.generic):
        return canon(v.item(), depth)
    if isinstance(v, (list, tuple)):
        return [canon(x, depth + 1) for x in v]
    if isinstance(v, dict):
        return {str(k): canon(x, depth + 1) for k, x in sorted(v.items())}
    return params(v, depth + 1)


def params(obj, depth=0, exclude=()):
    """Public, non-callable attributes of a parameter object (instance and
    class attributes, as the reference's Numerics keeps its settings on the
    class), except those named in `exclude`."""
    out = {"__class__": f"{type(obj).__module__}.{type(obj).__qualname__}"}
    for k in sorted(dir(obj)):
        if k.startswith("_") or k in exclude:
            continue
        v = getattr(obj, k)
        if callable(v):
            continue
        out[k] = canon(v, depth)
    return out


def digest(obj, exclude=()) -> str:
    """sha256 of the canonical JSON of params(obj)."""
    blob = json.dumps(params(obj, exclude=exclude), sort_keys=True, separators=(",", ":")).encode()
    return hashlib.sha256(blob).hexdigest()
