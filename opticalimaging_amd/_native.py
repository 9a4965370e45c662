"""ctypes binding of libopticalimaging_amd.so (declared in include/opticalimaging_amd.h).

The library is the only compute path: every public imaging function of this
package runs its numerical work through these entry points on the current HIP
device and raises if the library or the device is missing.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

LIB_NAME = "libopticalimaging_amd.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

OI_OK = 0
OI_ERR_INVALID = -1
OI_ERR_RUNTIME = -2
OI_ERR_UNSUPPORTED = -3
OI_ERR_NOMEM = -4

OI_FIELD_REAL = 0
OI_FIELD_C64_CHLAST = 1

OI_TCC_SPLIT_BF16 = 0
OI_TCC_FP32 = 1
OI_TCC_FP64 = 2

_lib = None
_lock = threading.Lock()

c_int = ctypes.c_int
c_double = ctypes.c_double
c_float = ctypes.c_float
c_size_t = ctypes.c_size_t
c_void_p = ctypes.c_void_p
P_int = ctypes.POINTER(ctypes.c_int)
P_float = ctypes.POINTER(ctypes.c_float)
P_double = ctypes.POINTER(ctypes.c_double)


class Geometry(ctypes.Structure):
    """Mirror of struct oi_geometry."""
    _fields_ = [(n, c_int) for n in (
        "ny", "nx", "k0r", "wr", "k0c", "wc", "band_r", "band_c", "nchan", "max_batch",
        "max_groups", "lr", "lc", "rows", "direct", "qr", "qc", "path", "max_kernels")]


class Optics(ctypes.Structure):
    """Mirror of struct oi_optics."""
    _fields_ = [("wavelength", c_double), ("na", c_double), ("magnification", c_double),
                ("index_image", c_double), ("focus", c_double), ("zernike", c_double * 37),
                ("pitch_r", c_double), ("pitch_c", c_double), ("vector", c_int)]


# name -> (restype, argtypes); the complete exported ABI.
SIGNATURES = {
    "oi_last_error": (ctypes.c_char_p, []),
    "oi_version": (c_int, []),
    "oi_plan_create": (c_int, [ctypes.POINTER(Geometry), ctypes.POINTER(c_void_p)]),
    "oi_plan_destroy": (c_int, [c_void_p]),
    "oi_plan_bytes": (c_size_t, [c_void_p]),
    "oi_plan_set_timing": (c_int, [c_void_p, c_int]),
    "oi_plan_read_timing": (c_int, [c_void_p, P_double, P_int]),
    "oi_field_flags": (c_int, [c_void_p, c_int, ctypes.c_longlong, c_int, c_void_p, c_void_p]),
    "oi_spectrum_window": (c_int, [c_void_p, c_void_p, c_int, c_int, P_int, c_int, c_void_p, c_void_p]),
    "oi_coherent_image": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "oi_socs_image": (c_int, [c_void_p, c_void_p, c_int, c_int, P_int, c_int, c_void_p, c_int,
                              c_void_p, c_void_p]),
    "oi_socs_image_noalloc": (c_int, [ctypes.POINTER(Geometry), c_void_p, c_int, c_int, P_int, c_int,
                                      c_void_p, c_int, c_void_p, c_void_p]),
    "oi_spectrum_full": (c_int, [c_void_p, c_int, c_int, P_int, c_int, c_int, c_int, c_int, c_float,
                                 c_void_p, c_void_p, c_void_p]),
    "oi_kernel_support": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "oi_kernel_window": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                 c_void_p, c_void_p]),
    "oi_kernel_dense": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                c_void_p, c_void_p]),
    "oi_mask_decode_u8": (c_int, [c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_float), c_int,
                                  c_void_p, c_void_p]),
    "oi_mask_decode_bits": (c_int, [c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_float),
                                    c_void_p, c_void_p]),
    "oi_pupil_kernels": (c_int, [ctypes.POINTER(Optics), c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "oi_pupil_kernels_foci": (c_int, [ctypes.POINTER(Optics), ctypes.POINTER(ctypes.c_double), c_int,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                      c_int, c_int, c_int, c_void_p, c_void_p]),
    "oi_tcc_work_bytes": (c_size_t, [c_int, c_int, c_int, c_int]),
    "oi_tcc_build": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_int, c_void_p,
                             c_void_p, c_void_p]),
    "oi_tcc_gram": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "oi_kernels_from_gram": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_int,
                                     c_void_p, c_void_p]),
    "oi_eigh_work_bytes": (c_size_t, [c_int]),
    "oi_eigh_hermitian": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_double, c_void_p,
                                  P_int, c_void_p]),
    "oi_zernike": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "oi_char_matrix": (c_int, [c_void_p, c_void_p, c_int, c_double, c_double, c_void_p, c_void_p]),
    "oi_socs_adjoint_work_bytes": (c_size_t, [c_void_p, c_int, c_int]),
    "oi_socs_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                c_void_p, c_void_p]),
    "oi_socs_residual_adjoint": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "oi_ilt_target_band_bytes": (c_size_t, [c_void_p, c_int]),
    "oi_ilt_target_band": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "oi_socs_band_residual_adjoint": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                              c_void_p, c_void_p, c_void_p, c_void_p]),
    "oi_ilt_band_step": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_float,
                                 c_void_p, c_void_p, c_void_p, c_void_p]),
    "oi_ilt_work_floats": (c_size_t, [c_int]),
    "oi_socs_residual": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                 c_void_p, c_void_p]),
    "oi_ilt_residual": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_size_t, c_void_p,
                                c_void_p]),
    "oi_rcwa_convmat": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "oi_rcwa_pq": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                           c_void_p]),
    "oi_rcwa_modes": (c_int, [c_double, c_double, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                              c_void_p]),
    "oi_rcwa_fields": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_double, c_void_p,
                               c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "oi_ilt_update": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_int, c_size_t, c_void_p,
                              c_void_p, c_void_p]),
}


def load():
    """Load the shared library (no GPU needed) and declare its signatures."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_NAME} is not built; run `python -m opticalimaging_amd.csrc.build` "
                    "or __graft_entry__.build()")
            lib = ctypes.CDLL(LIB_PATH)
            missing = []
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name, None)
                if fn is None:
                    missing.append(name)
                    continue
                fn.restype = res
                fn.argtypes = args
            lib.oi_missing = tuple(missing)
            _lib = lib
    return _lib


def missing_symbols():
    """Declared entry points the loaded library does not export."""
    return load().oi_missing


def check(rc: int) -> None:
    """Map a status code to the reference's exception classes."""
    if rc == OI_OK:
        return
    msg = load().oi_last_error().decode(errors="replace")
    if rc == OI_ERR_INVALID:
        raise ValueError(msg)
    raise RuntimeError(f"{msg} (status {rc})")


def device() -> torch.device:
    """The HIP device all work runs on; raises when none is present."""
    if not torch.cuda.is_available():
        raise RuntimeError("opticalimaging_amd requires a HIP (ROCm) GPU; none is visible")
    load()
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> c_void_p:
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None) -> c_void_p:
    if t is None:
        return c_void_p(0)
    if not t.is_cuda:
        raise ValueError("device tensor expected")
    return c_void_p(t.data_ptr())


def int_array(vals):
    arr = (c_int * max(1, len(vals)))(*vals)
    return ctypes.cast(arr, P_int), arr
