"""Builds libopticalimaging_amd.so for gfx950 with hipcc (in-tree).

Usage:  python -m opticalimaging_amd.csrc.build   (or call build()).
Each .hip source compiles to an object in parallel, then one shared library is
linked next to the package (opticalimaging_amd/libopticalimaging_amd.so).
Packed-fp32 VALU forms are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libopticalimaging_amd.so"
LIB_PATH = os.path.join(PKG, LIB_NAME)
ARCH = os.environ.get("OI_OFFLOAD_ARCH", "gfx950")

SOURCES = ["oi_api.hip", "oi_imaging.hip", "oi_fastpath.hip", "oi_pupil.hip", "oi_tcc.hip",
           "oi_eigh.hip", "oi_adjoint.hip", "oi_ilt.hip", "oi_rcwa.hip"]

COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-I", INCLUDE, "-I", HERE,
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP toolchain is required to build")
    return path


def _compile(src: str, obj: str) -> None:
    cmd = [_hipcc(), *COMMON_FLAGS, "-c", os.path.join(HERE, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    with tempfile.TemporaryDirectory(prefix="oi_build_") as tmp:
        objs = [os.path.join(tmp, s.replace(".hip", ".o")) for s in srcs]
        workers = min(len(srcs), max(1, (os.cpu_count() or 2) // 2), 8)
        with concurrent.futures.ThreadPoolExecutor(workers) as ex:
            list(ex.map(_compile, srcs, objs))
        tmp_lib = os.path.join(tmp, LIB_NAME)
        cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", *objs,
               "-o", tmp_lib]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed:\n{res.stderr}")
        shutil.move(tmp_lib, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
