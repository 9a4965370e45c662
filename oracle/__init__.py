"""FP64 CPU oracle for the aerial-image path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import, call or
execute anything under ``oracle/``; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only to check results
(or as the timed CPU baseline), never as the thing measured or shipped.

It is a float64 restatement of the reference imaging path of
gary21978/OpticalImaging.  Every function names the reference file:line it
follows.  The reference itself computes in complex64/float32 throughout
(``Scatter.py:14``, ``CalculateAerialImage.py:47,55,69``,
``CalculateTCCMatrix.py:35``, ``DecomposeTCC_SOCS.py:24-25``), so no float64
path exists in the reference; this module is that path, written here.

Discrete decisions (which source-grid points are lit, TCC grid sizes) are taken
exactly as the reference takes them, in float32 torch arithmetic, because a
float64 decision selects a different point set (eight annular points sit at
r = 0.5 exactly).  Continuous values are float64.

Pins (what each function is checked against) are listed in the docstring of
``oracle.imaging_fp64`` and exercised in ``tests/test_oracle_*.py``.
"""
from .imaging_fp64 import *  # noqa: F401,F403
