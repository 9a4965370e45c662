"""Seeded synthetic masks and optics for the benchmark configurations.

All generators use numpy's PCG64 ``default_rng(seed)``; ``digest`` gives the
SHA-256 of a generated array so a workload can be identified exactly.

Configurations (sizes as stated):
  scalar_abbe_128   128 x 128 binary mask, 1000 nm period, 6 rectangles 40-300 nm,
                    193 nm, NA 0.85, annular 0.5/0.8, PntNum 21, focus 0, scalar.
  vector_abbe_512   512 x 512, 2048 nm, Manhattan 1D/2D lines CD 45-90 nm, att-PSM
                    (intensity transmission 0.06 at 180 degrees: amplitude -sqrt(0.06)),
                    NA 1.35, n 1.44, quasar 0.6/0.9/30 deg, PntNum 41, t_pol,
                    37 Zernikes ~ N(0, 0.01 waves), 5 foci -60..60 nm, M = 1.
  socs_1024         1024 x 1024, 4096 nm, binary rectangles 40-400 nm to >= 35 %
                    coverage; config-1 optics; TCC grid 101 x 101; K = 24.
"""
from __future__ import annotations

import hashlib
import math

import numpy as np
import torch

from .Numerics import Numerics
from .Projection import Projection
from .Source import Source


def digest(a) -> str:
    a = np.ascontiguousarray(np.asarray(a))
    return hashlib.sha256(a.tobytes()).hexdigest()


def _wrap_range(n, pitch, start, length, period):
    """Pixels whose centre lies in [start, start + length) modulo the period."""
    centres = (np.arange(n) + 0.5) * pitch
    return np.mod(centres - start, period) < length


def rectangles(n, period, count, size_range, rng, coverage=None, value=1.0, background=0.0):
    """n x n mask of axis-aligned rectangles (pixel-centre inclusion, periodic
    wrap).  Adds `count` rectangles, or keeps adding until the covered fraction
    reaches `coverage`."""
    pitch = period / n
    m = np.zeros((n, n), dtype=bool)
    added = 0
    while True:
        if coverage is None and added >= count:
            break
        if coverage is not None and m.mean() >= coverage:
            break
        w, h = rng.uniform(size_range[0], size_range[1], 2)
        x0, y0 = rng.uniform(0, period, 2)
        rows = _wrap_range(n, pitch, y0, h, period)
        cols = _wrap_range(n, pitch, x0, w, period)
        m |= rows[:, None] & cols[None, :]
        added += 1
    out = np.full((n, n), background, dtype=np.float32)
    out[m] = value
    return out


def manhattan(n, period, rng, cd_range=(45.0, 90.0), lines=None):
    """Boolean n x n layout of 1D/2D Manhattan lines: full-length vertical and
    horizontal lines plus shorter 2D segments, widths ~ U(cd_range) nm."""
    pitch = period / n
    m = np.zeros((n, n), dtype=bool)
    cols = np.zeros(n, dtype=bool)      # union of vertical 1D lines
    rows = np.zeros(n, dtype=bool)      # union of horizontal 1D lines
    count = lines if lines is not None else max(4, int(period / 200))
    for i in range(count):
        cd = rng.uniform(*cd_range)
        start = rng.uniform(0, period)
        kind = rng.integers(0, 3)
        if kind == 0:       # vertical 1D line
            cols |= _wrap_range(n, pitch, start, cd, period)
        elif kind == 1:     # horizontal 1D line
            rows |= _wrap_range(n, pitch, start, cd, period)
        else:               # 2D segment
            length = rng.uniform(2 * cd, max(4 * cd, period / 2))
            other = rng.uniform(0, period)
            if rng.integers(0, 2):
                r = _wrap_range(n, pitch, other, length, period)
                c = _wrap_range(n, pitch, start, cd, period)
            else:
                r = _wrap_range(n, pitch, start, cd, period)
                c = _wrap_range(n, pitch, other, length, period)
            m[np.ix_(np.nonzero(r)[0], np.nonzero(c)[0])] = True
    m |= cols[None, :]
    m |= rows[:, None]
    return m


ATT_PSM_AMPLITUDE = -math.sqrt(0.06)   # 6 % intensity transmission, 180 degree phase


def att_psm_mask(n, period, seed):
    """Attenuated-PSM mask: clear background 1, absorber lines -sqrt(0.06)."""
    rng = np.random.default_rng(seed)
    m = manhattan(n, period, rng)
    out = np.ones((n, n), dtype=np.float32)
    out[m] = ATT_PSM_AMPLITUDE
    return out


def scalar_abbe_128(seed=1):
    rng = np.random.default_rng(seed)
    mask = rectangles(128, 1000.0, 6, (40.0, 300.0), rng)
    src = Source()
    src.PntNum = 21
    src.Wavelength = 193
    src.Shape = "annular"
    src.SigmaIn = 0.5
    src.SigmaOut = 0.8
    proj = Projection()
    proj.NA = 0.85
    proj.IndexImage = 1.0
    proj.FocusRange = torch.tensor([0])
    num = Numerics()
    num.ImageCalculationMode = "scalar"
    num.ImageCalculationMethod = "abbe"
    return mask, src, proj, num, 1000.0


def vector_abbe_512(seed=2, n=512):
    period = 4.0 * n
    mask = att_psm_mask(n, period, seed)
    rng = np.random.default_rng(seed + 1000)
    src = Source()
    src.PntNum = 41
    src.Wavelength = 193
    src.Shape = "quasar"
    src.SigmaIn = 0.6
    src.SigmaOut = 0.9
    src.OpenAngle = math.pi / 6
    src.PolarizationType = "t_pol"
    proj = Projection()
    proj.NA = 1.35
    proj.IndexImage = 1.44
    proj.Magnification = 1
    proj.Aberration_Zernike = torch.tensor(rng.normal(0.0, 0.01, 37), dtype=torch.float32)
    proj.FocusRange = torch.linspace(-60, 60, 5)
    num = Numerics()
    num.ImageCalculationMode = "vector"
    num.ImageCalculationMethod = "abbe"
    return mask, src, proj, num, period


def socs_1024_masks(batch=64, n=1024, seed=3, coverage=0.35):
    period = 4.0 * n
    out = np.empty((batch, n, n), dtype=np.float32)
    for b in range(batch):
        rng = np.random.default_rng(seed * 100003 + b)
        out[b] = rectangles(n, period, 0, (40.0, 400.0), rng, coverage=coverage)
    return out, period


def socs_1024_optics(order=24, grid_halfwidth=50):
    _, src, proj, num, _ = scalar_abbe_128()
    num.ImageCalculationMethod = "hopkins"
    num.Hopkins_SettingType = "order"
    num.Hopkins_Order = order
    num.Hopkins_GridHalfWidth = grid_halfwidth
    num.TCC_Precision = "split_bf16"
    return src, proj, num


def field_from_mask(mask):
    """ScatterField (Ny, Nx, 3): channel 0 the mask, channels 1-2 zero."""
    m = torch.as_tensor(mask)
    f = torch.zeros(m.shape + (3,), dtype=m.dtype)
    f[..., 0] = m
    return f


# Differentiation workload: SOCS kernels of the vector quasar optics above at
# focus 0, continuous masks sigmoid(theta), theta ~ N(0, 1), binary targets.
ILT_STEP = 0.5          # step of normalised gradient descent in theta units


def ilt_optics(order=24):
    _, src, proj, num, _ = vector_abbe_512()
    proj.FocusRange = torch.tensor([0])
    num.ImageCalculationMethod = "hopkins"
    num.Hopkins_SettingType = "order"
    num.Hopkins_Order = order
    num.TCC_Precision = "split_bf16"
    return src, proj, num


def ilt_problem(batch=16, n=1024, seed=5):
    """(theta [batch, n, n] float32 ~ N(0, 1), targets [batch, n, n] float32 in {0, 1})
    over an n x n grid of 4 nm pixels; targets are seeded Manhattan layouts."""
    period = 4.0 * n
    rng = np.random.default_rng(seed)
    theta = rng.standard_normal((batch, n, n)).astype(np.float32)
    T = np.empty((batch, n, n), dtype=np.float32)
    for b in range(batch):
        T[b] = manhattan(n, period, np.random.default_rng(seed * 7919 + b)).astype(np.float32)
    return theta, T


# Serving workload: Poisson arrivals, clip sizes uniform over SERVING_SIZES,
# masks as the attenuated-PSM layouts above (4 nm pixels) from seed = request id.
SERVING_SIZES = (256, 512, 1024, 2048)


def serving_arrivals(count, rate, sizes=SERVING_SIZES, seed=17):
    """(arrival seconds, clip size) per request: exponential inter-arrival
    times at `rate` requests/s and sizes drawn uniformly, both seeded."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=count)
    t = np.cumsum(gaps) - gaps[0]
    n = rng.choice(np.asarray(sizes), size=count)
    return t, n


def serving_mask(rid, n):
    return att_psm_mask(n, 4.0 * n, seed=int(rid))


ATT_PSM_LEVELS = (1.0, float(np.float32(ATT_PSM_AMPLITUDE)))


def serving_mask_codes(rid, n):
    """The same clip as serving_mask as (uint8 codes, levels): code 1 marks
    absorber pixels; levels[code] is bit-identical to serving_mask's float32."""
    rng = np.random.default_rng(int(rid))
    return manhattan(n, 4.0 * n, rng).astype(np.uint8), ATT_PSM_LEVELS
