"""Float64 restatement of the reference imaging path (test infrastructure only).

Reference: gary21978/OpticalImaging.  Each function cites the file:line it
restates.  Parameter objects are duck-typed on the reference attribute names
(``Source.py:16-45``, ``Projection.py:5-10``, ``Scatter.py:4-8``,
``Numerics.py:3-13``); nothing from the product package is imported.

Pins (tests/test_oracle_*.py):
  * source_points         -- point counts / weight sums / 4-fold symmetry for
                             the annular and quasar sets, and pixel-source
                             pass-through (closed forms).
  * zernike_fringe        -- general radial-polynomial formula checked against
                             the textbook Fringe list Z1..Z37, values at rho=0
                             and rho=1, and orthogonality on the unit disk.
  * characteristic_matrix -- transversality alpha*Mx + beta*My + gamma*Mz = 0
                             and identity at normal incidence.
  * spectrum              -- brute-force DFT on small grids, DC term.
  * abbe_image            -- brute-force DFT image, clear-field closed form,
                             dark mask, circular-shift equivariance.
  * tcc_matrix            -- brute-force definition, Hermitian PSD, trace,
                             clear-field entry.
  * decompose_tcc_socs    -- reconstruction of the TCC at full rank, trace,
                             Gram-route equivalence.
  * socs_image            -- full-order SOCS equals Abbe (even grids),
                             clear-field value.
  * calculate_socs        -- one order: constant |FM phi|^2; two orders: the
                             closed-form fringe (sign, axis order, half-pixel
                             offset); channel and kernel sums.
"""
from __future__ import annotations

import math

import numpy as np
import torch

__all__ = [
    "source_points", "polarization", "zernike_fringe", "FRINGE_NM",
    "characteristic_matrix", "spectrum", "abbe_image", "tcc_grid",
    "shifted_pupil_matrix", "get_tcc_matrix", "tcc_matrix", "eig_desc",
    "socs_count", "decompose_tcc_socs", "decompose_lowrank", "embed_kernels",
    "socs_image", "hopkins_image", "dft2_bruteforce", "clear_field_intensity",
    "calculate_socs", "optimized_image", "abbe_source_terms",
]


# ---------------------------------------------------------------------------
# Source (Source.py).  Decisions in float32 torch, exactly as the reference.
# ---------------------------------------------------------------------------

def _source_matrix(src):
    """P x P source weights; restates Source.Calc_SourceAll (Source.py:64-149)
    and the shape generators (Source.py:229-286).  Returns float32 torch
    tensors X, Y, V of shape (P, P)."""
    P = int(src.PntNum)
    coord = torch.linspace(-1, 1, P)                       # Source.py:65
    shape = str(src.Shape).lower()
    if shape == "pixel":                                    # Source.py:66-89
        for name in ("SPointX", "SPointY", "SPointValue"):
            if getattr(src, name).numel() != P * P:
                raise ValueError("pixel source arrays must hold PntNum**2 values")
        X = src.SPointX.clone().float()
        Y = src.SPointY.clone().float()
        V = src.SPointValue.clone().float()
        if V.dim() != 2:
            raise ValueError("pixel source arrays must be PntNum x PntNum")
    else:
        X, Y = torch.meshgrid(coord, coord, indexing="ij")
        V = torch.zeros((P, P))
        if shape == "annular":                              # Source.py:229-238
            r = torch.sqrt(X.pow(2) + Y.pow(2))
            lit = (r <= src.SigmaOut) & (r >= src.SigmaIn)
            V[lit] = 1.0 / ((src.SigmaOut ** 2 - src.SigmaIn ** 2) * math.pi)
        elif shape == "quasar":                             # Source.py:97-112,240-256
            oa, ra = src.OpenAngle, src.RotateAngle
            if ra > math.pi / 2 or ra < -math.pi / 2:
                raise ValueError("rotate angle must be in [-pi/2, pi/2]")
            if oa <= 0 or oa >= math.pi / 2:
                raise ValueError("open angle must be in (0, pi/2)")
            r = torch.sqrt(X.pow(2) + Y.pow(2))
            th = torch.atan2(Y, X)
            a = torch.abs(th)
            on_x = (a <= oa / 2) | ((a >= math.pi - oa / 2) & (a <= math.pi))
            on_y = (torch.abs(th - math.pi / 2) <= oa / 2) | (torch.abs(th + math.pi / 2) <= oa / 2)
            lit = (r <= src.SigmaOut) & (r >= src.SigmaIn) & (on_x | on_y)
            V[lit] = 0.5 / ((src.SigmaOut ** 2 - src.SigmaIn ** 2) * oa)
        elif shape == "dipolecirc":                         # Source.py:114-128,258-270
            oa, ra = src.OpenAngle, src.RotateAngle
            if ra > math.pi / 2 or ra < -math.pi / 2:
                raise ValueError("rotate angle must be in [-pi/2, pi/2]")
            if oa < 0 or oa > math.pi / 2:
                raise ValueError("open angle must be in [0, pi/2]")
            r = torch.sqrt(X.pow(2) + Y.pow(2))
            th = torch.atan2(Y, X)
            lit = (r <= src.SigmaOut) & (r >= src.SigmaIn) & (torch.abs(torch.cos(th - ra)) >= math.cos(oa / 2))
            V[lit] = 1.0 / ((src.SigmaOut ** 2 - src.SigmaIn ** 2) * oa)
        elif shape == "multipole":                          # Source.py:130-139,273-286
            ra = src.RotateAngle
            if ra > math.pi / 2 or ra < -math.pi / 2:
                raise ValueError("rotate angle must be in [-pi/2, pi/2]")
            step = 2 * math.pi / src.PoleNumber
            for i in range(src.PoleNumber):
                xc = src.SigmaCenter * math.cos(ra + i * step)
                yc = src.SigmaCenter * math.sin(ra + i * step)
                lit = ((X - xc).pow(2) + (Y - yc).pow(2)) <= src.SigmaRadius ** 2
                V[lit] = 1.0 / src.PoleNumber / (src.SigmaRadius ** 2)
        else:
            raise ValueError("unsupported illumination")
    if src.Source_Mask is not None:                         # Source.py:146-149
        V = V * src.Source_Mask
    sx, sy = V.shape
    if sx == sy:
        if src.PSFEnable:                                   # Source.py:153-176
            ks = round(P / 10) * 2 + 1
            edge = 1 / (P - 1) * (ks - 1)
            kx, ky = torch.meshgrid(torch.linspace(-edge, edge, ks),
                                    torch.linspace(-edge, edge, ks), indexing="ij")
            ker = 1 / math.sqrt(2 * math.pi) / src.PSFSigma * torch.exp(
                -(kx.pow(2) + ky.pow(2)) / src.PSFSigma ** 2)
            ker = ker[~torch.all(ker < 1e-6, 0)]
            kt = torch.transpose(ker, 0, 1)
            ker = kt[~torch.all(kt < 1e-6, 1)]
            ker = ker / torch.sum(ker)
            V = torch.nn.functional.conv2d(V.reshape(1, 1, sx, sy),
                                           ker.reshape(1, 1, *ker.shape),
                                           padding="same").reshape(sx, sy)
        V = V.clone()
        V[int((sx - 1) / 2), int((sy - 1) / 2)] = 0         # Source.py:178
        V[(X ** 2 + Y ** 2) > 1] = 0                         # Source.py:289-291
    return X, Y, V


def source_points(src):
    """Lit source points (X, Y, V) as float64 arrays holding the reference's
    float32 values.  Restates Calc_SourceSimple (Source.py:47-54): flatten
    (Source.py:293-295) and drop weights below 1e-5 (Source.py:49)."""
    X, Y, V = _source_matrix(src)
    X = X.reshape(-1)
    Y = Y.reshape(-1)
    V = V.reshape(-1)
    keep = ~(V < 1e-5)
    return (X[keep].double().numpy(), Y[keep].double().numpy(),
            V[keep].double().numpy())


def polarization(src, theta, rho):
    """Per-source polarisation (PX, PY), complex128.

    Live reference behaviour (Source.py:222-226) broadcasts the constant
    ``PolarizationVector``.  When the source carries a non-None
    ``PolarizationType`` the per-type formulas that the reference keeps only
    as a specification string (Source.py:184-221) are used, including the
    zeroing at rho < machine epsilon (Source.py:215-218)."""
    theta = np.asarray(theta, dtype=np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    ptype = getattr(src, "PolarizationType", None)
    one = np.ones_like(theta)
    if ptype is None:
        return (complex(src.PolarizationVector[0]) * one.astype(np.complex128),
                complex(src.PolarizationVector[1]) * one.astype(np.complex128))
    if ptype == "x_pol":
        px, py = one, 0 * one
    elif ptype == "y_pol":
        px, py = 0 * one, one
    elif ptype == "c_pol":
        px, py = math.sqrt(.5) * one, 1j * math.sqrt(.5) * one
    elif ptype == "d_pol":
        px, py = math.sqrt(.5) * one, math.sqrt(.5) * one
    elif ptype == "r_pol":
        px, py = np.cos(theta), np.sin(theta)
    elif ptype == "t_pol":
        px, py = np.sin(theta), -np.cos(theta)
    elif ptype == "line_pol":
        ang = float(src.PolarizationParameters.Angle)
        px, py = math.sin(ang) * one, math.cos(ang) * one
    else:
        raise ValueError("unsupported polarization type")
    px = np.array(px, dtype=np.complex128)
    py = np.array(py, dtype=np.complex128)
    tiny = rho < np.finfo(np.float64).eps
    px[tiny] = 0
    py[tiny] = 0
    return px, py


# ---------------------------------------------------------------------------
# Optics
# ---------------------------------------------------------------------------

# Fringe ordering Z1..Z37 as (n, m, kind); kind 0: rotationally symmetric,
# 1: cos(m theta), -1: sin(m theta).  This is the ordering behind the explicit
# polynomials of Projection.py:56-127.
FRINGE_NM = [
    (0, 0, 0), (1, 1, 1), (1, 1, -1), (2, 0, 0), (2, 2, 1), (2, 2, -1),
    (3, 1, 1), (3, 1, -1), (4, 0, 0), (3, 3, 1), (3, 3, -1), (4, 2, 1),
    (4, 2, -1), (5, 1, 1), (5, 1, -1), (6, 0, 0), (4, 4, 1), (4, 4, -1),
    (5, 3, 1), (5, 3, -1), (6, 2, 1), (6, 2, -1), (7, 1, 1), (7, 1, -1),
    (8, 0, 0), (5, 5, 1), (5, 5, -1), (6, 4, 1), (6, 4, -1), (7, 3, 1),
    (7, 3, -1), (8, 2, 1), (8, 2, -1), (9, 1, 1), (9, 1, -1), (10, 0, 0),
    (12, 0, 0),
]


def _radial(n, m, rho):
    """Zernike radial polynomial R_n^m by its factorial sum."""
    out = np.zeros_like(rho)
    for s in range((n - m) // 2 + 1):
        c = ((-1) ** s * math.factorial(n - s)
             / (math.factorial(s) * math.factorial((n + m) // 2 - s)
                * math.factorial((n - m) // 2 - s)))
        out = out + c * rho ** (n - 2 * s)
    return out


def zernike_fringe(rho, theta, coeffs):
    """Wavefront W(rho, theta) in waves; restates
    Projection.CalculateAberrationFast (Projection.py:13-128) with
    Orientation = 0 (the only value on the imaging path,
    CalculateAerialImage.py:66, CalculateTCCMatrix.py:34)."""
    rho = np.asarray(rho, dtype=np.float64)
    theta = np.asarray(theta, dtype=np.float64)
    coeffs = np.asarray(coeffs, dtype=np.float64).reshape(-1)
    w = np.zeros(np.broadcast(rho, theta).shape)
    for c, (n, m, kind) in zip(coeffs, FRINGE_NM):
        if c == 0:
            continue
        r = _radial(n, m, rho)
        if kind == 1:
            r = r * np.cos(m * theta)
        elif kind == -1:
            r = r * np.sin(m * theta)
        w = w + c * r
    return w


def characteristic_matrix(f, g, NA, n):
    """Polarisation transfer matrix; restates
    CalculateCharacteristicMatrix.py:3-16 in complex128."""
    a = (NA / n) * np.asarray(f, dtype=np.complex128)
    b = (NA / n) * np.asarray(g, dtype=np.complex128)
    gam = np.sqrt(1 - a ** 2 - b ** 2)
    mxx = 1 - a ** 2 / (1 + gam)
    myx = -a * b / (1 + gam)
    myy = 1 - b ** 2 / (1 + gam)
    return mxx, myx, myx, myy, -a, -b


def _pupil(rho2, theta, proj, wavelength, focus):
    """Scalar pupil value at squared radius rho2: obliquity x exp(2 pi i W) x
    defocus (CalculateAerialImage.py:62-67,79; CalculateTCCMatrix.py:33-38).
    NaN where the obliquity ratio is negative (NA > 1 and M > 1), as in the
    reference."""
    NA, M, n = float(proj.NA), float(proj.Magnification), float(proj.IndexImage)
    rho = np.sqrt(rho2)
    with np.errstate(invalid="ignore", divide="ignore"):
        ob = np.power((1 - (NA / M) ** 2 * rho2) / (1 - NA ** 2 * rho2), 0.25)
        w = zernike_fringe(rho, theta, _coeffs(proj))
        defocus = np.exp(1j * 2 * np.pi / wavelength
                         * np.sqrt(n ** 2 - (NA / M) ** 2 * rho2 + 0j) * focus)
    return ob * np.exp(1j * 2 * np.pi * w) * defocus


def _coeffs(proj):
    z = proj.Aberration_Zernike
    if isinstance(z, torch.Tensor):
        z = z.detach().cpu().double().numpy()
    return np.asarray(z, dtype=np.float64)


def _foci(proj):
    fr = proj.FocusRange
    if isinstance(fr, torch.Tensor):
        return [float(v) for v in fr.reshape(-1).tolist()]
    return [float(v) for v in np.asarray(fr).reshape(-1)]


def _mode(numerics):
    mode = numerics.ImageCalculationMode
    if mode not in ("scalar", "vector"):
        raise ValueError("ImageCalculationMode must be 'scalar' or 'vector'")
    return mode


# ---------------------------------------------------------------------------
# Spectrum (Scatter.py:10-36)
# ---------------------------------------------------------------------------

def _field(scatter):
    h = scatter.ScatterField
    if isinstance(h, torch.Tensor):
        h = h.detach().cpu().numpy()
    return np.asarray(h).astype(np.complex128)


def spectrum(field, period_x, period_y, NA, wavelength):
    """Returns (Hk, f, g, PupilImage); restates Scatter.CalculateSpectrum
    (Scatter.py:10-36) in float64: Hk = px*py*fftshift(ifft2(ifftshift(h)))."""
    h = np.asarray(field).astype(np.complex128)
    ny, nx, _ = h.shape
    px = period_x / (wavelength / NA)
    py = period_y / (wavelength / NA)
    f = (1 / px) * np.arange(-(nx // 2), (nx + 1) // 2)
    g = (1 / py) * np.arange(-(ny // 2), (ny + 1) // 2)
    hk = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(h, axes=(0, 1)), axes=(0, 1)),
                         axes=(0, 1)) * (px * py)
    pupil = hk.copy()
    gm, fm = np.meshgrid(g, f, indexing="ij")
    pupil[(gm ** 2 + fm ** 2) > 1, :] = np.nan
    ox = int(px) + 3
    oy = int(py) + 3
    pupil = pupil[(ny // 2) - oy:(ny // 2) + oy, (nx // 2) - ox:(nx // 2) + ox]
    return hk, f, g, pupil


# ---------------------------------------------------------------------------
# Abbe (CalculateAerialImage.py:18-103)
# ---------------------------------------------------------------------------

def _abbe_setup(src, scatter, proj, numerics):
    mode = _mode(numerics)
    h = _field(scatter)
    X, Y, V = source_points(src)
    lam = float(src.Wavelength)
    NA, n = float(proj.NA), float(proj.IndexImage)
    hk, f, g, pupil_img = spectrum(h, scatter.Period_X, scatter.Period_Y, NA, lam)
    fm, gm = np.meshgrid(g, f, indexing="ij")            # CalculateAerialImage.py:31
    dfmdg = (f[1] - f[0]) * (g[1] - g[0])                 # :38
    rs, ts = np.hypot(X, Y), np.arctan2(Y, X)
    PX, PY = polarization(src, ts, rs)
    live = [c for c in range(h.shape[2]) if np.any(h[:, :, c] != 0)]   # zero channels add 0
    return dict(mode=mode, h=h, X=X, Y=Y, V=V, lam=lam, NA=NA, n=n, hk=hk, fm=fm, gm=gm,
                dfmdg=dfmdg, PX=PX, PY=PY, live=live, pupil_img=pupil_img)


def _abbe_source_term(c, j, z):
    """sum over field channels and polarisation components of |fft2(S)|^2 for
    source point j at focus z, unweighted and unshifted
    (CalculateAerialImage.py:55-87)."""
    ny, nx = c["h"].shape[:2]
    fc = c["fm"] + c["X"][j]
    gc = c["gm"] + c["Y"][j]
    rho2 = fc ** 2 + gc ** 2
    valid = rho2 <= 1                                      # :55-57
    fv, gv = fc[valid], gc[valid]
    r2 = rho2[valid]
    th = np.arctan2(gv, fv)
    pv = _pupil(r2, th, c["proj"], c["lam"], z)
    if c["mode"] == "vector":                              # :68-74
        mxx, myx, mxy, myy, mxz, myz = characteristic_matrix(fv, gv, c["NA"], c["n"])
        PX, PY = c["PX"][j], c["PY"][j]
        obl = [PX * mxx + PY * myx, PX * mxy + PY * myy, PX * mxz + PY * myz]
    else:
        obl = [1.0]
    accj = np.zeros((ny, nx))
    for ch in c["live"]:                                   # :76-87
        base = c["hk"][:, :, ch][valid] * pv
        for o in obl:
            S = np.zeros((ny, nx), dtype=np.complex128)
            S[valid] = base * o
            accj += np.abs(np.fft.fft2(S)) ** 2
    return accj


def abbe_image(src, scatter, proj, numerics):
    """Partially coherent Abbe image, float64.  Returns
    (Intensity (nFocus, Ny, Nx), ImageX, ImageY, ImageZ, PupilImage).

    Validity of a pupil sample is decided as rho^2 <= 1 in float64
    (the reference decides it in float32 at CalculateAerialImage.py:55-57)."""
    c = _abbe_setup(src, scatter, proj, numerics)
    c["proj"] = proj
    ny, nx = c["h"].shape[:2]
    V = c["V"]
    W = V.sum()
    foci = _foci(proj)
    out = np.zeros((len(foci), ny, nx))
    for fi, z in enumerate(foci):
        acc = np.zeros((ny, nx))
        for j in range(len(V)):
            acc += V[j] * _abbe_source_term(c, j, z)       # :89
        out[fi] = c["n"] / W * c["dfmdg"] ** 2 * np.fft.fftshift(acc)   # :90-92
    image_x = np.linspace(-scatter.Period_X / 2, scatter.Period_X / 2, nx)
    image_y = np.linspace(-scatter.Period_Y / 2, scatter.Period_Y / 2, ny)
    return out, image_x, image_y, np.array(foci), c["pupil_img"]


def abbe_source_terms(src, scatter, proj, numerics):
    """Per-source unweighted intensities J (nFocus, Nsrc, Ny, Nx) with the
    Abbe image I = (n / sum V) sum_s V_s J_s (CalculateAerialImage.py:46-92),
    and (V, n): the image as an explicit function of the source weights, for
    source-weight gradients.  Small grids only (Nsrc images per focus)."""
    c = _abbe_setup(src, scatter, proj, numerics)
    c["proj"] = proj
    ny, nx = c["h"].shape[:2]
    foci = _foci(proj)
    J = np.zeros((len(foci), len(c["V"]), ny, nx))
    for fi, z in enumerate(foci):
        for j in range(len(c["V"])):
            J[fi, j] = c["dfmdg"] ** 2 * np.fft.fftshift(_abbe_source_term(c, j, z))
    return J, c["V"], c["n"]


def clear_field_intensity(src, proj, numerics=None):
    """Closed form of the scalar, aberration-free clear-field image, a
    constant: (n/W) sum_j V_j |P(s_j)|^2 with |P(s)|^2 the squared obliquity
    factor at rho = |s| (only the DC order passes; CalculateAerialImage.py:64)."""
    X, Y, V = source_points(src)
    NA, M, n = float(proj.NA), float(proj.Magnification), float(proj.IndexImage)
    r2 = X ** 2 + Y ** 2
    ob2 = np.sqrt((1 - (NA / M) ** 2 * r2) / (1 - NA ** 2 * r2))
    return n / V.sum() * np.sum(V * ob2)


# ---------------------------------------------------------------------------
# TCC (CalculateTCCMatrix.py)
# ---------------------------------------------------------------------------

def tcc_grid(xypitch, NA, wavelength, halfwidth=None):
    """Pupil-frequency grid of the TCC (CalculateTCCMatrix.py:12-22).

    Grid half-widths Nf = ceil(2 p) are decided in float32 as the reference
    does (torch.tensor of a Python float, :17-20); the frequency values use the
    float64 pitches.  ``halfwidth`` overrides Nf and Ng (a zero-padded grid,
    e.g. 50 for a 101 x 101 grid)."""
    px = xypitch[0] / (wavelength / NA)
    py = xypitch[1] / (wavelength / NA)
    if halfwidth is None:
        nf = int(math.ceil(np.float32(2) * np.float32(px)))
        ng = int(math.ceil(np.float32(2) * np.float32(py)))
    else:
        nf = ng = int(halfwidth)
    f = (1 / px) * np.arange(-nf, nf + 1)
    g = (1 / py) * np.arange(-ng, ng + 1)
    return f, g


def shifted_pupil_matrix(src, xypitch, proj, focus, numerics, halfwidth=None):
    """Per-component shifted-pupil matrices G_e (M x Nsrc), the source weights
    V and the grid; restates CalculateTCCMatrix.py:9-51 (before GetTCCMatrix).
    Row m = a*len(g) + b holds frequency (f[a], g[b]) (meshgrid 'ij', :23)."""
    mode = _mode(numerics)
    lam = float(src.Wavelength)
    NA, n = float(proj.NA), float(proj.IndexImage)
    X, Y, V = source_points(src)
    f, g = tcc_grid(xypitch, NA, lam, halfwidth)
    ff, gg = np.meshgrid(f, g, indexing="ij")
    nf_ = ff.reshape(-1, 1) + X.reshape(1, -1)
    ng_ = gg.reshape(-1, 1) + Y.reshape(1, -1)
    rho2 = nf_ ** 2 + ng_ ** 2
    valid = rho2 <= 1                                       # :29
    th = np.arctan2(ng_, nf_)
    G = np.zeros(rho2.shape, dtype=np.complex128)
    G[valid] = _pupil(rho2[valid], th[valid], proj, lam, focus)   # :33-38
    if mode == "vector":                                    # :39-47
        mxx, myx, mxy, myy, mxz, myz = characteristic_matrix(nf_, ng_, NA, n)
        rs, ts = np.hypot(X, Y), np.arctan2(Y, X)
        PX, PY = polarization(src, ts, rs)
        Gs = [(PX * mxx + PY * myx) * G, (PX * mxy + PY * myy) * G,
              (PX * mxz + PY * myz) * G]
    else:
        Gs = [G]
    return Gs, V, f, g


def get_tcc_matrix(V, G):
    """(G diag(V) G^H) / sum(V); restates GetTCCMatrix (CalculateTCCMatrix.py:58-76)."""
    return (G * V[None, :]) @ G.conj().T / V.sum()


def tcc_matrix(src, xypitch, proj, focus, numerics, halfwidth=None):
    """Returns (TCC, [len(g), len(f)]); restates CalculateTCCMatrix
    (CalculateTCCMatrix.py:9-56)."""
    Gs, V, f, g = shifted_pupil_matrix(src, xypitch, proj, focus, numerics, halfwidth)
    tcc = sum(get_tcc_matrix(V, G) for G in Gs)
    return float(proj.IndexImage) * tcc, [len(g), len(f)]


# ---------------------------------------------------------------------------
# SOCS decomposition (DecomposeTCC_SOCS.py) and image (CalculateAerialImage_SOCS.py)
# ---------------------------------------------------------------------------

def eig_desc(tcc):
    """Singular values (descending) and left vectors of a Hermitian PSD
    matrix via float64 eigh; equals torch.svd's U, S (DecomposeTCC_SOCS.py:7)
    up to per-vector phase."""
    lam, vec = np.linalg.eigh(tcc)
    s = np.abs(lam)
    order = np.argsort(-s, kind="stable")
    return s[order], vec[:, order]


def socs_count(s, M, numerics):
    """Number of kernels (DecomposeTCC_SOCS.py:8-22).  'order':
    min(Hopkins_Order, M - 1).  'threshold': the reference loop raises a
    RuntimeError (it indexes a 2-D torch.diag(S)); its evident intent is kept:
    the count of leading values whose running sum, including the next value,
    first exceeds Hopkins_Threshold * sum(S)."""
    kind = str(numerics.Hopkins_SettingType).lower()
    if kind == "order":
        return int(min(numerics.Hopkins_Order, M - 1))
    if kind == "threshold":
        total = float(np.sum(s))
        acc = 0.0
        for i in range(len(s) - 1):
            acc += float(s[i])
            if acc > numerics.Hopkins_Threshold * total:
                return i
        return M - 1
    raise ValueError("Hopkins_SettingType should be order or threshold")


def embed_kernels(U, s, K, fg, nfg):
    """Embed the first K vectors (scaled by sqrt(s)) into (Ny, Nx, K) and
    fftshift; restates DecomposeTCC_SOCS.py:24-35."""
    nf, ng = nfg[0], nfg[1]
    lenf, leng = fg[1], fg[0]
    out = np.zeros((ng, nf, K), dtype=np.complex128)
    r0 = (ng - lenf + 1) // 2
    c0 = (nf - leng + 1) // 2
    tmp = np.zeros((ng, nf), dtype=np.complex128)
    for i in range(K):
        tmp[r0:(ng + lenf + 1) // 2, c0:(nf + leng + 1) // 2] = U[:, i].reshape(lenf, leng)
        out[:, :, i] = np.fft.fftshift(tmp)
    return out * np.sqrt(s[:K]).reshape(1, 1, -1)


def decompose_tcc_socs(tcc, fg, nfg, numerics):
    """Dense float64 route of DecomposeTCC_SOCS (DecomposeTCC_SOCS.py:3-37)."""
    s, U = eig_desc(tcc)
    K = socs_count(s, tcc.shape[0], numerics)
    return embed_kernels(U, s, K, fg, nfg)


def decompose_lowrank(Gs, V, n, fg, nfg, numerics):
    """Same kernels through the Gram matrix of B = [G_e sqrt(n V / sum V)]:
    TCC = B B^H, so B^H B = Q L Q^H gives TCC's non-zero eigenpairs as
    u_i sqrt(l_i) = B q_i.  Pinned against the dense route at small M."""
    w = np.sqrt(n * V / V.sum())
    B = np.concatenate([G * w[None, :] for G in Gs], axis=1)
    H = B.conj().T @ B
    lam, Q = np.linalg.eigh(H)
    lam = np.abs(lam)
    order = np.argsort(-lam, kind="stable")
    lam, Q = lam[order], Q[:, order]
    M = B.shape[0]
    s = np.zeros(M)
    r = min(len(lam), M)
    s[:r] = lam[:r]
    K = socs_count(s, M, numerics)
    Keff = min(K, r)
    U = np.zeros((M, K), dtype=np.complex128)
    nz = lam[:Keff] > 0
    U[:, :Keff][:, nz] = (B @ Q[:, :Keff][:, nz]) / np.sqrt(lam[:Keff][nz])
    return embed_kernels(U, s, K, fg, nfg)


def socs_image(field, kernels, period_x, period_y, NA, wavelength):
    """SOCS image, float64; restates CalculateAerialImage_SOCS
    (CalculateAerialImage_SOCS.py:3-11):
    fftshift(sum_k,c |df dg fft2(phi_k * ifftshift(Hk_c))|^2)."""
    hk, f, g, _ = spectrum(field, period_x, period_y, NA, wavelength)
    sp = np.fft.ifftshift(hk, axes=(0, 1))
    ny, nx, K = kernels.shape
    acc = np.zeros((ny, nx))
    scale = (f[1] - f[0]) * (g[1] - g[0])
    for c in range(hk.shape[2]):
        if not np.any(sp[:, :, c] != 0):
            continue
        for k in range(K):
            acc += np.abs(scale * np.fft.fft2(kernels[:, :, k] * sp[:, :, c])) ** 2
    return np.fft.fftshift(acc)


def hopkins_image(src, scatter, proj, numerics, halfwidth=None, lowrank=False):
    """Hopkins image per focus; restates CalculateHopkinsImage
    (CalculateAerialImage.py:105-130).  Returns (Intensity, kernels of the last
    focus)."""
    h = _field(scatter)
    ny, nx, _ = h.shape
    pitch = [scatter.Period_Y, scatter.Period_X]
    nfg = [nx, ny]
    foci = _foci(proj)
    out = np.zeros((len(foci), ny, nx))
    kern = None
    for fi, z in enumerate(foci):
        if lowrank:
            Gs, V, f, g = shifted_pupil_matrix(src, pitch, proj, z, numerics, halfwidth)
            kern = decompose_lowrank(Gs, V, float(proj.IndexImage), [len(g), len(f)], nfg, numerics)
        else:
            tcc, fg = tcc_matrix(src, pitch, proj, z, numerics, halfwidth)
            kern = decompose_tcc_socs(tcc, fg, nfg, numerics)
        out[fi] = socs_image(h, kern, scatter.Period_X, scatter.Period_Y,
                             float(proj.NA), float(src.Wavelength))
    return out, kern


def dft2_bruteforce(x, sign):
    """Direct 2-D DFT sum_n x[n] exp(sign 2 pi i k.n / N) over axes (0, 1)."""
    x = np.asarray(x, dtype=np.complex128)
    ny, nx = x.shape[:2]
    wy = np.exp(sign * 2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    wx = np.exp(sign * 2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    return np.einsum("ab,bc...->ac...", wy, np.einsum("cd,bd...->bc...", wx, x))


# ---------------------------------------------------------------------------
# Floquet-mode SOCS (CalculateAerialImage_SOCS.py:13-32, CalculateAerialImage.py:131-152)
# ---------------------------------------------------------------------------

def calculate_socs(floquet_mode, kernels, grid_x, grid_y, order_x, order_y):
    """Direct float64 sum of CalculateSOCS (CalculateAerialImage_SOCS.py:13-32):
    S[x, y, c, k] = sum_o exp(2 pi i (ox (ix + 1/2)/Gx + oy (iy + 1/2)/Gy))
    FM[o, c] phi[o, k]; I[x, y] = sum_{c,k} |S|^2.  Orders o run over the
    'ij' meshgrid of [-Ox, Ox] x [-Oy, Oy] flattened row-major; phi[o] is the
    kernel array transposed (dims 0, 1) and ifftshifted, flattened the same way.
    Returns (Gx, Gy)."""
    fm = np.asarray(floquet_mode).astype(np.complex128).reshape(-1, 3)
    K = kernels.shape[2]
    kt = np.fft.ifftshift(np.transpose(np.asarray(kernels).astype(np.complex128), (1, 0, 2)),
                          axes=(0, 1)).reshape(-1, K)
    ox = np.arange(-order_x, order_x + 1)
    oy = np.arange(-order_y, order_y + 1)
    gx, gy = np.meshgrid(ox, oy, indexing="ij")
    ox_d, oy_d = gx.reshape(-1), gy.reshape(-1)
    x = (np.arange(grid_x) + 0.5) / grid_x
    y = (np.arange(grid_y) + 0.5) / grid_y
    ex = np.exp(2j * np.pi * np.outer(x, ox_d))            # (Gx, O)
    ey = np.exp(2j * np.pi * np.outer(y, oy_d))            # (Gy, O)
    out = np.zeros((grid_x, grid_y))
    for c in range(3):
        for k in range(K):
            a = fm[:, c] * kt[:, k]                          # (O,)
            S = np.einsum("xo,yo,o->xy", ex, ey, a)
            out += S.real ** 2 + S.imag ** 2
    return out


def optimized_image(src, scatter, proj, numerics):
    """CalculateOptimized (CalculateAerialImage.py:131-152), float64: per focus
    TCC -> SOCS kernels on the Floquet grid -> calculate_socs.  Returns
    (nFocus, ScatterGrid_X, ScatterGrid_Y)."""
    fm = scatter.FloquetMode
    if isinstance(fm, torch.Tensor):
        fm = fm.detach().cpu().numpy()
    fm = np.asarray(fm)
    pitch = [scatter.Period_Y, scatter.Period_X]
    nfg = [fm.shape[0], fm.shape[1]]
    foci = _foci(proj)
    out = np.zeros((len(foci), numerics.ScatterGrid_X, numerics.ScatterGrid_Y))
    for fi, z in enumerate(foci):
        tcc, fg = tcc_matrix(src, pitch, proj, z, numerics)
        kern = decompose_tcc_socs(tcc, fg, nfg, numerics)
        out[fi] = calculate_socs(fm, kern, numerics.ScatterGrid_X, numerics.ScatterGrid_Y,
                                 numerics.ScatterOrder_X, numerics.ScatterOrder_Y)
    return out
