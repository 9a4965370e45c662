"""Device-side engine: imaging plans, SOCS kernel sets and batched imaging.

A ``KernelSet`` is the device-resident, windowed form of SOCS kernels: the
c64 array [K][wr][wc] holding each kernel at spectrum index kappa =
(k0r + i, k0c + j), on a grid of ny x nx samples.  It is what
``DecomposeTCC_SOCS`` produces (in dense form, ``DecomposeTCC_SOCS.py:24-35``)
restricted to the kernels' support, so imaging never touches the dense
(ny, nx, K) array.

``Plan`` owns one library plan per geometry (cached by ``plan_for`` per
host thread and HIP stream); ``socs_image`` runs ``oi_socs_image`` on the
current stream.
"""
from __future__ import annotations

import collections
import ctypes
import threading
import weakref
from dataclasses import dataclass, field

import torch

from . import _native as nat


class _Attached:
    """A derived object attached to a tensor, valid while the tensor holds the
    storage and version it had when the object was attached."""
    __slots__ = ("obj", "version", "data_ptr")

    def __init__(self, obj, t: torch.Tensor):
        self.obj = obj
        self.version = t._version
        self.data_ptr = t.data_ptr()


def attach(t: torch.Tensor, name: str, obj) -> None:
    """Attach `obj` (derived from t's values) to tensor t under `name`."""
    setattr(t, name, _Attached(obj, t))


def attached(t, name: str):
    """The object attached to t under `name`, or None when there is none or t
    was modified in place (or re-pointed) after it was attached."""
    a = getattr(t, name, None)
    if not isinstance(a, _Attached) or not isinstance(t, torch.Tensor):
        return None
    if t._version != a.version or t.data_ptr() != a.data_ptr:
        return None
    return a.obj


def centred(idx: torch.Tensor, n: int) -> torch.Tensor:
    """Centred spectrum index of array positions (DC-at-0 layout)."""
    return torch.remainder(idx + n // 2, n) - n // 2


@dataclass
class KernelSet:
    """Windowed SOCS kernels on the device."""
    win: torch.Tensor      # c64 [K][wr][wc], cuda
    k0r: int
    k0c: int
    ny: int
    nx: int
    # (rows, cols) extent of one shifted pupil when the set holds every
    # eigenpair of its TCC: the image band is then the TCC band, 2 floor(2 p),
    # narrower than twice the kernels' support.  None otherwise.
    band: tuple | None = None
    # kernel set used for real masks (real_mask_kernels), built on first use
    _real: object = field(default=None, repr=False, compare=False)

    def for_real_masks(self) -> "KernelSet":
        """The kernel set that images real masks exactly like this one with
        the fewest kernels (real_mask_kernels; cached)."""
        if self._real is None:
            self._real = real_mask_kernels(self)
        return self._real

    @property
    def K(self) -> int:
        return int(self.win.shape[0])

    @property
    def wr(self) -> int:
        return int(self.win.shape[1])

    @property
    def wc(self) -> int:
        return int(self.win.shape[2])

    @staticmethod
    def from_dense(dense: torch.Tensor) -> "KernelSet":
        """Window of dense kernels (ny, nx, K) in the layout of
        DecomposeTCC_SOCS.py:31 (fftshifted embedding), found from their exact
        non-zero support."""
        dev = nat.device()
        lib = nat.load()
        d = dense.to(device=dev, dtype=torch.complex64).contiguous()
        if d.dim() != 3:
            raise ValueError("SOCS kernels must be a (Ny, Nx, K) tensor")
        ny, nx, K = d.shape
        rowflag = torch.zeros(ny, dtype=torch.int32, device=dev)
        colflag = torch.zeros(nx, dtype=torch.int32, device=dev)
        nat.check(lib.oi_kernel_support(nat.ptr(d), ny, nx, K, nat.ptr(rowflag),
                                        nat.ptr(colflag), nat.stream()))
        rows = torch.nonzero(rowflag).reshape(-1)
        cols = torch.nonzero(colflag).reshape(-1)
        if rows.numel() == 0:
            return KernelSet(torch.zeros((K, 1, 1), dtype=torch.complex64, device=dev), 0, 0, ny, nx)
        kr = centred(rows, ny)
        kc = centred(cols, nx)
        k0r, k1r = int(kr.min()), int(kr.max())
        k0c, k1c = int(kc.min()), int(kc.max())
        wr, wc = k1r - k0r + 1, k1c - k0c + 1
        win = torch.empty((K, wr, wc), dtype=torch.complex64, device=dev)
        nat.check(lib.oi_kernel_window(nat.ptr(d), ny, nx, K, k0r, wr, k0c, wc, nat.ptr(win),
                                       nat.stream()))
        return KernelSet(win, k0r, k0c, ny, nx)

    def dense(self) -> torch.Tensor:
        """Dense (ny, nx, K) c64 form (DecomposeTCC_SOCS.py:24-35 layout)."""
        dev = self.win.device
        out = torch.empty((self.ny, self.nx, self.K), dtype=torch.complex64, device=dev)
        nat.check(nat.load().oi_kernel_dense(nat.ptr(self.win), self.K, self.k0r, self.wr,
                                             self.k0c, self.wc, self.ny, self.nx, nat.ptr(out),
                                             nat.stream()))
        return out


class _DenseCache:
    """KernelSets extracted from dense kernel tensors, keyed on the tensor
    object itself (a weak reference, so a new tensor at a freed tensor's
    address never matches) and its version (in-place edits invalidate)."""

    def __init__(self, capacity=16):
        self.capacity = capacity
        self.items = collections.OrderedDict()
        self.lock = threading.Lock()

    def get(self, t: torch.Tensor):
        with self.lock:
            e = self.items.get(id(t))
            if e is None:
                return None
            ref, version, ptr, ks = e
            if ref() is not t or t._version != version or t.data_ptr() != ptr:
                del self.items[id(t)]
                return None
            self.items.move_to_end(id(t))
            return ks

    def put(self, t: torch.Tensor, ks) -> None:
        with self.lock:
            self.items[id(t)] = (weakref.ref(t), t._version, t.data_ptr(), ks)
            while len(self.items) > self.capacity:
                self.items.popitem(last=False)


_DENSE = _DenseCache()


def kernel_set_of(kernels) -> "KernelSet":
    """KernelSet for SOCS kernels given as a KernelSet, as DecomposeTCC_SOCS
    output (its attached set while unmodified), or as any dense (Ny, Nx, K)
    complex tensor (torch tensors: extracted once and cached while the same
    tensor object is unmodified; other arrays: extracted on every call)."""
    if isinstance(kernels, KernelSet):
        return kernels
    ks = attached(kernels, "_oi_kernels")
    if ks is not None:
        return ks
    if isinstance(kernels, torch.Tensor):
        ks = _DENSE.get(kernels)
        if ks is None:
            ks = KernelSet.from_dense(kernels)
            _DENSE.put(kernels, ks)
        return ks
    return KernelSet.from_dense(torch.as_tensor(kernels))


PAIR_MAX_K = 512        # larger sets are imaged as they are (the 2K x 2K eigenproblem grows)


def real_mask_kernels(ks: KernelSet, rel_tol: float = 1e-10) -> KernelSet:
    """Kernel set with at most ks.K kernels whose SOCS image of every real mask
    equals that of `ks` (CalculateAerialImage_SOCS.py:3-11).

    For a real mask the spectrum is Hermitian, S(-kappa) = conj S(kappa), so
    the antilinear map (J phi)(kappa) = conj phi(-kappa) sends a kernel's field
    to its conjugate, E_{J phi} = conj E_phi.  A kernel with J chi = chi
    ("J-real") therefore has a real field, and two of them share one complex
    kernel exactly: |E_{chi_a + i chi_b}|^2 = E_{chi_a}^2 + E_{chi_b}^2.

    Every kernel splits as phi = u + i w with u = (phi + J phi)/2 and
    w = (phi - J phi)/(2i) both J-real, and |E_phi|^2 = |E_u + i E_w|^2 =
    E_u^2 + E_w^2, so the image is sum_a E_{x_a}^2 over the 2K J-real vectors
    x = (u_1..u_K, w_1..w_K).  Their Gram matrix G = conj(X) X^T is real symmetric;
    with G = Q diag(l) Q^T the vectors chi_j = sum_a Q_aj x_a are J-real,
    orthogonal with |chi_j|^2 = l_j, and sum_j chi_j chi_j^H = sum_a x_a x_a^H,
    so the image is sum_j E_{chi_j}^2.  Keeping the r eigenpairs that hold all
    but rel_tol of trace(G) and pairing chi_{2j} + i chi_{2j+1} gives
    ceil(r/2) kernels.

    When the optics are symmetric under kappa -> -kappa with a pupil whose
    phase is odd (no defocus or even aberrations) the TCC commutes with J,
    r = K and the image costs half the kernels; otherwise r is usually 2K and
    `ks` itself is returned.  float64 throughout (eigenpairs from the
    library's Jacobi solver), result rounded to complex64 like the input."""
    K, wr, wc = ks.K, ks.wr, ks.wc
    if K < 2 or K > PAIR_MAX_K or 2 * ks.k0r + wr - 1 != 0 or 2 * ks.k0c + wc - 1 != 0:
        return ks        # J maps the window onto itself only when it is centred
    from .DecomposeTCC_SOCS import eigen_desc
    phi = ks.win.to(torch.complex128).reshape(K, -1)
    jphi = ks.win.flip(1, 2).conj().to(torch.complex128).reshape(K, -1)
    X = torch.cat([(phi + jphi) * 0.5, (phi - jphi) * (-0.5j)], 0)
    G = (X.conj() @ X.transpose(0, 1)).real
    G = 0.5 * (G + G.transpose(0, 1))
    lam, Q = eigen_desc(G.to(torch.complex128))
    if float(Q.imag.abs().max()) > 1e-9:
        return ks
    lam = lam.clamp_min(0)
    total = float(lam.sum())
    if total <= 0:
        return ks
    tail = torch.flip(torch.cumsum(torch.flip(lam, [0]), 0), [0])      # tail[j] = sum_{i >= j}
    keep = int((tail > rel_tol * total).sum())
    npk = (keep + 1) // 2
    if npk >= K:
        return ks
    chi = Q.real[:, :keep].transpose(0, 1).to(torch.complex128) @ X      # (keep, M), J-real rows
    if keep % 2:
        chi = torch.cat([chi, torch.zeros_like(chi[:1])], 0)
    psi = chi[0::2] + 1j * chi[1::2]
    win = psi.reshape(npk, wr, wc).to(torch.complex64).contiguous()
    out = KernelSet(win, ks.k0r, ks.k0c, ks.ny, ks.nx, band=ks.band)
    out._real = out
    return out


class Plan:
    """One oi_plan (twiddles + scratch) for a fixed geometry."""

    def __init__(self, ny, nx, k0r, wr, k0c, wc, nchan=1, max_batch=1, max_groups=0,
                 band_r=0, band_c=0, lr=0, lc=0, max_kernels=0):
        lib = nat.load()
        g = nat.Geometry(ny=ny, nx=nx, k0r=k0r, wr=wr, k0c=k0c, wc=wc, band_r=band_r,
                         band_c=band_c, nchan=nchan, max_batch=max_batch, max_groups=max_groups,
                         lr=lr, lc=lc, max_kernels=max_kernels)
        h = ctypes.c_void_p()
        nat.check(lib.oi_plan_create(ctypes.byref(g), ctypes.byref(h)))
        self.handle = h
        self.geo = g
        self.key = (ny, nx, k0r, wr, k0c, wc, nchan, max_batch, max_groups, band_r, band_c, lr, lc,
                    max_kernels)

    @property
    def max_batch(self) -> int:
        return int(self.geo.max_batch)

    @property
    def max_kernels(self) -> int:
        return int(self.geo.max_kernels)

    @property
    def bytes(self) -> int:
        return int(nat.load().oi_plan_bytes(self.handle))

    def info(self) -> dict:
        return {n: getattr(self.geo, n) for n, _ in nat.Geometry._fields_}

    STAGES = ("spectrum_rows", "spectrum_cols", "coherent", "interp_cols", "interp_rows")

    def set_timing(self, slots: int) -> None:
        """Record HIP events between the five launches of the next `slots`
        imaging calls (0 turns recording off)."""
        nat.check(nat.load().oi_plan_set_timing(self.handle, int(slots)))

    def read_timing(self):
        """(per-stage summed ms dict, number of recorded calls)."""
        ms = (ctypes.c_double * 5)()
        n = ctypes.c_int(0)
        nat.check(nat.load().oi_plan_read_timing(self.handle, ms, ctypes.byref(n)))
        return dict(zip(self.STAGES, list(ms))), int(n.value)

    def close(self):
        if self.handle:
            torch.cuda.synchronize()
            nat.load().oi_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            if self.handle and torch.cuda.is_available():
                self.close()
        except Exception:
            pass


class _PlanCache:
    """Plans by geometry, one set per (host thread, HIP stream): a plan's
    scratch is written by every call on it, so two streams (or threads) never
    share one.  An evicted plan is only dropped from the cache; it is
    destroyed when the last caller holding it lets it go."""

    def __init__(self, capacity=16):
        self.capacity = capacity
        self.plans = collections.OrderedDict()
        self.lock = threading.Lock()

    def get(self, **kw) -> Plan:
        owner = (threading.get_ident(), torch.cuda.current_stream().cuda_stream)
        key = (owner, tuple(sorted(kw.items())))
        with self.lock:
            p = self.plans.get(key)
            if p is not None:
                self.plans.move_to_end(key)
                return p
        p = Plan(**kw)
        with self.lock:
            self.plans[key] = p
            while len(self.plans) > self.capacity:
                self.plans.popitem(last=False)
        return p


PLANS = _PlanCache()


class FieldEncoding(tuple):
    """(tensor, kind, cin, chans) as the C-ABI takes them; ``source_channels``
    are the ScatterField channels that chans[i] stands for (they differ when
    one real channel is carried as a mask)."""
    source_channels: list


def _encoding(parts, source_channels) -> FieldEncoding:
    e = FieldEncoding(parts)
    e.source_channels = list(source_channels)
    return e


def field_encoding(field) -> FieldEncoding:
    """Choose the cheapest device encoding of a ScatterField (Ny, Nx, C), or a
    batch of them (B, Ny, Nx, C): returns (tensor, kind, cin, live channels).
    Channels that are identically zero contribute nothing to any image
    (CalculateAerialImage.py:76-87, CalculateAerialImage_SOCS.py:5-9) and are
    skipped; a single live channel with zero imaginary part is carried as a
    float32 mask [B, Ny, Nx].  The channel flags are attached to a torch
    tensor `field`, so a repeated call on the unchanged tensor reads them
    without the device-to-host transfer (an in-place edit drops them)."""
    dev = nat.device()
    user = field if isinstance(field, torch.Tensor) else None
    f = torch.as_tensor(field).to(device=dev)
    if f.dim() not in (3, 4):
        raise ValueError("ScatterField must be (Ny, Nx, C) or (B, Ny, Nx, C)")
    if f.dim() == 3:
        f = f.unsqueeze(0)
    cin = int(f.shape[-1])
    cplx = torch.is_complex(f)
    g = f.to(torch.complex64 if cplx else torch.float32).contiguous()
    npix = g.numel() // cin
    # which channels are non-zero, and which carry an imaginary part: one pass
    # over the field (oi_field_flags) and one device-to-host transfer
    flags = attached(user, "_oi_field_flags") if user is not None else None
    if flags is not None:
        nonzero, imag = flags
    elif cin <= 16:
        word = torch.zeros(1, dtype=torch.int32, device=dev)
        nat.check(nat.load().oi_field_flags(nat.ptr(g), nat.OI_FIELD_C64_CHLAST if cplx else nat.OI_FIELD_REAL,
                                            npix, cin, nat.ptr(word), nat.stream()))
        bits = int(word.item()) & 0xFFFFFFFF
        nonzero = [bool(bits >> (2 * c) & 1) for c in range(cin)]
        imag = [bool(bits >> (2 * c + 1) & 1) for c in range(cin)]
    else:
        flat = g.reshape(-1, cin)
        stats = torch.stack([(flat != 0).any(0), (flat.imag != 0).any(0) if cplx
                             else torch.zeros_like(flat[0], dtype=torch.bool)]).tolist()
        nonzero, imag = stats
    if user is not None and flags is None:
        attach(user, "_oi_field_flags", (tuple(nonzero), tuple(imag)))
    live = [c for c in range(cin) if nonzero[c]]
    if not live:
        live = [0]
    if len(live) == 1 and not imag[live[0]]:
        h = g[..., live[0]]
        h = h.real if cplx else h
        return _encoding((h.contiguous(), nat.OI_FIELD_REAL, 1, [0]), live)
    if not cplx:
        g = g.to(torch.complex64)
    return _encoding((g, nat.OI_FIELD_C64_CHLAST, cin, live), live)


def plan_for(ks: KernelSet, nchan: int = 1, max_batch: int = 1) -> Plan:
    """Cached plan for imaging with kernel set `ks` (or any set on its window
    with at most ks.K kernels, such as ks.for_real_masks()), for the calling
    thread and the current stream."""
    br, bc = ks.band if ks.band is not None else (0, 0)
    return PLANS.get(ny=ks.ny, nx=ks.nx, k0r=ks.k0r, wr=ks.wr, k0c=ks.k0c, wc=ks.wc, nchan=nchan,
                     max_batch=max_batch, band_r=min(br, ks.wr), band_c=min(bc, ks.wc),
                     max_kernels=ks.K)


def _device_tensor(name: str, t, dtype, shape) -> torch.Tensor:
    """t itself when it is a contiguous `dtype` tensor of `shape` on the
    current HIP device; ValueError otherwise (the library reads raw pointers)."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor on the HIP device")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the HIP device, got {t.device}")
    if t.device.index != torch.cuda.current_device():
        raise ValueError(f"{name} is on {t.device}, the current HIP device is "
                         f"cuda:{torch.cuda.current_device()}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def _check_plan(plan: Plan, ks: KernelSet, nchan: int, B: int) -> None:
    g = plan.geo
    if (g.ny, g.nx, g.k0r, g.wr, g.k0c, g.wc) != (ks.ny, ks.nx, ks.k0r, ks.wr, ks.k0c, ks.wc):
        raise ValueError("plan geometry does not match the kernel set's grid and window")
    if g.nchan != nchan:
        raise ValueError(f"plan carries {g.nchan} channels, the call {nchan}")
    if B > g.max_batch:
        raise ValueError(f"batch {B} exceeds the plan's max_batch {g.max_batch}")


def _real_set(ks: KernelSet, pack_real: bool) -> KernelSet:
    """The kernel set that images real masks: ks.for_real_masks() (the same
    image with up to half the kernels, a fixed function of ks) unless
    pack_real is False.  The choice never depends on earlier calls."""
    return ks.for_real_masks() if pack_real else ks


def socs_image(masks: torch.Tensor, ks: KernelSet, kind: int = nat.OI_FIELD_REAL, cin: int = 1,
               chans=(0,), out: torch.Tensor | None = None, max_batch: int | None = None,
               plan: Plan | None = None, pack_real: bool = True) -> torch.Tensor:
    """Batched SOCS intensity (CalculateAerialImage_SOCS.py:3-11 per mask).

    masks: float32 [B, ny, nx] (kind REAL) or c64 [B, ny, nx, cin] on the
    device, contiguous.  Returns float32 [B, ny, nx] in the reference's output
    layout (after its final fftshift); `out`, when given, must be exactly
    that.  Real masks are always imaged with ks.for_real_masks() (the same
    image with up to half the kernels; built on first use, then reused)
    unless pack_real is False, so a mask's image does not depend on which
    calls came before."""
    if not isinstance(masks, torch.Tensor) or masks.dim() not in (3, 4):
        raise ValueError("masks must be a (B, ny, nx) or (B, ny, nx, cin) device tensor")
    B = int(masks.shape[0])
    ny, nx = int(masks.shape[1]), int(masks.shape[2])
    if (ny, nx) != (ks.ny, ks.nx):
        raise ValueError(f"mask grid {ny}x{nx} does not match kernel grid {ks.ny}x{ks.nx}")
    if kind == nat.OI_FIELD_REAL:
        _device_tensor("masks", masks, torch.float32, (B, ny, nx))
    elif kind == nat.OI_FIELD_C64_CHLAST:
        _device_tensor("masks", masks, torch.complex64, (B, ny, nx, cin))
    else:
        raise ValueError("unknown field kind")
    if out is not None:
        _device_tensor("out", out, torch.float32, (B, ny, nx))
    if B == 0:
        return out if out is not None else torch.empty((0, ny, nx), dtype=torch.float32,
                                                       device=masks.device)
    if kind == nat.OI_FIELD_REAL:
        ks = _real_set(ks, pack_real)
    chans = list(chans)
    nchan = len(chans)
    if plan is None:
        plan = plan_for(ks, nchan, max(B, max_batch or 0))
    _check_plan(plan, ks, nchan, B)
    if out is None:
        out = torch.empty((B, ny, nx), dtype=torch.float32, device=masks.device)
    cptr, keep = nat.int_array(chans)
    nat.check(nat.load().oi_socs_image(plan.handle, nat.ptr(masks), kind, cin, cptr, B,
                                       nat.ptr(ks.win), ks.K, nat.ptr(out), nat.stream()))
    del keep
    return out


def socs_residual(masks: torch.Tensor, ks: KernelSet, target: torch.Tensor, out: torch.Tensor,
                  loss: torch.Tensor | None = None, plan: Plan | None = None,
                  pack_real: bool = True) -> bool:
    """out = 2 (I - target) for the SOCS image I of real masks [B, ny, nx], and
    loss[0] = sum (I - target)^2, fused into the pass that produces I
    (oi_socs_residual).  Returns False, with nothing computed, when the plan
    has no compile-time-size image pass; socs_image then does the work."""
    if not isinstance(masks, torch.Tensor) or masks.dim() != 3:
        raise ValueError("masks must be a (B, ny, nx) device tensor")
    B = int(masks.shape[0])
    shape = (B, ks.ny, ks.nx)
    if (int(masks.shape[1]), int(masks.shape[2])) != (ks.ny, ks.nx):
        raise ValueError(f"mask grid does not match kernel grid {ks.ny}x{ks.nx}")
    _device_tensor("masks", masks, torch.float32, shape)
    _device_tensor("target", target, torch.float32, shape)
    _device_tensor("out", out, torch.float32, shape)
    if loss is not None:
        if not (loss.is_cuda and loss.dtype == torch.float32 and loss.numel() >= 1
                and loss.is_contiguous()):
            raise ValueError("loss must be a contiguous float32 device tensor with one element")
    if B == 0:
        if loss is not None:
            loss.zero_()
        return True
    ks = _real_set(ks, pack_real)
    if plan is None:
        plan = plan_for(ks, 1, B)
    _check_plan(plan, ks, 1, B)
    rc = nat.load().oi_socs_residual(plan.handle, nat.ptr(masks), B, nat.ptr(ks.win), ks.K,
                                     nat.ptr(target), nat.ptr(out), nat.ptr(loss), nat.stream())
    if rc == nat.OI_ERR_UNSUPPORTED:
        return False
    nat.check(rc)
    return True


def socs_residual_adjoint(masks: torch.Tensor, ks: KernelSet, target: torch.Tensor,
                          resid: torch.Tensor, grad: torch.Tensor, work: torch.Tensor,
                          loss: torch.Tensor | None = None, plan: Plan | None = None) -> bool:
    """resid = 2 (I - target), loss[0] = sum (I - target)^2 and grad = the
    mask gradient of sum(resid * I), for real masks [B, ny, nx], sharing the
    mask spectrum (and on two-pass plans the column-stage fields) between the
    two passes (oi_socs_residual_adjoint: resid and loss equal socs_residual's
    bit for bit, grad equals socs_adjoint's to float32 rounding).  work: uint8 device tensor of at least
    oi_socs_adjoint_work_bytes(plan, B, K) bytes.  Returns False, with
    nothing computed, when the plan has no compile-time-size image pass."""
    if not isinstance(masks, torch.Tensor) or masks.dim() != 3:
        raise ValueError("masks must be a (B, ny, nx) device tensor")
    B = int(masks.shape[0])
    shape = (B, ks.ny, ks.nx)
    if (int(masks.shape[1]), int(masks.shape[2])) != (ks.ny, ks.nx):
        raise ValueError(f"mask grid does not match kernel grid {ks.ny}x{ks.nx}")
    for name, t in (("masks", masks), ("target", target), ("resid", resid), ("grad", grad)):
        _device_tensor(name, t, torch.float32, shape)
    if loss is not None and not (loss.is_cuda and loss.dtype == torch.float32 and loss.numel() >= 1
                                 and loss.is_contiguous()):
        raise ValueError("loss must be a contiguous float32 device tensor with one element")
    if B == 0:
        if loss is not None:
            loss.zero_()
        return True
    ks = _real_set(ks, True)
    if plan is None:
        plan = plan_for(ks, 1, B)
    _check_plan(plan, ks, 1, B)
    lib = nat.load()
    need = int(lib.oi_socs_adjoint_work_bytes(plan.handle, B, ks.K))
    if not (isinstance(work, torch.Tensor) and work.is_cuda and work.dtype == torch.uint8
            and work.is_contiguous() and work.numel() >= need):
        raise ValueError(f"work must be a contiguous uint8 device tensor of at least {need} bytes")
    rc = lib.oi_socs_residual_adjoint(plan.handle, nat.ptr(masks), B, nat.ptr(ks.win), ks.K,
                                      nat.ptr(target), nat.ptr(resid), nat.ptr(loss), nat.ptr(work),
                                      nat.ptr(grad), nat.stream())
    if rc == nat.OI_ERR_UNSUPPORTED:
        return False
    nat.check(rc)
    return True


def ilt_target_band(target: torch.Tensor, ks: KernelSet, work: torch.Tensor,
                    plan: Plan | None = None) -> torch.Tensor | None:
    """The targets' band spectrum and per-image constants for
    socs_band_residual_adjoint (oi_ilt_target_band), as a uint8 device
    tensor; None, with nothing computed, when the plan has no band form of
    the residual (direct plans, plans without the residual pass).  work: an
    adjoint_work buffer of the same plan and batch."""
    B = int(target.shape[0]) if isinstance(target, torch.Tensor) and target.dim() == 3 else -1
    if B < 1:
        raise ValueError("target must be a (B, ny, nx) device tensor with B >= 1")
    _device_tensor("target", target, torch.float32, (B, ks.ny, ks.nx))
    ks = _real_set(ks, True)
    if plan is None:
        plan = plan_for(ks, 1, B)
    _check_plan(plan, ks, 1, B)
    lib = nat.load()
    n = int(lib.oi_ilt_target_band_bytes(plan.handle, B))
    if n == 0:
        return None
    need = int(lib.oi_socs_adjoint_work_bytes(plan.handle, B, ks.K))
    if not (isinstance(work, torch.Tensor) and work.is_cuda and work.dtype == torch.uint8
            and work.is_contiguous() and work.numel() >= need):
        raise ValueError(f"work must be a contiguous uint8 device tensor of at least {need} bytes")
    tband = torch.empty(n, dtype=torch.uint8, device=target.device)
    nat.check(lib.oi_ilt_target_band(plan.handle, nat.ptr(target), B, nat.ptr(work), nat.ptr(tband),
                                     nat.stream()))
    return tband


def socs_band_residual_adjoint(masks: torch.Tensor, ks: KernelSet, tband: torch.Tensor,
                               grad: torch.Tensor, work: torch.Tensor,
                               loss: torch.Tensor | None = None, plan: Plan | None = None) -> None:
    """grad = the mask gradient of sum (I - T)^2 and loss[0] = that sum, for
    real masks [B, ny, nx] against the targets `tband` was made from
    (ilt_target_band, same plan and batch), from the image's band without
    forming the image (oi_socs_band_residual_adjoint; equal to
    socs_residual_adjoint's gradient and loss to float32 rounding)."""
    if not isinstance(masks, torch.Tensor) or masks.dim() != 3:
        raise ValueError("masks must be a (B, ny, nx) device tensor")
    B = int(masks.shape[0])
    shape = (B, ks.ny, ks.nx)
    for name, t in (("masks", masks), ("grad", grad)):
        _device_tensor(name, t, torch.float32, shape)
    if loss is not None and not (loss.is_cuda and loss.dtype == torch.float32 and loss.numel() >= 1
                                 and loss.is_contiguous()):
        raise ValueError("loss must be a contiguous float32 device tensor with one element")
    if B == 0:
        if loss is not None:
            loss.zero_()
        return
    ks = _real_set(ks, True)
    if plan is None:
        plan = plan_for(ks, 1, B)
    _check_plan(plan, ks, 1, B)
    lib = nat.load()
    nb = int(lib.oi_ilt_target_band_bytes(plan.handle, B))
    if nb == 0:
        raise ValueError("the plan has no band form of the residual")
    if not (isinstance(tband, torch.Tensor) and tband.is_cuda and tband.dtype == torch.uint8
            and tband.is_contiguous() and tband.numel() >= nb):
        raise ValueError(f"tband must be an ilt_target_band tensor of {nb} bytes")
    need = int(lib.oi_socs_adjoint_work_bytes(plan.handle, B, ks.K))
    if not (isinstance(work, torch.Tensor) and work.is_cuda and work.dtype == torch.uint8
            and work.is_contiguous() and work.numel() >= need):
        raise ValueError(f"work must be a contiguous uint8 device tensor of at least {need} bytes")
    nat.check(lib.oi_socs_band_residual_adjoint(plan.handle, nat.ptr(masks), B, nat.ptr(ks.win),
                                                ks.K, nat.ptr(tband), nat.ptr(loss), nat.ptr(work),
                                                nat.ptr(grad), nat.stream()))


def ilt_band_step(theta: torch.Tensor, ks: KernelSet, tband: torch.Tensor, lr: float,
                  grad: torch.Tensor, work: torch.Tensor, loss: torch.Tensor | None = None,
                  plan: Plan | None = None) -> None:
    """One normalised gradient step on the logits theta [B, ny, nx] (updated
    in place) against the targets `tband` was made from (oi_ilt_band_step):
    loss[0] = sum (I(sigmoid(theta)) - T)^2 before the step, grad = the
    gradient with respect to theta, theta -= lr grad / rms(grad) per image."""
    if not isinstance(theta, torch.Tensor) or theta.dim() != 3:
        raise ValueError("theta must be a (B, ny, nx) device tensor")
    B = int(theta.shape[0])
    shape = (B, ks.ny, ks.nx)
    for name, t in (("theta", theta), ("grad", grad)):
        _device_tensor(name, t, torch.float32, shape)
    if B > 0 and theta.data_ptr() == grad.data_ptr():
        raise ValueError("grad must not be theta itself")
    if loss is not None and not (loss.is_cuda and loss.dtype == torch.float32 and loss.numel() >= 1
                                 and loss.is_contiguous()):
        raise ValueError("loss must be a contiguous float32 device tensor with one element")
    if B == 0:
        if loss is not None:
            loss.zero_()
        return
    ks = _real_set(ks, True)
    if plan is None:
        plan = plan_for(ks, 1, B)
    _check_plan(plan, ks, 1, B)
    lib = nat.load()
    nb = int(lib.oi_ilt_target_band_bytes(plan.handle, B))
    if nb == 0:
        raise ValueError("the plan has no band form of the residual")
    if not (isinstance(tband, torch.Tensor) and tband.is_cuda and tband.dtype == torch.uint8
            and tband.is_contiguous() and tband.numel() >= nb):
        raise ValueError(f"tband must be an ilt_target_band tensor of {nb} bytes")
    need = int(lib.oi_socs_adjoint_work_bytes(plan.handle, B, ks.K))
    if not (isinstance(work, torch.Tensor) and work.is_cuda and work.dtype == torch.uint8
            and work.is_contiguous() and work.numel() >= need):
        raise ValueError(f"work must be a contiguous uint8 device tensor of at least {need} bytes")
    nat.check(lib.oi_ilt_band_step(plan.handle, nat.ptr(theta), B, nat.ptr(ks.win), ks.K,
                                   nat.ptr(tband), float(lr), nat.ptr(loss), nat.ptr(work),
                                   nat.ptr(grad), nat.stream()))


def adjoint_work(ks: KernelSet, batch: int, plan: Plan | None = None) -> torch.Tensor:
    """Device scratch for socs_adjoint / socs_residual_adjoint on `plan`."""
    ks = _real_set(ks, True)
    if plan is None:
        plan = plan_for(ks, 1, batch)
    n = int(nat.load().oi_socs_adjoint_work_bytes(plan.handle, batch, ks.K))
    return torch.empty(max(n, 1), dtype=torch.uint8, device=nat.device())


def socs_adjoint(masks: torch.Tensor, grad_image: torch.Tensor, ks: KernelSet,
                 out: torch.Tensor | None = None, plan: Plan | None = None,
                 pack_real: bool = True) -> torch.Tensor:
    """Mask gradient of the SOCS image: d/d masks of sum(grad_image * I(masks)).

    masks, grad_image: float32 [B, ny, nx] on the device (real masks).  The
    coherent fields are recomputed per kernel by the library (oi_socs_adjoint);
    no forward state is reused.  The image is the same function of real masks
    with ks.for_real_masks(), so its gradient is too; that set is used unless
    pack_real is False."""
    if not isinstance(masks, torch.Tensor) or masks.dim() != 3:
        raise ValueError("masks must be a (B, ny, nx) device tensor")
    B = int(masks.shape[0])
    ny, nx = int(masks.shape[1]), int(masks.shape[2])
    if (ny, nx) != (ks.ny, ks.nx):
        raise ValueError(f"mask grid {ny}x{nx} does not match kernel grid {ks.ny}x{ks.nx}")
    if masks.dtype != torch.float32 or not isinstance(grad_image, torch.Tensor) \
            or grad_image.dtype != torch.float32:
        raise ValueError("socs_adjoint takes float32 masks and image cotangents")
    if tuple(grad_image.shape) != (B, ny, nx):
        raise ValueError("grad_image must have the shape of masks")
    masks = _device_tensor("masks", masks.contiguous(), torch.float32, (B, ny, nx))
    grad_image = _device_tensor("grad_image", grad_image.contiguous(), torch.float32, (B, ny, nx))
    if out is not None:
        _device_tensor("out", out, torch.float32, (B, ny, nx))
    if B == 0:
        return out if out is not None else torch.empty((0, ny, nx), dtype=torch.float32,
                                                       device=masks.device)
    ks = _real_set(ks, pack_real)
    if plan is None:
        plan = plan_for(ks, 1, B)
    _check_plan(plan, ks, 1, B)
    lib = nat.load()
    nbytes = int(lib.oi_socs_adjoint_work_bytes(plan.handle, B, ks.K))
    work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=masks.device)
    if out is None:
        out = torch.empty((B, ny, nx), dtype=torch.float32, device=masks.device)
    nat.check(lib.oi_socs_adjoint(plan.handle, nat.ptr(masks), nat.ptr(grad_image), B,
                                  nat.ptr(ks.win), ks.K, nat.ptr(work), nat.ptr(out), nat.stream()))
    return out


def decode_mask_bits(bits: torch.Tensor, levels, n: int, out: torch.Tensor | None = None) -> torch.Tensor:
    """float32 masks of `n` pixels from one bit per pixel on the device
    (oi_mask_decode_bits): pixel i is levels[1] when bit 7 - i mod 8 of byte
    i / 8 is set, else levels[0] (numpy.packbits order).  `bits` is a uint8
    device tensor of at least ceil(n / 8) bytes."""
    if bits.dtype != torch.uint8 or not bits.is_cuda or not bits.is_contiguous():
        raise ValueError("mask bits must be a contiguous uint8 tensor on the device")
    if bits.numel() * 8 < n:
        raise ValueError(f"{bits.numel()} bytes hold fewer than {n} pixels")
    lv = [float(v) for v in levels]
    if len(lv) != 2:
        raise ValueError("bit-packed masks take exactly two levels")
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device=bits.device)
    if out.dtype != torch.float32 or out.numel() != n or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 tensor of n values")
    arr = (ctypes.c_float * 2)(*lv)
    nat.check(nat.load().oi_mask_decode_bits(nat.ptr(bits), n, arr, nat.ptr(out), nat.stream()))
    return out


def decode_mask_codes(codes: torch.Tensor, levels, out: torch.Tensor | None = None) -> torch.Tensor:
    """float32 masks levels[codes] from uint8 codes on the device
    (oi_mask_decode_u8): masks with at most 256 distinct values cross the host
    link as one byte per pixel.  `levels` is a sequence of floats."""
    if codes.dtype != torch.uint8 or not codes.is_cuda:
        raise ValueError("mask codes must be a uint8 tensor on the device")
    codes = codes.contiguous()
    if out is None:
        out = torch.empty(codes.shape, dtype=torch.float32, device=codes.device)
    if out.dtype != torch.float32 or out.numel() != codes.numel() or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 tensor with one value per code")
    lv = [float(v) for v in levels]
    arr = (ctypes.c_float * max(1, len(lv)))(*lv)
    nat.check(nat.load().oi_mask_decode_u8(nat.ptr(codes), codes.numel(), arr, len(lv), nat.ptr(out),
                                           nat.stream()))
    return out

