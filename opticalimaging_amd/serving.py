"""Continuous-batching aerial-image server over device-resident SOCS kernels.

Requests are mask clips of several sizes (one size class per grid size, each
with its own device-resident kernel set).  One host thread runs the loop:

  * arrivals join a FIFO; a request is *admitted* when its device bytes (mask
    in + image out, float32, plus one byte per pixel when the mask arrives as
    codes into a table of levels, or one bit per pixel when a two-level clip
    arrives packed) fit in what the HBM budget leaves after the resident
    kernel sets and plan scratch;
  * admitted requests wait per size class; whenever one of S execution slots
    (a HIP stream with its own plans) is free, the class holding the oldest
    admitted request launches a batch of up to its max_batch requests --
    requests that arrived while other batches were running join the next
    launch, so work of different classes overlaps on the device;
  * a batch's completion is detected with a HIP event query; its requests'
    bytes are released and their latency (arrival -> event complete, observed
    by the host) is recorded.

A request's image never depends on the batch it was computed in: plans fix
the kernel partitioning per size class, so alone and in the mix give the
same bits.

The device layer is an ``Executor``; ``GpuExecutor`` runs the library's
imaging path on HIP streams, and tests of the scheduling logic use a
host-only executor with simulated durations.
"""
from __future__ import annotations

import collections
import gc
import time
from dataclasses import dataclass, field

import numpy as np
import torch

__all__ = ["Request", "Completed", "SizeClass", "Server", "GpuExecutor", "latency_stats",
           "size_classes", "pinned_masks", "pinned_codes", "pinned_bits"]


@dataclass
class Request:
    rid: int
    mask: object            # (n, n) float32 host array (numpy or pinned torch), or uint8
                            # codes into `levels`, or (packed) the clip's bits
    arrival: float = 0.0    # seconds from the start of the run
    levels: tuple | None = None   # pixel value of each code when `mask` holds uint8 codes
    packed: bool = False    # mask is (n, n / 8) uint8, numpy.packbits of the 0/1 codes of
                            # each row, and levels holds the two pixel values

    @property
    def n(self) -> int:
        return int(self.mask.shape[0])

    @property
    def coded(self) -> bool:
        return self.levels is not None

    def code_bytes(self) -> int:
        """Bytes of the codes on the device (0 for a float32 mask)."""
        n = self.n
        return 0 if not self.coded else (n * n // 8 if self.packed else n * n)


@dataclass
class Completed:
    rid: int
    n: int
    image: object           # float32 (n, n) on the device (None unless status is "ok")
    arrival: float
    admitted: float
    launched: float
    done: float
    batch: int
    status: str = "ok"      # "ok" | "error" | "cancelled"
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    @property
    def ok(self) -> bool:
        return self.status == "ok"


@dataclass
class SizeClass:
    n: int
    kernels: object          # engine.KernelSet on the device
    max_batch: int

    @property
    def request_bytes(self) -> int:
        return 2 * 4 * self.n * self.n

    def bytes_for(self, req: "Request") -> int:
        """Device bytes of one request: float32 mask and image, plus the codes
        (a byte or a bit per pixel) when the mask arrives coded."""
        return self.request_bytes + req.code_bytes()


DEFAULT_MAX_BATCH = {128: 64, 256: 64, 512: 32, 1024: 16, 2048: 4}


def size_classes(sizes, source, projector, numerics, pixel_nm=4.0, focus=0.0, max_batch=None,
                 cache_dir=None):
    """One SizeClass per clip size n: kernels of the optics for an n x n clip
    of period n * pixel_nm (DecomposeTCC_SOCS.socs_kernel_set, resident on
    the device).  With cache_dir, each set is loaded from there when a file
    with the same optics fingerprint exists and saved there otherwise
    (checkpoint.cached_kernel_set: a warm start skips the TCC and its
    decomposition)."""
    from .DecomposeTCC_SOCS import socs_kernel_set
    from .checkpoint import cached_kernel_set
    out = []
    for n in sizes:
        period = n * pixel_nm
        if cache_dir is not None:
            ks, _ = cached_kernel_set(cache_dir, source, [period, period], projector, focus,
                                      numerics, [n, n])
        else:
            ks = socs_kernel_set(source, [period, period], projector, focus, numerics, [n, n])
        ks.for_real_masks()          # paired set for real masks, built before serving starts
        mb = (max_batch or {}).get(n, DEFAULT_MAX_BATCH.get(n, max(1, (1 << 22) // (n * n))))
        out.append(SizeClass(n, ks, mb))
    return out


def pinned_codes(rids, sizes, make):
    """Coded request masks in one pinned host allocation: make(rid, n) ->
    (uint8 codes (n, n), levels); returns a list of (pinned codes, levels)."""
    sizes = [int(n) for n in sizes]
    pool = torch.empty(sum(n * n for n in sizes), dtype=torch.uint8, pin_memory=True)
    out = []
    off = 0
    for rid, n in zip(rids, sizes):
        t = pool[off:off + n * n].view(n, n)
        codes, levels = make(rid, n)
        t.numpy()[...] = codes
        out.append((t, tuple(float(v) for v in levels)))
        off += n * n
    return out


def pinned_bits(rids, sizes, make):
    """Two-level request masks packed one bit per pixel in one pinned host
    allocation: make(rid, n) -> (0/1 uint8 codes (n, n), levels); returns a
    list of (pinned (n, n / 8) bits, levels), rows packed by numpy.packbits."""
    sizes = [int(n) for n in sizes]
    if any(n % 8 for n in sizes):
        raise ValueError("bit-packed clips need n divisible by 8")
    pool = torch.empty(sum(n * n // 8 for n in sizes), dtype=torch.uint8, pin_memory=True)
    out = []
    off = 0
    for rid, n in zip(rids, sizes):
        t = pool[off:off + n * n // 8].view(n, n // 8)
        codes, levels = make(rid, n)
        if len(levels) != 2:
            raise ValueError("bit-packed clips take two levels")
        t.numpy()[...] = np.packbits(np.asarray(codes, dtype=np.uint8), axis=1)
        out.append((t, tuple(float(v) for v in levels)))
        off += n * n // 8
    return out


def pinned_masks(rids, sizes, make):
    """Request masks generated straight into one pinned host allocation:
    returns a list of (n, n) float32 pinned tensors (make(rid, n) -> array)."""
    sizes = [int(n) for n in sizes]
    total = sum(n * n for n in sizes)
    pool = torch.empty(total, dtype=torch.float32, pin_memory=True)
    out = []
    off = 0
    for rid, n in zip(rids, sizes):
        t = pool[off:off + n * n].view(n, n)
        t.numpy()[...] = make(rid, n)
        out.append(t)
        off += n * n
    return out


class GpuExecutor:
    """Runs batches on `streams` HIP streams (default priority), one plan per
    (size class, stream).  Each (size class, stream) also owns its input
    buffers (float32 masks and byte codes for max_batch clips), allocated here
    once: a stream runs one batch at a time, so launches allocate only their
    output images."""

    def __init__(self, classes, streams: int = 4):
        from . import engine
        self.engine = engine
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.plans = {}
        self.inputs = {}
        self.bits = {}
        self.resident = 0
        for c in classes.values():
            ks = c.kernels
            self.resident += ks.win.numel() * 8
            if ks.for_real_masks() is not ks:
                self.resident += ks.for_real_masks().win.numel() * 8
            br, bc = ks.band if ks.band is not None else (0, 0)
            for si in range(streams):
                p = engine.Plan(ks.ny, ks.nx, ks.k0r, ks.wr, ks.k0c, ks.wc, nchan=1,
                                max_batch=c.max_batch, band_r=min(br, ks.wr),
                                band_c=min(bc, ks.wc), max_kernels=ks.K)
                self.plans[(c.n, si)] = p
                self.resident += p.bytes
                shape = (c.max_batch, c.n, c.n)
                self.inputs[(c.n, si)] = (torch.empty(shape, dtype=torch.float32, device="cuda"),
                                          torch.empty(shape, dtype=torch.uint8, device="cuda"))
                self.resident += 5 * c.max_batch * c.n * c.n
                if c.n % 8 == 0:
                    self.bits[(c.n, si)] = torch.empty((c.max_batch, c.n * c.n // 8),
                                                       dtype=torch.uint8, device="cuda")
                    self.resident += c.max_batch * c.n * c.n // 8

    @property
    def slots(self) -> int:
        return len(self.streams)

    def launch(self, cls: SizeClass, slot: int, reqs):
        s = self.streams[slot]
        b = len(reqs)
        n = cls.n
        srcs = [torch.as_tensor(r.mask) for r in reqs]
        keep = [srcs]
        fin, cin = self.inputs[(n, slot)]
        with torch.cuda.stream(s):
            dev_in = fin[:b]
            codes = cin[:b]
            bits = self.bits.get((n, slot))
            coded = [i for i, r in enumerate(reqs) if r.coded and not r.packed]
            packed = [i for i, r in enumerate(reqs) if r.packed]
            for i, m in enumerate(srcs):      # pinned host memory: asynchronous copies on s
                r = reqs[i]
                dst = bits[i].view(n, n // 8) if r.packed else (codes[i] if r.coded else dev_in[i])
                dst.copy_(m, non_blocking=m.is_pinned())
            if coded and len(coded) == b and all(r.levels == reqs[0].levels for r in reqs):
                self.engine.decode_mask_codes(codes, reqs[0].levels, out=dev_in)
            else:
                for i in coded:
                    self.engine.decode_mask_codes(codes[i], reqs[i].levels, out=dev_in[i])
            if packed and len(packed) == b and all(r.levels == reqs[0].levels for r in reqs):
                self.engine.decode_mask_bits(bits[:b].reshape(-1), reqs[0].levels, b * n * n,
                                             out=dev_in.view(-1))
            else:
                for i in packed:
                    self.engine.decode_mask_bits(bits[i], reqs[i].levels, n * n,
                                                 out=dev_in[i].view(-1))
            out = self.engine.socs_image(dev_in, cls.kernels, out=torch.empty_like(dev_in),
                                         plan=self.plans[(n, slot)])
            ev = torch.cuda.Event()
            ev.record(s)
        return {"event": ev, "out": out, "keep": keep}

    def done(self, handle) -> bool:
        return handle["event"].query()

    def outputs(self, handle):
        out = handle["out"]
        return [out[i] for i in range(out.shape[0])]


@dataclass
class _Batch:
    cls: SizeClass
    slot: int
    reqs: list
    launched: float
    handle: object


@dataclass
class _Entry:
    req: Request
    admitted: float = 0.0
    nbytes: int = 0


class Server:
    """Continuous batching under an HBM budget (bytes, including the resident
    kernel sets and plan scratch reported by the executor).

    Failures stay with their request: a request that cannot be served (no
    size class for its clip, a malformed mask or level table, more bytes
    than the budget can ever hold) completes at once with status "error";
    a batch whose launch raises is retried one request at a time, so only
    the requests that fail alone are reported as errors; a batch whose
    completion check raises fails its own requests only.  ``cancel(rid)``
    withdraws a request: before launch it never reaches the device and its
    bytes are released; once launched its result is discarded when the
    batch completes.  Cancelled and failed requests appear in ``completed``
    with their status and no image."""

    def __init__(self, classes, executor, hbm_budget: float = 8e9, keep_images: bool = True):
        self.classes = {c.n: c for c in classes} if not isinstance(classes, dict) else classes
        self.ex = executor
        self.budget = float(hbm_budget)
        self.resident = float(getattr(executor, "resident", 0))
        if self.resident >= self.budget:
            raise ValueError(f"resident kernels and plans ({self.resident:.3e} B) exceed the HBM budget")
        for c in self.classes.values():
            if c.request_bytes > self.budget - self.resident:
                raise ValueError(f"a {c.n}^2 request does not fit in the HBM budget")
        self.room = self.budget - self.resident
        self.in_use = 0.0
        self.peak_in_use = 0.0
        self.waiting = collections.deque()
        self.pending = {n: collections.deque() for n in self.classes}
        self.running = []
        self.free_slots = list(range(executor.slots))
        self.completed = []
        self.cancelled = set()       # rids withdrawn before they completed
        self.keep_images = keep_images      # False: Completed.image is None (timing runs)

    # -- per-request checks ---------------------------------------------------
    def problem(self, r: Request) -> str | None:
        """Why request r cannot be served, or None."""
        mask = r.mask
        shape = tuple(getattr(mask, "shape", ()))
        dt = str(getattr(mask, "dtype", ""))
        if r.packed:
            if r.levels is None or len(r.levels) != 2:
                return "a bit-packed mask takes a table of two levels"
            if len(shape) != 2 or shape[0] % 8 or shape[1] * 8 != shape[0]:
                return f"a bit-packed mask is (n, n / 8) with n divisible by 8, got shape {shape}"
            if "uint8" not in dt:
                return f"bit-packed masks are uint8, got {dt}"
        elif len(shape) != 2 or shape[0] != shape[1]:
            return f"mask must be a square (n, n) array, got shape {shape}"
        c = self.classes.get(int(shape[0]))
        if c is None:
            return f"no size class for {shape[0]}^2 requests"
        if r.coded and not r.packed:
            if "uint8" not in dt:
                return f"coded masks are uint8, got {dt}"
            if not 1 <= len(r.levels) <= 256:
                return "a level table holds 1 to 256 values"
        elif not r.coded and "float32" not in dt:
            return f"masks are float32 (or uint8 codes with levels), got {dt}"
        if c.bytes_for(r) > self.room:
            return f"request {r.rid} ({c.bytes_for(r)} B) does not fit in the HBM budget"
        return None

    def _fail(self, r: Request, now, why, admitted=None, launched=None, status="error"):
        self.completed.append(Completed(r.rid, int(getattr(r.mask, "shape", (0,))[0]), None,
                                        r.arrival, now if admitted is None else admitted,
                                        now if launched is None else launched, now, 0,
                                        status=status, error=why))

    def cancel(self, rid: int) -> None:
        """Withdraw request rid (no effect once it has completed)."""
        self.cancelled.add(rid)

    # -- one pass of the loop -------------------------------------------------
    def _complete(self, now):
        still = []
        for b in self.running:
            try:
                finished = self.ex.done(b.handle)
                outs = None
                if finished and self.keep_images:
                    outs = self.ex.outputs(b.handle)
            except Exception as exc:          # the batch's own requests fail, nothing else
                finished, outs = True, None
                err = f"{type(exc).__name__}: {exc}"
            else:
                err = None
            if not finished:
                still.append(b)
                continue
            outs = outs if outs is not None else [None] * len(b.reqs)
            for e, img in zip(b.reqs, outs):
                rid = e.req.rid
                if rid in self.cancelled:
                    status, img, why = "cancelled", None, "cancelled after launch"
                    self.cancelled.discard(rid)
                elif err is not None:
                    status, img, why = "error", None, err
                else:
                    status, why = "ok", None
                self.completed.append(Completed(rid, e.req.n, img, e.req.arrival, e.admitted,
                                                b.launched, now, len(b.reqs), status=status,
                                                error=why))
                self.in_use -= e.nbytes
            self.free_slots.append(b.slot)
        self.running = still

    def _drop_cancelled(self, now):
        if not self.cancelled:
            return
        keep = collections.deque()
        for r in self.waiting:
            if r.rid in self.cancelled:
                self.cancelled.discard(r.rid)
                self._fail(r, now, "cancelled before admission", status="cancelled")
            else:
                keep.append(r)
        self.waiting = keep
        for n, q in self.pending.items():
            keep = collections.deque()
            for e in q:
                if e.req.rid in self.cancelled:
                    self.cancelled.discard(e.req.rid)
                    self.in_use -= e.nbytes
                    self._fail(e.req, now, "cancelled before launch", admitted=e.admitted,
                               status="cancelled")
                else:
                    keep.append(e)
            self.pending[n] = keep

    def _admit(self, now):
        while self.waiting:
            r = self.waiting[0]
            why = self.problem(r)
            if why is not None:
                self.waiting.popleft()
                self._fail(r, now, why)
                continue
            nb = self.classes[r.n].bytes_for(r)
            if self.resident + self.in_use + nb > self.budget:
                break
            self.waiting.popleft()
            self.in_use += nb
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            self.pending[r.n].append(_Entry(r, now, nb))

    def _try_launch(self, c, slot, entries, now):
        """Launch `entries` on `slot`; on failure retry them one at a time and
        report those that fail alone.  Returns True when the slot was used."""
        try:
            handle = self.ex.launch(c, slot, [e.req for e in entries])
        except Exception as exc:
            if len(entries) == 1:
                e = entries[0]
                self.in_use -= e.nbytes
                self._fail(e.req, now, f"{type(exc).__name__}: {exc}", admitted=e.admitted)
                return False
            used = False
            for i, e in enumerate(entries):
                if used:                      # the slot is busy: the rest wait for the next pass
                    self.pending[c.n].extendleft(reversed(entries[i:]))
                    break
                used = self._try_launch(c, slot, [e], now)
            return used
        self.running.append(_Batch(c, slot, entries, now, handle))
        return True

    def _launch(self, now):
        while self.free_slots:
            heads = [(q[0].req.arrival, n) for n, q in self.pending.items() if q]
            if not heads:
                return
            _, n = min(heads)
            c = self.classes[n]
            q = self.pending[n]
            take = [q.popleft() for _ in range(min(c.max_batch, len(q)))]
            slot = self.free_slots.pop(0)
            if not self._try_launch(c, slot, take, now):
                self.free_slots.insert(0, slot)

    def step(self, now):
        self._complete(now)
        self._drop_cancelled(now)
        self._admit(now)
        self._launch(now)

    # -- driving a request stream ---------------------------------------------
    def run(self, requests, clock=None, idle_sleep: float = 2e-5, cancel_at=None):
        """Serve `requests` (sorted by arrival, seconds from the start) in real
        time; returns the Completed records in completion order.
        cancel_at: optional {rid: seconds} at which a request is withdrawn."""
        clock = clock or time.perf_counter
        cancels = sorted((t, rid) for rid, t in (cancel_at or {}).items())
        # the cyclic garbage collector is paused while serving: a full
        # collection over the run's records stalls the loop for milliseconds
        gc_on = gc.isenabled()
        gc.disable()
        try:
            t0 = clock()
            i = j = 0
            total = len(requests)
            while len(self.completed) < total:
                now = clock() - t0
                while i < total and requests[i].arrival <= now:
                    self.waiting.append(requests[i])
                    i += 1
                while j < len(cancels) and cancels[j][0] <= now:
                    self.cancel(cancels[j][1])
                    j += 1
                before = (len(self.completed), len(self.running))
                self.step(now)
                if (len(self.completed), len(self.running)) == before and idle_sleep:
                    time.sleep(idle_sleep)
        finally:
            if gc_on:
                gc.enable()
        return self.completed


def latency_stats(done):
    """p50/p99/mean/max latency (ms), images/s and mean batch over the
    requests that completed with status "ok"."""
    done = [c for c in done if getattr(c, "status", "ok") == "ok"]
    lat = np.array([c.latency for c in done]) * 1e3
    span = max(c.done for c in done) - min(c.arrival for c in done)
    return {"requests": len(done), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "max_ms": float(lat.max()), "images_per_s": len(done) / span if span > 0 else None,
            "mean_batch": float(np.mean([c.batch for c in done]))}
