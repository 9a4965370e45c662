"""Scheduling logic of the continuous-batching server, with a host-only
executor whose batches take a simulated time (no GPU)."""
import numpy as np
import pytest

from opticalimaging_amd.serving import Request, Server, SizeClass, latency_stats


class SimExecutor:
    """Batches finish `base + per_image * b` simulated seconds after launch."""

    def __init__(self, clock, slots=2, base=1e-3, per_image=2e-4, resident=0):
        self.clock = clock
        self.slots = slots
        self.base = base
        self.per_image = per_image
        self.resident = resident
        self.launches = []

    def launch(self, cls, slot, reqs):
        now = self.clock[0]
        scale = (cls.n / 256) ** 2
        end = now + self.base + self.per_image * scale * len(reqs)
        self.launches.append((now, end, cls.n, slot, [r.rid for r in reqs]))
        return {"end": end, "imgs": [("img", r.rid) for r in reqs]}

    def done(self, h):
        return self.clock[0] >= h["end"]

    def outputs(self, h):
        return h["imgs"]


def _classes(max_batch=8):
    return [SizeClass(n, None, max_batch) for n in (256, 512, 1024)]


def _requests(count, rate, seed=0):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1 / rate, count))
    sizes = rng.choice([256, 512, 1024], count)
    return [Request(i, np.zeros((n, n), np.float32), float(a)) for i, (a, n) in enumerate(zip(t, sizes))]


def _drive(server, ex, reqs, dt=5e-5, limit=1e6):
    clock = ex.clock
    i = 0
    steps = 0
    inflight_max = 0
    while len(server.completed) < len(reqs):
        while i < len(reqs) and reqs[i].arrival <= clock[0]:
            server.waiting.append(reqs[i])
            i += 1
        server.step(clock[0])
        inflight = sum(len(b.reqs) for b in server.running)
        inflight_max = max(inflight_max, inflight)
        assert server.resident + server.in_use <= server.budget + 1e-9
        clock[0] += dt
        steps += 1
        assert steps < limit, "server made no progress"
    return inflight_max


def test_all_requests_complete_with_ordered_timestamps():
    clock = [0.0]
    ex = SimExecutor(clock)
    srv = Server(_classes(), ex, hbm_budget=1e9)
    reqs = _requests(300, rate=2000)
    _drive(srv, ex, reqs)
    done = srv.completed
    assert sorted(c.rid for c in done) == list(range(300))
    for c in done:
        assert c.arrival <= c.admitted <= c.launched <= c.done
    st = latency_stats(done)
    assert st["requests"] == 300 and st["p50_ms"] <= st["p99_ms"]


def test_batches_are_single_class_and_bounded():
    clock = [0.0]
    ex = SimExecutor(clock, slots=3)
    srv = Server(_classes(max_batch=5), ex, hbm_budget=1e9)
    reqs = _requests(200, rate=5000, seed=1)
    _drive(srv, ex, reqs)
    by_rid = {r.rid: r.n for r in reqs}
    for _, _, n, _, rids in ex.launches:
        assert 1 <= len(rids) <= 5
        assert all(by_rid[r] == n for r in rids)
    # a slot is never used by two batches at once
    for slot in range(3):
        spans = sorted((a, b) for a, b, _, s, _ in ex.launches if s == slot)
        for (a0, b0), (a1, b1) in zip(spans, spans[1:]):
            assert a1 >= b0


def test_hbm_budget_limits_admission():
    clock = [0.0]
    ex = SimExecutor(clock, slots=4)
    cls = [SizeClass(1024, None, 16)]
    per = cls[0].request_bytes
    srv = Server(cls, ex, hbm_budget=3 * per + 10)      # three requests fit
    reqs = [Request(i, np.zeros((1024, 1024), np.float32), 0.0) for i in range(20)]
    inflight_max = _drive(srv, ex, reqs)
    assert inflight_max <= 3
    assert srv.peak_in_use <= 3 * per
    assert len(srv.completed) == 20


def test_continuous_batching_overlaps_classes():
    """While a long batch of one class runs, requests of another class that
    arrive later are launched on a free slot before it finishes."""
    clock = [0.0]
    ex = SimExecutor(clock, slots=2, base=0.0, per_image=1e-3)
    srv = Server(_classes(max_batch=4), ex, hbm_budget=1e9)
    reqs = [Request(i, np.zeros((1024, 1024), np.float32), 0.0) for i in range(4)]
    reqs += [Request(4 + i, np.zeros((256, 256), np.float32), 1e-3 + 1e-4 * i) for i in range(3)]
    _drive(srv, ex, reqs, dt=1e-5)
    big = [l for l in ex.launches if l[2] == 1024][0]
    small = [l for l in ex.launches if l[2] == 256]
    assert small[0][0] < big[1], "small clips waited for the running batch"
    done = {c.rid: c for c in srv.completed}
    assert done[4].done < done[0].done


def test_oldest_class_first():
    clock = [0.0]
    ex = SimExecutor(clock, slots=1, base=1e-3, per_image=0.0)
    srv = Server(_classes(max_batch=2), ex, hbm_budget=1e9)
    reqs = [Request(0, np.zeros((512, 512), np.float32), 0.0),
            Request(1, np.zeros((256, 256), np.float32), 1e-4),
            Request(2, np.zeros((512, 512), np.float32), 2e-4),
            Request(3, np.zeros((256, 256), np.float32), 3e-4)]
    _drive(srv, ex, reqs, dt=1e-5)
    order = [l[2] for l in ex.launches]
    assert order[0] == 512 and order[1] == 256


def test_budget_too_small_raises():
    clock = [0.0]
    with pytest.raises(ValueError):
        Server([SizeClass(2048, None, 4)], SimExecutor(clock), hbm_budget=1e6)
    with pytest.raises(ValueError):
        Server([SizeClass(256, None, 4)], SimExecutor(clock, resident=2e9), hbm_budget=1e9)


def test_coded_requests_count_their_code_bytes():
    """A request whose mask arrives as byte codes holds one more byte per
    pixel on the device until it completes."""
    clock = [0.0]
    ex = SimExecutor(clock, slots=4)
    cls = SizeClass(1024, None, 16)
    coded = Request(0, np.zeros((1024, 1024), np.uint8), 0.0, levels=(1.0, -0.25))
    plain = Request(1, np.zeros((1024, 1024), np.float32), 0.0)
    assert cls.bytes_for(coded) == cls.request_bytes + 1024 * 1024
    assert cls.bytes_for(plain) == cls.request_bytes
    per = cls.bytes_for(coded)
    srv = Server([cls], ex, hbm_budget=2 * per + 10)        # two coded requests fit
    reqs = [Request(i, np.zeros((1024, 1024), np.uint8), 0.0, levels=(1.0, 0.0)) for i in range(9)]
    inflight_max = _drive(srv, ex, reqs)
    assert inflight_max <= 2 and srv.peak_in_use <= 2 * per
    assert srv.in_use == 0 and len(srv.completed) == 9
    # a coded clip that cannot fit even alone fails alone, with a reason
    tight = Server([cls], SimExecutor(clock), hbm_budget=cls.request_bytes + 10)
    done = tight.run([coded])
    assert len(done) == 1 and done[0].status == "error" and "budget" in done[0].error
    assert tight.in_use == 0


class FlakyExecutor(SimExecutor):
    """Raises on launch whenever a batch holds one of the `bad` request ids."""

    def __init__(self, clock, bad, **kw):
        super().__init__(clock, **kw)
        self.bad = set(bad)

    def launch(self, cls, slot, reqs):
        if any(r.rid in self.bad for r in reqs):
            raise RuntimeError("device rejected the batch")
        return super().launch(cls, slot, reqs)


def test_failures_stay_with_their_request():
    """Malformed requests, requests of an unknown size and requests whose
    launch fails complete with status "error"; every other request of the
    stream, including the failing ones' batchmates, is served."""
    clock = [0.0]
    ex = FlakyExecutor(clock, bad={7, 30}, slots=2)
    srv = Server(_classes(max_batch=4), ex, hbm_budget=1e9)
    reqs = _requests(60, rate=20000, seed=3)
    reqs[3] = Request(3, np.zeros((300, 300), np.float32), reqs[3].arrival)        # no class
    reqs[11] = Request(11, np.zeros((256, 512), np.float32), reqs[11].arrival)     # not square
    reqs[12] = Request(12, np.zeros((256, 256), np.float64), reqs[12].arrival)     # wrong dtype
    reqs[13] = Request(13, np.zeros((256, 256), np.uint8), reqs[13].arrival, levels=())
    _drive(srv, ex, reqs)
    by = {c.rid: c for c in srv.completed}
    assert sorted(by) == list(range(60))
    bad = {3, 7, 11, 12, 13, 30}
    for rid, c in by.items():
        if rid in bad:
            assert c.status == "error" and c.image is None and c.error, rid
        else:
            assert c.status == "ok" and c.image == ("img", rid), rid
    # the failing launches' batchmates were relaunched and served
    launched = [rid for *_, rids in ex.launches for rid in rids]
    assert 7 not in launched and 30 not in launched
    assert srv.in_use == 0 and not srv.running
    st = latency_stats(srv.completed)
    assert st["requests"] == 60 - len(bad)


def test_cancellation_before_and_after_launch():
    clock = [0.0]
    ex = SimExecutor(clock, slots=1, base=1e-3, per_image=0.0)
    cls = [SizeClass(256, None, 2)]
    srv = Server(cls, ex, hbm_budget=1e9)
    reqs = [Request(i, np.zeros((256, 256), np.float32), 0.0) for i in range(6)]
    for r in reqs:
        srv.waiting.append(r)
    srv.step(0.0)                           # admits all six, launches 0 and 1
    assert [e.req.rid for e in srv.running[0].reqs] == [0, 1]
    srv.cancel(1)                           # running: result discarded at completion
    srv.cancel(4)                           # pending: never launched, bytes released
    srv.cancel(99)                          # unknown: no effect
    in_use = srv.in_use
    srv.step(1e-4)
    assert srv.in_use == in_use - cls[0].request_bytes
    t = 1e-4
    while len(srv.completed) < 6:
        t += 2e-4
        clock[0] = t
        srv.step(t)
    by = {c.rid: c for c in srv.completed}
    assert by[1].status == "cancelled" and by[1].image is None
    assert by[4].status == "cancelled" and by[4].batch == 0
    assert all(by[i].status == "ok" for i in (0, 2, 3, 5))
    launched = [rid for *_, rids in ex.launches for rid in rids]
    assert 4 not in launched and srv.in_use == 0


def test_run_cancel_at():
    """run(cancel_at=...) withdraws a request that is still waiting."""
    clock = [0.0]
    ex = SimExecutor(clock, slots=1, base=5e-3, per_image=0.0)
    srv = Server([SizeClass(256, None, 1)], ex, hbm_budget=1e9)
    reqs = [Request(i, np.zeros((256, 256), np.float32), 0.0) for i in range(3)]

    def tick():
        clock[0] += 1e-3
        return clock[0]
    done = srv.run(reqs, clock=tick, idle_sleep=0.0, cancel_at={2: 0.0})
    by = {c.rid: c.status for c in done}
    assert by == {0: "ok", 1: "ok", 2: "cancelled"}


def test_packed_requests_are_validated_and_sized():
    """Bit-packed clips: (n, n / 8) uint8 with two levels; their device bytes
    count one bit per pixel of codes."""
    clock = [0.0]
    ex = SimExecutor(clock)
    classes = _classes()
    srv = Server(classes, ex, hbm_budget=1e9)
    good = Request(0, np.zeros((256, 32), np.uint8), 0.0, levels=(0.0, 1.0), packed=True)
    assert srv.problem(good) is None
    c = {k.n: k for k in classes}[256]
    assert c.bytes_for(good) == 2 * 4 * 256 * 256 + 256 * 256 // 8
    bad = [
        Request(1, np.zeros((256, 256), np.uint8), 0.0, levels=(0.0, 1.0), packed=True),   # unpacked shape
        Request(2, np.zeros((256, 32), np.uint8), 0.0, levels=(0.0, 0.5, 1.0), packed=True),  # three levels
        Request(3, np.zeros((256, 32), np.float32), 0.0, levels=(0.0, 1.0), packed=True),  # dtype
        Request(4, np.zeros((260, 32), np.uint8), 0.0, levels=(0.0, 1.0), packed=True),   # n not /8
        Request(5, np.zeros((256, 32), np.uint8), 0.0, packed=True),                       # no levels
        Request(6, np.zeros((96, 12), np.uint8), 0.0, levels=(0.0, 1.0), packed=True),    # no class
    ]
    for r in bad:
        assert srv.problem(r) is not None, r.rid
    # a mixed stream of packed, coded and float32 clips completes
    reqs = [Request(10, np.zeros((256, 32), np.uint8), 0.0, levels=(0.0, 1.0), packed=True),
            Request(11, np.zeros((512, 512), np.uint8), 0.0, levels=(0.0, 1.0)),
            Request(12, np.zeros((1024, 1024), np.float32), 0.0)]
    _drive(srv, ex, reqs)
    assert sorted(d.rid for d in srv.completed if d.ok) == [10, 11, 12]
