"""Continuous-batching server on the GPU: every request's image is the same
alone and in a mixed-size stream (bit-exact: plans fix the kernel partition
per size class), matches the float64 oracle, and the HBM budget holds."""
import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import rel_err

pytestmark = pytest.mark.gpu


def test_mix_equals_alone_and_oracle(gpu):
    from opticalimaging_amd import serving, workloads
    src, proj, num = workloads.ilt_optics()
    classes = serving.size_classes((128, 256, 512), src, proj, num,
                                   max_batch={128: 8, 256: 8, 512: 4})
    t, sizes = workloads.serving_arrivals(60, rate=20000.0, sizes=(128, 256, 512), seed=3)
    masks = serving.pinned_masks(range(60), sizes, workloads.serving_mask)
    reqs = [serving.Request(i, masks[i], float(t[i])) for i in range(60)]
    ex = serving.GpuExecutor({c.n: c for c in classes}, streams=3)
    budget = ex.resident + 6 * 2 * 4 * 512 * 512          # a handful of clips in flight
    srv = serving.Server(classes, ex, hbm_budget=budget)
    done = srv.run(reqs)
    torch.cuda.synchronize()
    st = serving.latency_stats(done)
    print("mixed stream:", {k: round(v, 3) if isinstance(v, float) else v for k, v in st.items()},
          "peak in flight bytes", srv.peak_in_use)
    assert len(done) == 60 and srv.peak_in_use + srv.resident <= budget
    assert max(c.batch for c in done) > 1, "no batching happened"
    mixed = {c.rid: c.image.cpu().numpy() for c in done}
    by_n = {c.n: c for c in classes}
    worst = 0.0
    for i in range(60):
        alone = serving.Server(classes, ex, hbm_budget=budget).run([serving.Request(i, masks[i], 0.0)])
        a = alone[0].image.cpu().numpy()
        worst = max(worst, float(np.max(np.abs(a - mixed[i]))))
    print("max |alone - mixed| over 60 requests:", worst)
    assert worst == 0.0
    # oracle check of one request per size
    for n in (128, 256, 512):
        i = int(np.nonzero(sizes == n)[0][0])
        ks = by_n[n].kernels
        dense = ks.dense().cpu().numpy().astype(np.complex128)
        f3 = np.zeros((n, n, 3))
        f3[..., 0] = masks[i].numpy()
        ref = O.socs_image(f3, dense, 4.0 * n, 4.0 * n, proj.NA, src.Wavelength)
        err = rel_err(mixed[i], ref)
        print(f"served {n}^2 request {i} vs float64 SOCS: {err:.2e}")
        assert err < 1e-5


def test_mask_decode_kernel(gpu):
    """oi_mask_decode_u8: levels[codes] exactly, any length (16-code body and
    tail), codes past the table give 0, misaligned pointers are refused."""
    from opticalimaging_amd import engine
    rng = np.random.default_rng(1)
    levels = (0.25, -1.5, 3.0)
    for count in (1, 15, 16, 1000, 4096 * 16 + 7):
        codes = torch.as_tensor(rng.integers(0, 4, count).astype(np.uint8)).cuda()
        out = engine.decode_mask_codes(codes, levels)
        lut = torch.tensor(levels + (0.0,), dtype=torch.float32)
        assert torch.equal(out.cpu(), lut[codes.cpu().long()]), count
    buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError):
        engine.decode_mask_codes(buf[1:33], levels)


def test_coded_requests_equal_float_requests(gpu):
    """The same clips sent as byte codes and as float32 give bit-identical
    images (the decoded mask is the float32 mask exactly)."""
    from opticalimaging_amd import serving, workloads
    src, proj, num = workloads.ilt_optics()
    classes = serving.size_classes((128, 256), src, proj, num, max_batch={128: 4, 256: 4})
    t, sizes = workloads.serving_arrivals(24, rate=50000.0, sizes=(128, 256), seed=9)
    masks = serving.pinned_masks(range(24), sizes, workloads.serving_mask)
    codes = serving.pinned_codes(range(24), sizes, workloads.serving_mask_codes)
    ex = serving.GpuExecutor({c.n: c for c in classes}, streams=2)
    budget = ex.resident + 64 * 2**20
    a = serving.Server(classes, ex, hbm_budget=budget).run(
        [serving.Request(i, masks[i], float(t[i])) for i in range(24)])
    b = serving.Server(classes, ex, hbm_budget=budget).run(
        [serving.Request(i, codes[i][0], float(t[i]), levels=codes[i][1]) for i in range(24)])
    torch.cuda.synchronize()
    ia = {c.rid: c.image.cpu() for c in a}
    ib = {c.rid: c.image.cpu() for c in b}
    worst = max(float((ia[i] - ib[i]).abs().max()) for i in range(24))
    print("max |float - coded| over 24 requests:", worst)
    assert worst == 0.0


def test_mask_bits_decode_kernel(gpu):
    """oi_mask_decode_bits: levels[bit] exactly in numpy.packbits order, any
    pixel count (32-pixel body and tail); misaligned pointers are refused."""
    from opticalimaging_amd import engine
    rng = np.random.default_rng(2)
    levels = (0.125, -2.5)
    for count in (1, 7, 8, 31, 32, 33, 1000, 4096 * 32 * 256 + 13):
        b = rng.integers(0, 2, count).astype(np.uint8)
        packed = np.packbits(b)
        buf = torch.zeros(packed.size + 4, dtype=torch.uint8)
        buf[:packed.size] = torch.as_tensor(packed)
        out = engine.decode_mask_bits(buf.cuda(), levels, count)
        want = torch.tensor(levels, dtype=torch.float32)[torch.as_tensor(b).long()]
        assert torch.equal(out.cpu(), want), count
    buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError):
        engine.decode_mask_bits(buf[1:33], levels, 64)
    with pytest.raises(ValueError):
        engine.decode_mask_bits(buf, (0.0, 1.0, 2.0), 64)


def test_packed_requests_equal_float_requests(gpu):
    """The same clips sent one bit per pixel, as byte codes and as float32
    (mixed in one stream) give bit-identical images."""
    from opticalimaging_amd import serving, workloads
    src, proj, num = workloads.ilt_optics()
    classes = serving.size_classes((128, 256), src, proj, num, max_batch={128: 4, 256: 4})
    t, sizes = workloads.serving_arrivals(24, rate=50000.0, sizes=(128, 256), seed=11)
    masks = serving.pinned_masks(range(24), sizes, workloads.serving_mask)
    codes = serving.pinned_codes(range(24), sizes, workloads.serving_mask_codes)
    bits = serving.pinned_bits(range(24), sizes, workloads.serving_mask_codes)
    ex = serving.GpuExecutor({c.n: c for c in classes}, streams=2)
    budget = ex.resident + 64 * 2**20
    a = serving.Server(classes, ex, hbm_budget=budget).run(
        [serving.Request(i, masks[i], float(t[i])) for i in range(24)])

    def mixed(i):
        if i % 3 == 0:
            return serving.Request(i, bits[i][0], float(t[i]), levels=bits[i][1], packed=True)
        if i % 3 == 1:
            return serving.Request(i, codes[i][0], float(t[i]), levels=codes[i][1])
        return serving.Request(i, masks[i], float(t[i]))
    b = serving.Server(classes, ex, hbm_budget=budget).run([mixed(i) for i in range(24)])
    c = serving.Server(classes, ex, hbm_budget=budget).run(
        [serving.Request(i, bits[i][0], float(t[i]), levels=bits[i][1], packed=True)
         for i in range(24)])
    torch.cuda.synchronize()
    assert all(x.ok for x in a + b + c)
    ia = {x.rid: x.image.cpu() for x in a}
    for run in (b, c):
        ib = {x.rid: x.image.cpu() for x in run}
        worst = max(float((ia[i] - ib[i]).abs().max()) for i in range(24))
        print("max |float - packed or mixed| over 24 requests:", worst)
        assert worst == 0.0


def test_kernel_set_warm_start(gpu, tmp_path):
    """serving.size_classes with a cache directory: the first call builds and
    saves each size's kernel set, the second loads them (checkpoint) and the
    images through the loaded sets equal those through the built ones bit
    for bit; a changed optic misses the cache."""
    from opticalimaging_amd import checkpoint, engine, serving, workloads
    src, proj, num = workloads.ilt_optics()
    sizes = [256, 512]
    cold = serving.size_classes(sizes, src, proj, num, cache_dir=str(tmp_path))
    files = sorted(p.name for p in tmp_path.iterdir())
    assert len(files) == 2 and all(f.startswith("socs_") and f.endswith(".npz") for f in files)
    warm = serving.size_classes(sizes, src, proj, num, cache_dir=str(tmp_path))
    rng = np.random.default_rng(5)
    for c, w in zip(cold, warm):
        n = c.n
        assert torch.equal(c.kernels.win, w.kernels.win) and c.kernels.band == w.kernels.band
        m = torch.as_tensor((rng.random((2, n, n)) < 0.4).astype(np.float32)).cuda()
        assert torch.equal(engine.socs_image(m, c.kernels), engine.socs_image(m, w.kernels))
    ks, loaded = checkpoint.cached_kernel_set(str(tmp_path), src, [1024.0, 1024.0], proj, 0.0, num,
                                              [256, 256])
    assert loaded
    proj.NA = float(proj.NA) - 0.01
    ks2, loaded2 = checkpoint.cached_kernel_set(str(tmp_path), src, [1024.0, 1024.0], proj, 0.0, num,
                                                [256, 256])
    assert not loaded2 and len(list(tmp_path.iterdir())) == 3
