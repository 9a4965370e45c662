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
