"""The vector Abbe and serving workloads at their stated sizes.

* Vector Abbe (CalculateAerialImage.py:18-103, vector branch 68-74): 512 x 512
  attenuated-PSM Manhattan clip over 2048 nm, NA 1.35 / n 1.44 immersion,
  41-point quasar 0.6/0.9/30 deg (196 lit points) with t_pol, 37 seeded
  Zernike terms, five foci -60..60 nm, through CalculateAbbeImage against the
  float64 Abbe oracle on every focus (<= 1e-5).
* Serving: clips of 256^2, 512^2, 1024^2 and 2048^2 (4 nm pixels) against the
  device-resident K = 24 kernel sets of the same optics at focus 0, admitted
  under an 8 GB HBM budget; every request's image is bit-identical alone and
  in the mix, and one request per size matches the float64 SOCS oracle
  (CalculateAerialImage_SOCS.py:3-11) with the same kernels (<= 1e-5).
"""
import numpy as np
import pytest
import torch

import oracle as O
from gpu_cases import rel_err, scatter

pytestmark = pytest.mark.gpu


def test_vector_abbe_512_five_foci(gpu):
    from opticalimaging_amd import CalculateAbbeImage, workloads
    mask, src, proj, num, period = workloads.vector_abbe_512()
    field = workloads.field_from_mask(mask)
    sc = scatter(field.cuda(), period, period)
    img, pupil = CalculateAbbeImage(src, sc, proj, num)
    got = img.Intensity.double().cpu().numpy()
    sc_ref = scatter(field.numpy(), period, period)
    ref, _, _, iz, _ = O.abbe_image(src, sc_ref, proj, num)
    assert got.shape == (5, 512, 512) and ref.shape == (5, 512, 512)
    assert src.source_data.Value.numel() == 196
    errs = [rel_err(got[i], ref[i]) for i in range(5)]
    print("vector Abbe 512^2, 196 points x 3 components, 5 foci: max rel err per focus",
          [f"{e:.2e}" for e in errs], flush=True)
    assert max(errs) < 1e-5
    # the focus sweep is live at M = 1: the five planes differ
    assert rel_err(ref[0], ref[2]) > 1e-3


def test_serving_all_sizes_8gb(gpu):
    from opticalimaging_amd import serving, workloads
    sizes_all = (256, 512, 1024, 2048)
    src, proj, num = workloads.ilt_optics()
    classes = serving.size_classes(sizes_all, src, proj, num,
                                   max_batch={256: 8, 512: 8, 1024: 4, 2048: 2})
    count = 32
    t, sizes = workloads.serving_arrivals(count, rate=40000.0, sizes=sizes_all, seed=23)
    for n in sizes_all:                      # every class present in the stream
        assert (sizes == n).any(), n
    codes = serving.pinned_codes(range(count), sizes, workloads.serving_mask_codes)
    reqs = [serving.Request(i, codes[i][0], float(t[i]), levels=codes[i][1]) for i in range(count)]
    ex = serving.GpuExecutor({c.n: c for c in classes}, streams=4)
    budget = 8e9
    print(f"resident kernels + plans: {ex.resident / 1e9:.2f} GB of the {budget / 1e9:.0f} GB budget",
          flush=True)
    srv = serving.Server(classes, ex, hbm_budget=budget)
    done = srv.run(reqs)
    torch.cuda.synchronize()
    st = serving.latency_stats(done)
    print("mixed stream:", {k: round(v, 3) if isinstance(v, float) else v for k, v in st.items()},
          flush=True)
    assert len(done) == count and srv.peak_in_use + srv.resident <= budget
    assert max(c.batch for c in done) > 1, "no batching happened"
    mixed = {c.rid: c.image.cpu().numpy() for c in done}
    worst = 0.0
    for i in range(count):
        alone = serving.Server(classes, ex, hbm_budget=budget).run(
            [serving.Request(i, codes[i][0], 0.0, levels=codes[i][1])])
        worst = max(worst, float(np.max(np.abs(alone[0].image.cpu().numpy() - mixed[i]))))
    print(f"max |alone - mixed| over {count} requests of 4 sizes: {worst}", flush=True)
    assert worst == 0.0
    by_n = {c.n: c for c in classes}
    for n in sizes_all:
        i = int(np.nonzero(sizes == n)[0][0])
        ks = by_n[n].kernels
        dense = ks.dense().cpu().numpy().astype(np.complex128)
        lv = np.asarray(codes[i][1], dtype=np.float64)
        f3 = np.zeros((n, n, 3))
        f3[..., 0] = lv[codes[i][0].numpy()]
        ref = O.socs_image(f3, dense, 4.0 * n, 4.0 * n, proj.NA, src.Wavelength)
        del dense
        err = rel_err(mixed[i], ref)
        print(f"served {n}^2 request {i} vs float64 SOCS (K = {ks.K}): {err:.2e}", flush=True)
        assert err < 1e-5


@pytest.mark.parametrize("mode,nfoci", [("vector", 5), ("scalar", 37)])
def test_multi_focus_pupils_equal_per_focus_pupils(gpu, mode, nfoci):
    """oi_pupil_kernels_foci writes, for every focus, the kernels that
    oi_pupil_kernels writes for that focus alone, bit for bit; 37 foci cross
    the 32 foci that one launch carries."""
    from opticalimaging_amd import optics, workloads
    from opticalimaging_amd.Scatter import _pitches
    mask, src, proj, num, period = workloads.vector_abbe_512(n=128)
    sc = scatter(workloads.field_from_mask(mask).cuda(), period, period)
    sset = optics.source_set(src)
    px, py = _pitches(sc, proj, src)
    k0r, wr, k0c, wc, _, _ = optics.abbe_window(sset, py, px, 128, 128)
    vector = mode == "vector"
    foci = [-75.0 + 150.0 * i / (nfoci - 1) for i in range(nfoci)]
    scale = torch.sqrt(float(proj.IndexImage) * sset.V / sset.weight)
    opt = optics.optics_record(src, proj, 0.0, py, px, vector)
    G = optics.pupil_kernels_foci(sset, opt, foci, scale, k0r, wr, k0c, wc)
    assert G.shape == (nfoci, 3 if vector else 1, sset.n, wr, wc)
    for i, z in enumerate(foci):
        one = optics.pupil_kernels(sset, optics.optics_record(src, proj, z, py, px, vector), scale,
                                   k0r, wr, k0c, wc)
        assert torch.equal(torch.view_as_real(G[i]), torch.view_as_real(one)), (mode, i, z)
    assert not torch.equal(G[0], G[-1])                 # the defocus phase is live
