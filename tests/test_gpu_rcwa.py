"""RCWA near-field producer on the device (opticalimaging_amd.rcwa) against the
float64 oracle (oracle/rcwa_fp64.py, restating rcwa.py) and the reference's
own outputs (tests/golden/reference_rcwa.npz).  Both sides compute in
complex128; tolerances are relative to the largest entry."""
import math
import os
import time

import numpy as np
import pytest
import torch

import oracle as O
from oracle.rcwa_fp64 import RcwaOracle
from gpu_cases import rel_err, scatter

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                            "reference_rcwa.npz"))
EPS_HI = 2.25 + 0.01j
FILM = 4.5 + 0.2j


def _np(t):
    return t.detach().cpu().numpy()


def _golden_pair(fast=False, dtype=torch.complex128):
    from opticalimaging_amd.rcwa import rcwa
    d = rcwa(freq=1 / 193.0, order=[2, 2], L=[600.0, 400.0], dtype=dtype, fast=fast)
    o = RcwaOracle(1 / 193.0, [2, 2], [600.0, 400.0])
    for s in (d, o):
        s.add_input_layer(eps=1.0)
        s.set_incident_angle(math.radians(10), math.radians(30))
    d.add_layer(thickness=50.0, eps=torch.tensor(GOLD["geo"] * (EPS_HI - 1) + 1.0))
    d.add_layer(thickness=20.0, eps=FILM)
    o.add_layer(50.0, GOLD["geo"] * (EPS_HI - 1) + 1.0)
    o.add_layer(20.0, FILM)
    return d, o


def test_rcwa_matches_oracle_and_reference(gpu):
    d, o = _golden_pair()
    d.solve_global_smatrix()
    So = o.solve_global_smatrix()
    for i, name in enumerate(("S11", "S21", "S12", "S22")):
        e_o = rel_err(_np(d.S[i]), So[i])
        e_r = rel_err(_np(d.S[i]), GOLD["eig_" + name])
        print(f"global {name}: vs oracle {e_o:.2e}, vs reference {e_r:.2e}", flush=True)
        assert e_o < 1e-10 and e_r < 1e-10
    xs, ys = torch.tensor(GOLD["xs"]), torch.tensor(GOLD["ys"])
    d.source_planewave(amplitude=[0.8, 0.6], direction="forward")
    o.source_planewave((0.8, 0.6))
    fd = torch.stack(d.field_xy(xs, ys), 2)
    fo = np.stack(o.field_xy(GOLD["xs"], GOLD["ys"]), 2)
    qd = torch.stack(d.Floquet_mode(), 2)
    print(f"fields vs oracle {rel_err(_np(fd), fo):.2e}, vs reference "
          f"{rel_err(_np(fd), GOLD['eig_field']):.2e}; Floquet vs reference "
          f"{rel_err(_np(qd), GOLD['eig_floquet']):.2e}", flush=True)
    assert rel_err(_np(fd), fo) < 1e-10 and rel_err(_np(fd), GOLD["eig_field"]) < 1e-9
    assert rel_err(_np(qd), GOLD["eig_floquet"]) < 1e-9
    # complex64 requests: the same computation rounded at the end
    d32, _ = _golden_pair(dtype=torch.complex64)
    d32.solve_global_smatrix()
    assert d32.S[1].dtype == torch.complex64
    assert rel_err(_np(d32.S[1]), GOLD["eig_S21"]) < 1e-6


def test_rcwa_fast_flag_same_smatrix(gpu):
    d, _ = _golden_pair(fast=True)
    d.solve_global_smatrix()
    e_eig = rel_err(_np(d.S[1]), GOLD["eig_S21"])
    e_fast = rel_err(_np(d.S[1]), GOLD["fast_S21"])
    print(f"fast=True: vs reference eigen route {e_eig:.2e}, vs reference fast route {e_fast:.2e}")
    assert e_eig < 1e-10 and e_fast < 1e-5


def test_rcwa_energy_conservation_and_output_layer(gpu):
    from opticalimaging_amd.rcwa import rcwa
    geo = GOLD["geo"]
    d = rcwa(freq=1 / 193.0, order=[4, 2], L=[600.0, 400.0], dtype=torch.complex128)
    o = RcwaOracle(1 / 193.0, [4, 2], [600.0, 400.0])
    for s in (d, o):
        s.add_input_layer(eps=1.21)
        s.add_output_layer(eps=2.1)
        s.set_incident_angle(math.radians(17), math.radians(40))
    for s in (d, o):
        s.add_layer(50.0, torch.tensor(geo * 1.25 + 1.0) if s is d else geo * 1.25 + 1.0)
        s.add_layer(35.0, torch.tensor(np.roll(geo, 7, axis=1) * 2.1 + 1.0) if s is d
                    else np.roll(geo, 7, axis=1) * 2.1 + 1.0)
    d.solve_global_smatrix()
    So = o.solve_global_smatrix()
    o.S = [_np(s) for s in d.S]
    o.source_planewave((0.3, 0.9j))
    R, T = o.efficiencies()
    err = max(rel_err(_np(d.S[i]), So[i]) for i in range(4))
    print(f"input 1.21 / output 2.1: S vs oracle {err:.2e}; device R+T-1 {R.sum() + T.sum() - 1:.2e}")
    assert err < 1e-10 and abs(R.sum() + T.sum() - 1) < 1e-9


def test_rcwa_defect_and_repeated_block(gpu):
    from opticalimaging_amd.rcwa import rcwa
    d, o = _golden_pair()
    deps = GOLD["dgeo"] * (EPS_HI - 1)
    d.add_defect(layer_num=0, deps=torch.tensor(deps))
    d.solve_global_smatrix()
    exact = RcwaOracle(1 / 193.0, [2, 2], [600.0, 400.0])
    exact.add_input_layer(1.0)
    exact.set_incident_angle(math.radians(10), math.radians(30))
    exact.add_layer(50.0, GOLD["geo"] * (EPS_HI - 1) + 1.0 + deps)
    exact.add_layer(20.0, FILM)
    Se = exact.solve_global_smatrix()
    base = o.solve_global_smatrix()
    e_exact = rel_err(_np(d.S[1]), Se[1])
    change = rel_err(Se[1], base[1])
    e_pert = rel_err(GOLD["defect_S21"], Se[1])
    print(f"defect: device vs exact layer {e_exact:.2e}; reference perturbation vs exact "
          f"{e_pert:.2e} of a change of {change:.2e}")
    assert e_exact < 1e-10
    r = rcwa(freq=1 / 193.0, order=[2, 2], L=[600.0, 400.0], dtype=torch.complex128)
    r.add_input_layer(eps=1.0)
    r.set_incident_angle(inc_ang=0.0, azi_ang=0.0)
    r.add_layer(thickness=30.0, eps=torch.tensor(GOLD["geo"] * (EPS_HI - 1) + 1.0))
    r.add_layer(thickness=15.0, eps=FILM)
    r.add_layer(thickness=25.0, eps=torch.tensor(np.roll(GOLD["geo"], 5, axis=1) * (EPS_HI - 1) + 1.0))
    r.solve_global_smatrix_enhanced(1, 3, 4)
    r.source_planewave(amplitude=[0.8, 0.6], direction="forward")
    f = torch.stack(r.field_xy(torch.tensor(GOLD["xs"]), torch.tensor(GOLD["ys"])), 2)
    assert rel_err(_np(r.S[1]), GOLD["repeat_S21"]) < 1e-9
    assert rel_err(_np(f), GOLD["repeat_field"]) < 1e-9


def _pipeline_stack(mod, geometry, dtype=torch.complex128):
    """The stack of ImagingPipeline.Scattering (ImagingPipeline.py:55-98):
    365 nm, periods 3000 x 1500 nm, orders 5 x 5, oxide 1e5 nm / TiN 21 nm /
    TEOS 100 nm patterned, SiN 20 nm and substrate 5 nm homogeneous."""
    n_ox, n_tin, n_teos = 1.4745, 2.0669 + 1.0563j, 1.380
    n_sin, n_sub = 2.1222, 6.5271 + 2.6672j
    if mod == "device":
        from opticalimaging_amd.rcwa import rcwa
        s = rcwa(freq=1 / 365.0, order=[5, 5], L=[3000.0, 1500.0], dtype=dtype)
        s.add_input_layer(eps=1.0)
        s.set_incident_angle(inc_ang=0, azi_ang=0)
        for t, n in zip([1e5, 21, 100], [n_ox, n_tin, n_teos]):
            s.add_layer(thickness=t, eps=torch.tensor(geometry * (n ** 2 - 1) + 1.0))
        s.add_layer(thickness=20, eps=n_sin ** 2)
        s.add_layer(thickness=5, eps=n_sub ** 2)
        s.solve_global_smatrix()
        s.source_planewave(amplitude=[0.0, 1.0], direction="forward")
        return s
    s = RcwaOracle(1 / 365.0, [5, 5], [3000.0, 1500.0])
    s.add_input_layer(1.0)
    s.set_incident_angle(0.0, 0.0)
    for t, n in zip([1e5, 21, 100], [n_ox, n_tin, n_teos]):
        s.add_layer(t, geometry * (n ** 2 - 1) + 1.0)
    s.add_layer(20, n_sin ** 2)
    s.add_layer(5, n_sub ** 2)
    s.solve_global_smatrix()
    s.source_planewave((0.0, 1.0))
    return s


def test_rcwa_pipeline_stack_feeds_imaging(gpu):
    """The reference demo's stack through device RCWA, its near field as the
    ScatterField of CalculateAbbeImage, against the oracle end to end."""
    from opticalimaging_amd import CalculateAbbeImage, Numerics, Projection, Source
    geometry = np.zeros((200, 100))
    geometry[:, 25:50] = 1
    geometry[50:100, :] = 1
    torch.cuda.synchronize()
    t = time.perf_counter()
    d = _pipeline_stack("device", geometry)
    xs = (3000.0 / 256) * (torch.arange(256) + 0.5)
    ys = (1500.0 / 128) * (torch.arange(128) + 0.5)
    ex, ey, ez = d.field_xy(xs, ys)
    field = torch.stack((ex, ey, ez), 2)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    t = time.perf_counter()
    o = _pipeline_stack("oracle", geometry)
    fo = np.stack(o.field_xy(_np(xs), _np(ys)), 2)
    dto = time.perf_counter() - t
    e_field = rel_err(_np(field), fo)
    print(f"pipeline stack (5 layers, 121 orders): device {dt * 1e3:.1f} ms, float64 oracle "
          f"{dto * 1e3:.1f} ms; near field vs oracle {e_field:.2e}", flush=True)
    assert e_field < 1e-9
    src, proj, num = Source(), Projection(), Numerics()
    src.Wavelength = 365
    src.PolarizationVector = [0.0, 1.0]
    proj.Magnification, proj.NA = 100, 0.9
    proj.FocusRange = torch.tensor([0])
    num.ImageCalculationMode, num.ImageCalculationMethod = "scalar", "abbe"
    sc = scatter(field, 3000.0, 1500.0)
    img, _ = CalculateAbbeImage(src, sc, proj, num)
    ref, *_ = O.abbe_image(src, scatter(fo, 3000.0, 1500.0), proj, num)
    err = rel_err(_np(img.Intensity), ref)
    print(f"Abbe image of the RCWA near field vs float64 oracle chain: {err:.2e}")
    assert err < 1e-5
