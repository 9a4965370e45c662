"""Writes golden vectors from the reference's RCWA (rcwa.py), run in the build
container where the reference sources are mounted read-only.

Purpose: pin what no closed form or conservation law fixes -- which block of
the global S-matrix field_xy and Floquet_mode read, their phase and Ez
conventions and reshape order (rcwa.py:269-338), the defect perturbation as
written (rcwa.py:592-693), and the repeated-block global S-matrix
(rcwa.py:196-223).  Energy conservation, the Airy slab and the agreement of
the eigen and matrix-exponential routes pin the rest (tests/test_oracle_rcwa.py).

Cases (complex128, CPU): a binary grating over a homogeneous film, orders
[2, 2], periods 600 x 400 nm, 193 nm, input layer eps 1.0, incidence 10 deg,
azimuth 30 deg; eigen route and `fast=True`; the same with a defect added to
the grating layer; and a three-layer stack whose middle pair repeats four
times.  The reference's add_output_layer cannot be combined with
set_incident_angle (rcwa.py:407-431 reads an undefined self.Vo), so no case
has an output layer.  Output: tests/golden/reference_rcwa.npz.
"""
import math
import os
import sys

import numpy as np
import torch

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_rcwa.npz")


def grating(n, m, period_frac=0.4, shift=0):
    g = np.zeros((n, m))
    g[:, shift:shift + int(period_frac * m)] = 1.0
    g[: n // 3, :] = np.maximum(g[: n // 3, :], 0.0)
    g[n // 4: n // 2, m // 2:] = 1.0
    return g


def main():
    sys.path.insert(0, REF)
    from rcwa import rcwa

    dt = torch.complex128
    dev = torch.device("cpu")
    geo = grating(32, 24)
    dgeo = np.zeros_like(geo)
    dgeo[20:22, 5:7] = 1.0
    xs = np.linspace(0.0, 600.0, 16, endpoint=False) + 7.0
    ys = np.linspace(0.0, 400.0, 12, endpoint=False) + 3.0
    out = {"geo": geo, "dgeo": dgeo, "xs": xs, "ys": ys}
    eps_hi = 2.25 + 0.01j
    film = 4.5 + 0.2j

    def build(fast):
        s = rcwa(freq=1 / 193.0, order=[2, 2], L=[600.0, 400.0], dtype=dt, device=dev, fast=fast)
        s.add_input_layer(eps=1.0)
        s.set_incident_angle(inc_ang=math.radians(10), azi_ang=math.radians(30))
        s.add_layer(thickness=50.0, eps=torch.tensor(geo * (eps_hi - 1) + 1.0))
        s.add_layer(thickness=20.0, eps=film)
        return s

    def record(tag, s):
        s.source_planewave(amplitude=[0.8, 0.6], direction="forward")
        ex, ey, ez = s.field_xy(torch.tensor(xs), torch.tensor(ys))
        out[tag + "_field"] = torch.stack((ex, ey, ez), 2).numpy()
        fx, fy, fz = s.Floquet_mode()
        out[tag + "_floquet"] = torch.stack((fx, fy, fz), 2).numpy()

    s = build(False)
    s.solve_global_smatrix()
    for i, name in enumerate(("S11", "S21", "S12", "S22")):
        out["eig_" + name] = s.S[i].numpy()
    record("eig", s)
    s.add_defect(layer_num=0, deps=torch.tensor(dgeo * (eps_hi - 1)))
    s.solve_global_smatrix()
    record("defect", s)
    out["defect_S21"] = s.S[1].numpy()

    f = build(True)
    f.solve_global_smatrix()
    out["fast_S21"] = f.S[1].numpy()
    record("fast", f)

    r = rcwa(freq=1 / 193.0, order=[2, 2], L=[600.0, 400.0], dtype=dt, device=dev)
    r.add_input_layer(eps=1.0)
    r.set_incident_angle(inc_ang=0.0, azi_ang=0.0)
    r.add_layer(thickness=30.0, eps=torch.tensor(geo * (eps_hi - 1) + 1.0))
    r.add_layer(thickness=15.0, eps=film)
    r.add_layer(thickness=25.0, eps=torch.tensor(np.roll(geo, 5, axis=1) * (eps_hi - 1) + 1.0))
    r.solve_global_smatrix_enhanced(1, 3, 4)
    out["repeat_S21"] = r.S[1].numpy()
    record("repeat", r)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
