"""The reference's one published answer: its README states that a 1 m cube gives
Cd ~ 1.04 (experimental ~ 1.05), without the tunnel, resolution or averaging window.
This test runs the cube workload (22 cells per side in a 128x64x64 half tunnel,
u = 0.05, tau0 = 0.51, Cs = 0.1, 2000 steps, Cd averaged over steps 1001-2000) and
the reference app's own defaults (medium quality 232x68x52, u = 0.0925, tau0 = 0.508,
Cs = 0.17), and records Cd.  Both runs follow the reference's kernels bit for bit (exact
mode; see the parity tests); the measured values (about 1.5 and 1.3) match a cube at
the configurations' Reynolds numbers (about 330 and 770) rather than the README's
1.04, so the test bounds Cd to the physically plausible band for those Reynolds
numbers and prints the value."""
import numpy as np
import pytest

import aerolbm
from aerolbm import workloads as W

pytestmark = pytest.mark.gpu


def cd_series(e, chunks, chunk=100):
    out = []
    for _ in range(chunks):
        e.step(chunk)
        d, _, n = e.read_forces_raw(reset=True)
        out.append(W.cd_from_forces(d[0] + d[3] + d[6], n, e.h, e.u_lb, 1.0))
    return np.array(out)


def test_cube_cd_workload_2000_steps():
    wl = W.cube()
    e = aerolbm.HipAeroEngine(math="exact", **wl.engine_kwargs())
    e.set_shape(wl.shape)
    e.reset()
    s = cd_series(e, 20)
    cd = float(s[10:].mean())
    print(f"cube workload: Cd (steps 1001-2000) = {cd:.4f}; Re = {0.05 * 22 / ((0.51 - 0.5) / 3):.0f}", flush=True)
    assert 1.2 < cd < 1.8
    assert np.abs(s[10:] - cd).max() < 0.1 * cd  # settled: the window's 100-step means stay within 10%


def test_cube_cd_reference_app_defaults():
    e = aerolbm.HipAeroEngine("medium")
    e.set_shape(aerolbm.ShapeParams(kind=aerolbm.ShapeKind.Cube))
    e.reset()
    s = cd_series(e, 40)
    cd = float(s[20:].mean())
    print(f"app defaults (medium): Cd (steps 2001-4000) = {cd:.4f}; README 1.04", flush=True)
    assert 1.0 < cd < 1.6
