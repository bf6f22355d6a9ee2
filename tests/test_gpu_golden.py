"""The library against the committed golden fixtures (tests/golden): exact-mode steps
reproduce the stored populations, fields, fixed-point and FP64 force sums bit for bit;
fast mode within 1e-5 relative; heat bit for bit; the voxelizer bit for bit."""
import math
from pathlib import Path

import numpy as np
import pytest

import aerolbm
from aerolbm._lib import Shape

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLD / name, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("math_", ["exact", "fast"])
def test_lbm_fixture(math_):
    d = load("lbm_mixed_24x12x8.npz")
    u, tau0, smag = (float(v) for v in d["sim"])
    e = aerolbm.HipAeroEngine(dims=tuple(int(v) for v in d["dims"]), h=float(d["h"]), origin=(0, 0, 0), u_lb=u,
                              tau0=tau0, smag=smag, math=math_)
    e.set_flags(d["solid"], d["emit"])
    for steps in (1, 100):
        e.reset()
        e.write_populations(d["f0"])
        e.read_forces_raw(reset=True)
        e.step(steps)
        f = e.populations()
        dsum, q, n = e.read_forces_raw(reset=True)
        want = d[f"f_{steps}"]
        rel = np.abs(f - want) / np.maximum(np.abs(want), 1e-30)
        print(f"{math_} {steps} steps: max rel {rel.max():.2e}", flush=True)
        if math_ == "exact":
            assert np.array_equal(f, want)
            assert np.array_equal(e.fields(), d[f"fields_{steps}"])
            assert np.array_equal(q, d[f"forces_q_{steps}"][:9])
            assert np.abs(dsum - d[f"forces_d_{steps}"]).max() <= 1e-9 * np.abs(d[f"forces_d_{steps}"]).max()
        else:
            assert rel.max() <= 1e-5
        assert n == steps and e.guard_count() == int(d[f"guard_{steps}"])
    e.close()


def test_heat_fixture():
    d = load("heat_24x12x8.npz")
    lbm = load("lbm_mixed_24x12x8.npz")
    e = aerolbm.HipAeroEngine(dims=(24, 12, 8), h=0.1, origin=(0, 0, 0), u_lb=0.06, tau0=0.51, smag=0.1,
                              math="exact")
    e.set_flags(lbm["solid"], lbm["emit"])
    e.reset()
    e.write_populations(lbm["f0"])
    e.step(100)  # fields of the fixture are those after 100 exact steps
    assert np.array_equal(e.fields(), d["fields"])
    e.write_temps(d["tin"])
    e.heat_step(1, dt=float(d["heat"][0]))
    assert np.array_equal(e.temps(), d["tout"])
    e.close()


@pytest.mark.parametrize("name,kind", [("sphere", 1), ("cube", 2), ("wing", 3), ("f1", 0)])
def test_voxel_fixture(name, kind):
    d = load("voxelize_small.npz")
    dims = tuple(int(v) for v in d[f"{name}_dims"])
    e = aerolbm.HipAeroEngine(dims=dims, h=float(d[f"{name}_h"]), origin=tuple(float(v) for v in d[f"{name}_origin"]))
    s = Shape()
    s.kind, s.pw, s.ph, s.p_scale = kind, 256, 128, 0.02
    s.ride_h, s.rear_ride, s.aoa = 0.10, 0.12, 6 * math.pi / 180
    s.front_flap, s.rear_main, s.rear_flap = 14 * math.pi / 180, 9 * math.pi / 180, 32 * math.pi / 180
    s.heat_brakes, s.heat_engine, s.heat_rad, s.duct = 520.0, 620.0, 75.0, 1.0
    e.set_shape(s)
    solid, emit = e.flags()
    want = d[f"{name}_solid"].astype(np.uint32)
    assert np.array_equal(solid, want)
    assert np.array_equal(emit, d[f"{name}_emit"])
    e.close()
