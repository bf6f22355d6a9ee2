"""The committed golden fixtures (tests/golden, written by make_golden.py from the
oracle): their SHA-256 manifest, and the oracle reproducing every stored answer bit
for bit from the stored inputs, so that a change to the oracle cannot pass unnoticed."""
import hashlib
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLD / name, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_manifest_hashes():
    lines = (GOLD / "MANIFEST.sha256").read_text().split("\n")
    entries = [ln.split() for ln in lines if ln.strip()]
    assert len(entries) == 6
    for digest, name in entries:
        assert hashlib.sha256((GOLD / name).read_bytes()).hexdigest() == digest, name


def test_oracle_reproduces_lbm_fixture():
    d = load("lbm_mixed_24x12x8.npz")
    g = orc.grid(tuple(int(v) for v in d["dims"]), float(d["h"]), (0.0, 0.0, 0.0))
    sim = orc.OSim(*[float(v) for v in d["sim"]], 0.0)
    F32 = orc.LBMOracle("f32")
    for steps in (1, 100):
        f, fields, st = F32.run(g, sim, d["f0"], d["solid"], d["fields0"], d["emit"], steps)
        assert np.array_equal(f, d[f"f_{steps}"])
        assert np.array_equal(fields, d[f"fields_{steps}"])
        assert np.array_equal(np.array(st.forces_q), d[f"forces_q_{steps}"])
        assert np.array_equal(np.array(st.forces_d), d[f"forces_d_{steps}"])
        assert st.guard_hits == int(d[f"guard_{steps}"])
    # every group carries force, and the FP64 oracle agrees to 1e-4 of the largest sum
    fd, f64 = d["forces_d_100"], d["forces_d64_100"]
    assert np.abs(fd[[0, 3, 6]]).min() > 0
    assert np.abs(fd - f64).max() <= 1e-4 * np.abs(f64).max()


def test_oracle_reproduces_heat_fixture():
    d = load("heat_24x12x8.npz")
    g = orc.grid((24, 12, 8), 0.1, (0.0, 0.0, 0.0))
    hp = orc.OHeat(*[float(v) for v in d["heat"]], 0.0)
    tout = orc.LBMOracle("f32").heat(g, hp, d["fields"], d["emit"], d["tin"])
    assert np.array_equal(tout, d["tout"])


@pytest.mark.parametrize("name,kind", [("sphere", 1), ("cube", 2), ("wing", 3), ("f1", 0)])
def test_oracle_reproduces_voxel_fixture(name, kind):
    d = load("voxelize_small.npz")
    dims = tuple(int(v) for v in d[f"{name}_dims"])
    g = orc.grid(dims, float(d[f"{name}_h"]), tuple(float(v) for v in d[f"{name}_origin"]))
    s = orc.OShape()
    s.kind = kind
    s.pw, s.ph, s.pScale = 256, 128, 0.02
    s.rideH, s.rearRide, s.aoa = 0.10, 0.12, 6 * math.pi / 180
    s.frontFlap, s.rearMain, s.rearFlap = 14 * math.pi / 180, 9 * math.pi / 180, 32 * math.pi / 180
    s.heatBrakes, s.heatEngine, s.heatRad = 520.0, 620.0, 75.0
    s.duct = 1.0
    solid, emit, sdf = orc.SDFOracle().voxelize(g, s)
    assert np.array_equal(solid.astype(np.uint8), d[f"{name}_solid"])
    assert np.array_equal(emit, d[f"{name}_emit"])
    assert np.array_equal(sdf, d[f"{name}_sdf"])
    assert d[f"{name}_solid"].any()
