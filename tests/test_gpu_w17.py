"""W17 baked car (kind 5) on the GPU against the oracle, on grids in the shipped
format (f16 body 368x84x67, part grids placed by the reference's rig data,
tests/golden/w17_rig.json) with analytic content: flags bit-exact (the per-point part
rotation goes through the same FP64 sin/cos on both sides), over setups that move every
rig input."""
import json
from pathlib import Path

import numpy as np
import pytest

import aerolbm
from aerolbm import w17 as W17
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

RIG = json.loads((Path(__file__).resolve().parent / "golden" / "w17_rig.json").read_text())
SDF = orc.SDFOracle()


@pytest.mark.parametrize("setup", [
    dict(),
    dict(frontFlapDeg=6, rearWing=1, rearFlapDeg=-4, rearMainDeg=3),
    dict(rearWing=2, rearShape=1, thinWings=True),
    dict(sidepods=1, cooling=2, frontWing=1, rideH=0.16, rearRideH=0.22),
    dict(frontWingOn=False, rearShape=2, cooling=0),
])
def test_w17_voxelize_matches_oracle(setup):
    print(f"W17 {setup}", flush=True)
    body, parts = W17.synthetic_car(RIG)
    e = aerolbm.HipAeroEngine("medium")
    e.set_car_model(body, parts, RIG)
    p = aerolbm.ShapeParams(kind=aerolbm.ShapeKind.W17, **setup)
    e.set_shape(p)
    solid, emit = e.flags()
    dims, atlas = W17.parts_atlas(RIG, parts)
    SDF.set_car(body, atlas, dims, W17.rig_for_setup(RIG, p))
    g = orc.grid(e.dims, e.h, e.origin)
    osolid, oemit, osdf = SDF.voxelize(g, aerolbm.shape_uniform(p))
    diff = solid != osolid
    print(f"  {int((osolid != 0).sum())} solid, groups {[int((osolid == k).sum()) for k in (1, 2, 3)]}, "
          f"{int(diff.sum())} differ, duct cells {int((oemit == -1).sum())}", flush=True)
    assert np.array_equal(solid, osolid)
    assert np.array_equal(emit, oemit)
    assert (osolid != 0).sum() > 5000
    if setup.get("frontWingOn", True):
        assert (osolid == 2).sum() > 0
    else:
        assert (osolid == 2).sum() == 0
    e.close()


def test_w17_needs_car_model_and_valid_rig():
    e = aerolbm.HipAeroEngine("low")
    with pytest.raises(aerolbm.AeroError, match="car"):
        e.set_shape(aerolbm.ShapeParams(kind=aerolbm.ShapeKind.W17))
    body, parts = W17.synthetic_car(RIG)
    e.set_car_model(body, parts, RIG)
    bad = W17.rig_uniform(RIG, [0, 0, 0, 0])
    bad.grids[1][3] = 1e6  # part 1 placed beyond the atlas
    with pytest.raises(aerolbm.AeroError, match="atlas"):
        e.set_rig(bad)
    e.close()
