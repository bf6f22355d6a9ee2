"""GPU vortex-lattice wing loads (aero_vlm_loads, csrc/vlm.hip) against the FP64 oracle,
bit for bit: the W17 front and rear wings over many setups in one batched call,
uneven lattices from empty to the maximum size, row-swapping and singular systems, and
the reference README's Gurney-flap gain (README.md:47) on the GPU's own numbers."""
import json
import math
import time
from pathlib import Path

import numpy as np
import pytest

from aerolbm import vlm as V
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
VO = orc.VLMOracle()
RIG = json.loads((GOLD / "w17_rig.json").read_text())


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def assert_same(got_out, got_gamma, ref_out, ref_gamma, what):
    """Bit-identical, NaNs matched as NaNs."""
    for g, r, name in ((got_out, ref_out, "out"), (got_gamma, ref_gamma, "gamma")):
        g, r = np.asarray(g, np.float64), np.asarray(r, np.float64)
        assert g.shape == r.shape, (what, name)
        nan = np.isnan(r)
        assert np.array_equal(np.isnan(g), nan), (what, name)
        assert np.array_equal(bits(g)[~nan], bits(r)[~nan]), (what, name, g[~nan][:4], r[~nan][:4])


def check_batch(lattices, speeds, what):
    print(f"vlm {what}: {len(lattices)} lattices, sizes {[len(h) for h in lattices]}", flush=True)
    loads, gammas = V.solve_lattices(lattices, speeds, want_gamma=True)
    for k, (h, s) in enumerate(zip(lattices, speeds)):
        ref_out, ref_g = VO.loads(h, s)
        got = [loads[k].downforceN, loads[k].dragN, loads[k].copX, loads[k].dragY]
        assert_same(got, gammas[k], ref_out, ref_g, f"{what}[{k}]")
    return loads


def flat_wing(b, c, alpha, H, strips=10, panels=4):
    st = [{"z": 0.0, "le": [0.0, 0.0], "mid": [0.5 * c, 0.0], "te": [c, 0.0]},
          {"z": b / 2, "le": [0.0, 0.0], "mid": [0.5 * c, 0.0], "te": [c, 0.0]}]
    ca, sa = math.cos(alpha), math.sin(alpha)
    pose = V.WingPose(lambda e, z, p: (p[0] * ca + p[1] * sa, -p[0] * sa + p[1] * ca),
                      lambda p: (p[0], p[1] + H, p[2]))
    return V.build_lattice([st], pose, (0.0, b / 2), strips, panels)


SETUPS = [
    V.W17WingSetup(),
    V.W17WingSetup(gurney=True),
    V.W17WingSetup(frontDeg=7.0, rearFlapDeg=9.0, rearMainDeg=3.0),
    V.W17WingSetup(frontDeg=-9.0, rearFlapDeg=-13.0, rearMainDeg=-4.0),
    V.W17WingSetup(flapScale=0.72, rearTwistDeg=-16.0),
    V.W17WingSetup(rearTwistDeg=6.0, gurney=True, frontDeg=3.5),
    V.W17WingSetup(rideH=0.08, rearRideH=0.26),
    V.W17WingSetup(rideH=0.2, rearRideH=0.1, flapScale=0.72),
]


def test_flat_wing_bit_exact():
    check_batch([flat_wing(6.0, 1.0, 2 * V.DEG, 0.7)], [50.0], "flat")


def test_w17_setups_one_batch():
    lattices, speeds = [], []
    for k, st in enumerate(SETUPS):
        fl, rl = V.w17_lattices(RIG, st)
        lattices += [fl, rl]
        speeds += [250 / 3.6 + k, 250 / 3.6 + k]
    loads = check_batch(lattices, speeds, "w17")
    assert all(ld.downforceN > 0 for ld in loads)


def test_w17_wing_loads_api():
    """w17_wing_loads / w17_wing_loads_many: a wing that is off gives NO_LOAD, the
    others the same numbers as the single-lattice path."""
    speed = 250 / 3.6
    one = V.w17_wing_loads(RIG, V.W17WingSetup(), speed)
    fl, rl = V.w17_lattices(RIG, V.W17WingSetup())
    assert one["front"] == V.wing_loads(RIG["wings"]["front"], V.w17_poses(RIG, V.W17WingSetup())[0],
                                        V.FRONT_SPAN, speed)
    many = V.w17_wing_loads_many(RIG, [V.W17WingSetup(frontOn=False), V.W17WingSetup(rearOn=False)],
                                 [speed, speed])
    assert many[0]["front"] == V.NO_LOAD and many[1]["rear"] == V.NO_LOAD
    assert many[0]["rear"] == one["rear"] and many[1]["front"] == one["front"]
    share = V.front_axle_share(one)
    assert 0.0 < share < 1.0


def test_uneven_sizes_to_the_maximum():
    """Empty, one-horseshoe, small and maximum-size lattices in one call."""
    lats = [np.zeros((0, 20)), flat_wing(2.0, 0.5, 3 * V.DEG, 0.3, 1, 1)[:1], flat_wing(4.0, 1.0, 4 * V.DEG, 0.4, 1, 1),
            flat_wing(3.0, 0.8, 5 * V.DEG, 0.25, 3, 2), flat_wing(6.0, 1.0, -3 * V.DEG, 0.5, 64, 8)]
    assert [len(h) for h in lats] == [0, 1, 2, 12, V.MAX_HORSESHOES]
    check_batch(lats, [30.0, 40.0, 50.0, 60.0, 70.0], "uneven")


def test_row_swaps_and_singular_systems():
    """A shuffled lattice (the elimination then swaps rows at many steps), a lattice
    with repeated horseshoes (zero pivots, replaced by 1e-12), and one holding a NaN."""
    rng = np.random.default_rng(7)
    base = V.w17_lattices(RIG, V.W17WingSetup())[1]
    shuffled = base[rng.permutation(len(base))]
    dup = np.concatenate([base[:30], base[:30]])
    bad = base[:40].copy()
    bad[5, 13] = np.nan
    check_batch([shuffled, dup, bad], [69.0, 69.0, 69.0], "pivots")


def test_gurney_gain():
    """README.md:47: +10-15 % rear-wing downforce with the Gurney flap."""
    speed = 250 / 3.6
    res = V.w17_wing_loads_many(RIG, [V.W17WingSetup(), V.W17WingSetup(gurney=True)], [speed, speed])
    gain = res[1]["rear"].downforceN / res[0]["rear"].downforceN - 1
    assert 0.10 <= gain <= 0.15
    assert res[1]["front"] == res[0]["front"]


def test_ride_height_map_batch():
    """The 25 ride-height pairs of the aero map, front and rear wings: 50 lattices in
    one call, bit-identical to the oracle on a sample, and faster than one call per
    lattice."""
    from aerolbm import aeromap as AM

    pairs = [(f, r) for f in AM.REFERENCE_FRONT for r in AM.REFERENCE_REAR]
    setups = [V.W17WingSetup(rideH=f, rearRideH=r) for f, r in pairs]
    lattices, speeds = [], []
    for st in setups:
        fl, rl = V.w17_lattices(RIG, st)
        lattices += [fl, rl]
        speeds += [250 / 3.6, 250 / 3.6]
    V.solve_lattices(lattices[:2], speeds[:2])  # warm
    t0 = time.perf_counter()
    loads, _ = V.solve_lattices(lattices, speeds)
    batched = time.perf_counter() - t0
    t0 = time.perf_counter()
    for h, s in zip(lattices[:10], speeds[:10]):
        V.solve_lattices([h], [s])
    single = (time.perf_counter() - t0) / 10
    print(f"vlm map: 50 lattices in {batched * 1e3:.2f} ms, one lattice per call {single * 1e3:.2f} ms", flush=True)
    for k in (0, 17, 49):
        ref = VO.loads(lattices[k], speeds[k])[0]
        got = [loads[k].downforceN, loads[k].dragN, loads[k].copX, loads[k].dragY]
        assert np.array_equal(bits(got), bits(ref)), k
    assert batched < 50 * single
    # ground effect across the map: lower front ride height, more front-wing load
    front = np.array([loads[2 * k].downforceN for k in range(25)]).reshape(5, 5)
    assert (np.diff(front[:, 0]) < 0).all()
    bal = AM.wing_balance(RIG, V.W17WingSetup(), pairs, 250 / 3.6)
    assert len(bal) == 25 and all(0.0 < b < 1.0 for b in bal)


def test_w17_map_cells_post_processing():
    """w17_map_cells: each point's forces through withPanelWings with the panel loads at
    its ride heights (one batched solve), bal their front-axle share (main.ts:669-675,
    aeromap.ts:96-101)."""
    from aerolbm import aeromap as AM
    from aerolbm import forces as FP
    from aerolbm.engine import ForcePair, Forces

    pairs = [(f, r) for f in AM.REFERENCE_FRONT for r in AM.REFERENCE_REAR]
    lbm = [Forces(dragN=0.0, downforceN=0.0, body=ForcePair(800.0 + 10 * k, -200.0 + 5 * k),
                  frontWing=ForcePair(0.0, 0.0), rearWing=ForcePair(0.0, 0.0)) for k in range(25)]
    kmh = 250.0
    cells = FP.w17_map_cells(lbm, pairs, RIG, V.W17WingSetup(), kmh, sidepods=1, cooling=0)
    q = 0.5 * 1.225 * (kmh / 3.6) ** 2
    for k, (c, (f, r)) in enumerate(zip(cells, pairs)):
        ld = V.w17_wing_loads(RIG, V.W17WingSetup(rideH=f, rearRideH=r), kmh / 3.6)
        body_down = (lbm[k].body.down / q + 0.31 + 0.035 + 0.01) * q
        body_drag = (lbm[k].body.drag / q * 0.69 - 0.05 - 0.03) * q
        down = body_down + ld["front"].downforceN + ld["rear"].downforceN
        drag = body_drag + ld["front"].dragN + ld["rear"].dragN
        assert c["down"] == pytest.approx(down / 9.81, rel=1e-12)
        assert c["eff"] == pytest.approx(down / max(drag, 1.0), rel=1e-12)
        assert c["bal"] == V.front_axle_share(ld)
