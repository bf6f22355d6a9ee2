"""Vortex-lattice wing loads of the W17 (the reference's panel method), solved on the
GPU by aero_vlm_loads, many wing setups per call.

Reference: src/engine/vlm.ts — station (:68-77), camberPoint (:80-84), buildLattice
(:86-113), wingLoads (:140-182, here on the GPU), withGurney (:204-221) and
w17WingLoads (:234-271) with its wing poses; the setup a UI state maps to is
src/main.ts:550-571, the front-axle share of the wing load :618-625.

The lattice is built here in FP64 with the reference's operations in its order;
Math.hypot is V8's (js_hypot).  Each horseshoe is one row of 20 doubles:
a, b, ta, tb, cp, n, area, 0.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check, ptr
from .w17 import FRONT_LEVEL_FLAP, REAR_LEVEL_FLAP, REAR_LEVEL_MAIN, REAR_SHAPE_TWIST, THIN_FLAP_SCALE

MAX_HORSESHOES = 1024  # AERO_VLM_MAX_HORSESHOES
REC = 20
# vlm.ts:31-33
FRONT_SPAN = (0.09, 0.62)
REAR_SPAN = (0.0, 0.53)
GURNEY_HEIGHT = 0.004  # vlm.ts:204
DEG = math.pi / 180
# car-model.ts:15-20, :37-46
FRONT_AXLE_X = 1.0
BASE_RIDE_H = 0.092
REAR_AXLE_X = 4.4
FRONT_DZ = 0.01
FRONT_BINS = 71
FRONT_FADE_IN = (0.16, 0.22)
FRONT_FADE_OUT = (0.5, 0.56)
FRONT_THIN_FACTOR = 0.65

Pt = Tuple[float, float]
V3 = Tuple[float, float, float]


def js_hypot(*v: float) -> float:
    """Math.hypot as V8 evaluates it: NaN / infinity first, then the magnitudes scaled
    by the largest and their squares Kahan-summed."""
    nan = any(x != x for x in v)
    a = [0.0 if x != x else abs(x) for x in v]
    m = 0.0
    for x in a:
        if x > m:
            m = x
    if m == math.inf:
        return math.inf
    if nan:
        return math.nan
    if m == 0:
        return 0.0
    s = comp = 0.0
    for x in a:
        n = x / m
        t = n * n - comp
        p = s + t
        comp = (p - s) - t
        s = p
    return math.sqrt(s) * m


def _or_one(x: float) -> float:
    """JavaScript's `x || 1` for a number: 1 when x is 0 or NaN."""
    return 1.0 if x == 0 or x != x else x


def _lerp(a: float, b: float, t: float) -> float:
    return a + (b - a) * t


def station(stations: Sequence[dict], z: float) -> dict:
    """station (vlm.ts:68-77): the section at |z|, linear between the stored ones and
    held at the ends."""
    if z <= stations[0]["z"]:
        return stations[0]
    if z >= stations[-1]["z"]:
        return stations[-1]
    i = 0
    while stations[i + 1]["z"] < z:
        i += 1
    s0, s1 = stations[i], stations[i + 1]
    t = (z - s0["z"]) / (s1["z"] - s0["z"])
    mix = lambda a, b: [_lerp(a[0], b[0], t), _lerp(a[1], b[1], t)]  # noqa: E731
    return {"z": z, "le": mix(s0["le"], s1["le"]), "mid": mix(s0["mid"], s1["mid"]), "te": mix(s0["te"], s1["te"])}


def camber_point(s: dict, t: float) -> Pt:
    """camberPoint (vlm.ts:80-84): the quadratic Bezier through le, mid and te at t."""
    le, mid, te = s["le"], s["mid"], s["te"]
    c = (2 * mid[0] - 0.5 * (le[0] + te[0]), 2 * mid[1] - 0.5 * (le[1] + te[1]))
    u = 1 - t
    return (u * u * le[0] + 2 * u * t * c[0] + t * t * te[0], u * u * le[1] + 2 * u * t * c[1] + t * t * te[1])


@dataclass
class WingPose:
    """WingPose (vlm.ts:11-16): rotate(element, |z|, section point) and the car ->
    tunnel map."""
    rotate: Callable[[int, float, Pt], Pt]
    to_world: Callable[[V3], V3]


def build_lattice(elements: Sequence[Sequence[dict]], pose: WingPose, span: Tuple[float, float],
                  span_strips: int = 10, chord_panels: int = 4) -> np.ndarray:
    """buildLattice (vlm.ts:86-113): per element, both sides, span_strips strips of
    chord_panels panels; (n, 20) float64."""
    rows: List[List[float]] = []
    for e, stations in enumerate(elements):
        z0 = min(stations[0]["z"], span[0])
        z1 = max(stations[-1]["z"], span[1])

        def point(z, t, side, e=e, stations=stations):
            local = pose.rotate(e, z, camber_point(station(stations, z), t))
            return pose.to_world((local[0], local[1], z * side))

        for side in (1, -1):
            for i in range(span_strips):
                za = _lerp(z0, z1, i / span_strips)
                zb = _lerp(z0, z1, (i + 1) / span_strips)
                zl, zr = (za, zb) if side > 0 else (zb, za)
                for j in range(chord_panels):
                    t0, t1 = j / chord_panels, (j + 1) / chord_panels
                    tq, tc = _lerp(t0, t1, 0.25), _lerp(t0, t1, 0.75)
                    a, b = point(zl, tq, side), point(zr, tq, side)
                    zm = (zl + zr) / 2
                    cp = point(zm, tc, side)
                    p00, p01 = point(zl, t0, side), point(zr, t0, side)
                    p10, p11 = point(zl, t1, side), point(zr, t1, side)
                    u = (p11[0] - p00[0], p11[1] - p00[1], p11[2] - p00[2])
                    w = (p01[0] - p10[0], p01[1] - p10[1], p01[2] - p10[2])
                    n = (u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0])
                    ln = _or_one(js_hypot(*n))
                    ta, tb = point(zl, 1, side), point(zr, 1, side)
                    rows.append([*a, *b, *ta, *tb, *cp, n[0] / ln, n[1] / ln, n[2] / ln, ln / 2, 0.0])
    return np.array(rows, dtype=np.float64).reshape(-1, REC)


@dataclass
class WingLoads:
    """WingLoads (vlm.ts:18-25)."""
    downforceN: float
    dragN: float
    copX: float
    dragY: float


NO_LOAD = WingLoads(0.0, 0.0, 0.0, 0.0)


def solve_lattices(lattices: Sequence[np.ndarray], speeds: Sequence[float], want_gamma: bool = False):
    """aero_vlm_loads over any number of lattices in one call: (list of WingLoads,
    list of circulation arrays or None)."""
    lib = _lib.load()
    count = len(lattices)
    if count == 0:
        return [], ([] if want_gamma else None)
    if len(speeds) != count:
        raise ValueError("one speed per lattice")
    sizes = [len(h) for h in lattices]
    off = np.zeros(count + 1, np.uint32)
    off[1:] = np.cumsum(sizes)
    hs = np.ascontiguousarray(np.concatenate([np.asarray(h, np.float64).reshape(-1, REC) for h in lattices]))
    sp = np.ascontiguousarray(speeds, dtype=np.float64)
    out = np.zeros((count, 4), np.float64)
    gamma = np.zeros(max(int(off[-1]), 1), np.float64) if want_gamma else None
    check(lib.aero_vlm_loads(ptr(hs) if hs.size else None, ptr(off), count, ptr(sp), ptr(out),
                             ptr(gamma) if want_gamma else None), None, "aero_vlm_loads")
    loads = [WingLoads(*map(float, row)) for row in out]
    gam = [gamma[off[k]:off[k + 1]].copy() for k in range(count)] if want_gamma else None
    return loads, gam


def wing_loads(elements, pose: WingPose, span, speed: float, span_strips: int = 10, chord_panels: int = 4) -> WingLoads:
    """wingLoads (vlm.ts:140-182)."""
    return solve_lattices([build_lattice(elements, pose, span, span_strips, chord_panels)], [speed])[0][0]


@dataclass
class W17WingSetup:
    """W17WingSetup (vlm.ts:184-197)."""
    frontDeg: float = 0.0
    rearFlapDeg: float = 0.0
    rearMainDeg: float = 0.0
    rideH: float = 0.12
    rearRideH: float = 0.14
    gurney: bool = False
    frontOn: bool = True
    rearOn: bool = True
    rearTwistDeg: float = 0.0
    flapScale: float = 1.0


def setup_from_shape(p) -> W17WingSetup:
    """The setup main.ts:554-569 sends for a shape state (ShapeParams fields)."""
    fw = int(p.frontWing)
    rw = int(p.rearWing)
    front = FRONT_LEVEL_FLAP[fw] if 0 <= fw < 3 else FRONT_LEVEL_FLAP[0]
    rear_main = REAR_LEVEL_MAIN[rw] if 0 <= rw < 3 else REAR_LEVEL_MAIN[0]
    rear_flap = REAR_LEVEL_FLAP[rw] if 0 <= rw < 3 else REAR_LEVEL_FLAP[0]
    return W17WingSetup(frontDeg=p.frontFlapDeg + front, rearFlapDeg=p.rearFlapDeg + rear_flap,
                        rearMainDeg=rear_main + p.rearMainDeg, rideH=p.rideH, rearRideH=p.rearRideH,
                        gurney=bool(p.gurney), frontOn=bool(p.frontWingOn), rearOn=bool(p.rearWingOn),
                        rearTwistDeg=REAR_SHAPE_TWIST.get(int(p.rearShape), 0.0),
                        flapScale=THIN_FLAP_SCALE if p.thinWings else 1.0)


def _smooth(a: float, b: float, x: float) -> float:
    t = min(1.0, max(0.0, (x - a) / (b - a)))
    return t * t * (3 - 2 * t)


def _rotate_about(p: Pt, h: Pt, a: float) -> Pt:
    if a == 0:
        return p
    c, s = math.cos(a), math.sin(a)
    dx, dy = p[0] - h[0], p[1] - h[1]
    return (h[0] + c * dx - s * dy, h[1] + s * dx + c * dy)


def _scale_about(p: Pt, h: Pt, k: float) -> Pt:
    return p if k == 1 else (h[0] + (p[0] - h[0]) * k, h[1] + (p[1] - h[1]) * k)


def with_gurney(elements: Sequence[Sequence[dict]]) -> List[Sequence[dict]]:
    """withGurney (vlm.ts:207-221): a tab of GURNEY_HEIGHT normal to the last
    element's chord at its trailing edge, as one more element."""
    tab = []
    for s in elements[-1]:
        dx, dy = s["te"][0] - s["le"][0], s["te"][1] - s["le"][1]
        ln = _or_one(js_hypot(dx, dy))
        up = (-dy / ln, dx / ln)
        te = s["te"]
        tab.append({"z": s["z"], "le": te,
                    "mid": [te[0] + up[0] * GURNEY_HEIGHT * 0.5, te[1] + up[1] * GURNEY_HEIGHT * 0.5],
                    "te": [te[0] + up[0] * GURNEY_HEIGHT, te[1] + up[1] * GURNEY_HEIGHT]})
    return [*elements, tab]


def w17_poses(rig: Dict, setup: W17WingSetup) -> Tuple[WingPose, WingPose]:
    """The front and rear wing poses of w17WingLoads (vlm.ts:235-265): the articulation
    the mesh shader and the distance field apply, then ride height and rake."""
    rake = math.atan((setup.rearRideH - setup.rideH) / (REAR_AXLE_X - FRONT_AXLE_X))
    c, s = math.cos(rake), math.sin(rake)

    def to_world(p: V3) -> V3:
        bx, by = p[0] - FRONT_AXLE_X, p[1] - BASE_RIDE_H
        return (FRONT_AXLE_X + c * bx - s * by, setup.rideH + s * bx + c * by, p[2])

    def hinge(table, z: float) -> Pt:
        f = min(max(z / FRONT_DZ, 0), FRONT_BINS - 1)
        i = math.floor(f)
        j = min(i + 1, FRONT_BINS - 1)
        t = f - i
        return (_lerp(table[i][0], table[j][0], t), _lerp(table[i][1], table[j][1], t))

    def front_weight(z: float) -> float:
        return _smooth(FRONT_FADE_IN[0], FRONT_FADE_IN[1], z) * (1 - _smooth(FRONT_FADE_OUT[0], FRONT_FADE_OUT[1], z))

    def front_rotate(e: int, z: float, p: Pt) -> Pt:
        if e == 0:
            return p
        w = front_weight(z)
        h = hinge(rig["frontHinges"][e - 1], z)
        k = 1 - (1 - setup.flapScale) * FRONT_THIN_FACTOR
        return _rotate_about(_scale_about(p, h, k), h, setup.frontDeg * DEG * (0.8 if e == 1 else 1) * w)

    main_h = tuple(rig["rearMainHinge"])
    flap_h = tuple(rig["rearFlapHinge"])

    def rear_rotate(e: int, z: float, p: Pt) -> Pt:
        twist = setup.rearTwistDeg * _smooth(0.12, 0.48, z) * DEG
        if e == 0:
            return _rotate_about(p, main_h, setup.rearMainDeg * DEG + twist)
        return _rotate_about(_scale_about(p, flap_h, setup.flapScale), flap_h, setup.rearFlapDeg * DEG + twist)

    return WingPose(front_rotate, to_world), WingPose(rear_rotate, to_world)


def w17_lattices(rig: Dict, setup: W17WingSetup) -> Tuple[Optional[np.ndarray], Optional[np.ndarray]]:
    """The front and rear lattices w17WingLoads solves (None for a wing that is off)."""
    front, rear = w17_poses(rig, setup)
    rear_elements = with_gurney(rig["wings"]["rear"]) if setup.gurney else rig["wings"]["rear"]
    fl = build_lattice(rig["wings"]["front"], front, FRONT_SPAN) if setup.frontOn else None
    rl = build_lattice(rear_elements, rear, REAR_SPAN) if setup.rearOn else None
    return fl, rl


def w17_wing_loads_many(rig: Dict, setups: Sequence[W17WingSetup], speeds: Sequence[float]) -> List[Dict[str, WingLoads]]:
    """w17WingLoads (vlm.ts:234-271) for many setups, every lattice in one GPU call."""
    lattices, slots = [], []
    for k, (st, sp) in enumerate(zip(setups, speeds)):
        fl, rl = w17_lattices(rig, st)
        for name, lat in (("front", fl), ("rear", rl)):
            if lat is not None:
                slots.append((k, name))
                lattices.append((lat, sp))
    res = [{"front": NO_LOAD, "rear": NO_LOAD} for _ in setups]
    if lattices:
        loads, _ = solve_lattices([l for l, _ in lattices], [s for _, s in lattices])
        for (k, name), ld in zip(slots, loads):
            res[k][name] = ld
    return res


def w17_wing_loads(rig: Dict, setup: W17WingSetup, speed: float) -> Dict[str, WingLoads]:
    """w17WingLoads (vlm.ts:234-271): {"front": WingLoads, "rear": WingLoads}."""
    return w17_wing_loads_many(rig, [setup], [speed])[0]


def front_axle_share(loads: Dict[str, WingLoads]) -> Optional[float]:
    """frontAxleShare (main.ts:618-625): the front axle's share of the wing downforce
    from each wing's centre of pressure; None below 100 N in total."""
    wb = REAR_AXLE_X - FRONT_AXLE_X
    wings = [loads["front"], loads["rear"]]
    total = 0.0
    for w in wings:
        total += w.downforceN
    if total < 100:
        return None
    share = 0.0
    for w in wings:
        share += w.downforceN * (REAR_AXLE_X - w.copX) / wb
    return share / total
