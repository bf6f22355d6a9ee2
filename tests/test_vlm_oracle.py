"""The FP64 vortex-lattice oracle (oracle/vlm_oracle.c) pinned against what fixes it
from outside: Math.hypot values written by Node's engine, the classical Biot-Savart
law, the ground image, numpy's solver, Helmbold's lift slope, thin-aerofoil centre of
pressure, induced drag, ground effect and the reference's README claim for the Gurney
flap (README.md:47).  Also the host lattice builder's geometry and the C ABI's
argument checks (which return before any device work)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from aerolbm import vlm as V
from oracle.oracle import VLMOracle

GOLD = Path(__file__).resolve().parent / "golden"
RHO = 1.225


@pytest.fixture(scope="module")
def vo():
    return VLMOracle()


@pytest.fixture(scope="module")
def rig():
    return json.loads((GOLD / "w17_rig.json").read_text())


def _same_bits(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    nan = np.isnan(b)
    return np.array_equal(np.isnan(a), nan) and np.array_equal(a.view(np.uint64)[~nan], b.view(np.uint64)[~nan])


def test_hypot_matches_v8(vo):
    d = np.load(GOLD / "js_hypot.npz")
    v = d["inputs"]
    assert _same_bits([vo.hypot3(*r) for r in v], d["hypot3"])
    assert _same_bits([V.js_hypot(*r) for r in v], d["hypot3"])
    assert _same_bits([V.js_hypot(r[0], r[1]) for r in v], d["hypot2"])
    # a plain square root of the sum of squares differs from V8 on many of them
    with np.errstate(over="ignore", invalid="ignore"):
        plain = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    fin = np.isfinite(plain) & np.isfinite(d["hypot3"])
    assert (plain[fin] != d["hypot3"][fin]).sum() > 100


def test_segment_biot_savart(vo):
    """A segment of length L seen from its perpendicular bisector at distance d:
    |v| = G L / (4 pi d sqrt(L^2/4 + d^2)), along r0 x (p - a) (right-hand rule)."""
    L, d, G = 2.0, 10.0, 3.0
    a, b = np.array([0.0, 0.0, -L / 2]), np.array([0.0, 0.0, L / 2])
    p = np.array([d, 0.0, 0.0])
    v = vo.segment(p, a, b, G)
    mag = G * L / (4 * math.pi * d * math.sqrt(L * L / 4 + d * d))
    assert v[0] == 0 and v[2] == 0
    assert v[1] > 0  # vortex along +z seen from +x turns the flow towards +y
    assert v[1] == pytest.approx(mag, rel=1e-6)  # core 0.01 m: relative change (0.01/d)^2
    # near the core the regularisation caps the velocity
    near = vo.segment(np.array([0.005, 0.0, 0.0]), a, b, G)
    assert near[1] < G / (2 * math.pi * 0.005) / 2


def test_segment_infinite_line(vo):
    """A long segment acts as an infinite line vortex: |v| = G / (2 pi d)."""
    G, d = 2.0, 2.0
    a, b = np.array([-1e5, 0.0, 0.0]), np.array([1e5, 0.0, 0.0])
    v = vo.segment(np.array([0.0, 0.0, d]), a, b, G)
    assert np.linalg.norm(v) == pytest.approx(G / (2 * math.pi * d), rel=5e-5)
    # arbitrary orientation: velocity perpendicular to both the line and the offset
    rng = np.random.default_rng(3)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    off = np.cross(u, rng.normal(size=3))
    off *= d / np.linalg.norm(off)
    v = vo.segment(off, -1e5 * u, 1e5 * u, G)
    assert abs(v @ u) < 1e-12 and abs(v @ off) < 1e-12
    assert np.linalg.norm(v) == pytest.approx(G / (2 * math.pi * d), rel=5e-5)
    assert np.cross(u, off) @ v > 0


def test_segment_degenerate(vo):
    a, b = np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0])
    assert not vo.segment(a, a, b).any()           # on an end point: skipped
    assert not vo.segment([3.0, 0, 0], a, b).any()  # on the line: zero cross product


def _horseshoe(a, b, ta, tb, cp=(0, 0, 0), n=(0, 1, 0), area=1.0):
    return np.array([*a, *b, *ta, *tb, *cp, *n, area, 0.0], np.float64)


def test_ground_image_cancels_normal_flow(vo):
    """With its image (y -> -y, circulation negated) a horseshoe induces no flow
    through the ground plane y = 0; with the image's circulation not negated it
    would double the normal flow instead."""
    h = _horseshoe((0.1, 0.3, -0.4), (0.1, 0.32, 0.45), (0.6, 0.27, -0.4), (0.6, 0.29, 0.45))
    rng = np.random.default_rng(5)
    for _ in range(20):
        p = np.array([rng.uniform(-2, 3), 0.0, rng.uniform(-1.5, 1.5)])
        v = vo.horseshoe(p, h, 1.7)
        assert abs(v[1]) <= 1e-12 * max(1.0, np.linalg.norm(v))
        assert np.linalg.norm(v) > 1e-4
    # above the ground the image matters: a point at y = 0.1 differs from the lone vortex
    p = np.array([0.2, 0.1, 0.0])
    lone = sum(vo.segment(p, s, e, 1.7) for s, e in (((60.6, 0.27, -0.4), (0.6, 0.27, -0.4)),
                                                     ((0.6, 0.27, -0.4), (0.1, 0.3, -0.4)),
                                                     ((0.1, 0.3, -0.4), (0.1, 0.32, 0.45)),
                                                     ((0.1, 0.32, 0.45), (0.6, 0.29, 0.45)),
                                                     ((0.6, 0.29, 0.45), (60.6, 0.29, 0.45))))
    assert np.linalg.norm(vo.horseshoe(p, h, 1.7) - lone) > 0.1 * np.linalg.norm(lone)


def _biot_savart_brute(p, pts, gamma, pieces=200_000):
    """Velocity at p of the polyline pts (vortex running along it), by midpoint-rule
    quadrature of G/(4 pi) * dl x (p - l) / |p - l|^3."""
    v = np.zeros(3)
    for a, b in zip(pts[:-1], pts[1:]):
        a, b = np.asarray(a, float), np.asarray(b, float)
        t = (np.arange(pieces) + 0.5) / pieces
        l = a + np.outer(t, b - a)
        dl = (b - a) / pieces
        r = p - l
        v += gamma / (4 * math.pi) * (np.cross(dl, r) / np.linalg.norm(r, axis=1)[:, None] ** 3).sum(0)
    return v


def _horseshoe_brute(p, h, gamma):
    a, b, ta, tb = h[0:3], h[3:6], h[6:9], h[9:12]
    v = np.zeros(3)
    for sign in (1.0, -1.0):
        m = lambda q: np.array([q[0], q[1] * sign, q[2]])  # noqa: E731
        chain = [m(ta + [60.0, 0, 0]), m(ta), m(a), m(b), m(tb), m(tb + [60.0, 0, 0])]
        v += _biot_savart_brute(p, chain, gamma * sign)
    return v


def test_single_swept_horseshoe(vo):
    """One skewed horseshoe: the circulation and the Kutta-Joukowski force from a
    brute-force quadrature of the Biot-Savart law (with the ground image); a single
    panel's centre of pressure and drag height are its bound vortex midpoint."""
    # distances of a metre or more, where the 0.01 m core changes velocities by < 1e-4
    H, speed = 3.0, 30.0
    a, b = np.array([0.0, H, -2.0]), np.array([1.4, H + 0.08, 2.0])
    ta, tb = a + [1.6, -0.12, 0.0], b + [1.6, -0.12, 0.0]
    n = np.array([0.06, -1.0, 0.01])
    n /= np.linalg.norm(n)
    cp = (a + b) / 2 + [1.2, -0.08, 0.0]
    h = _horseshoe(a, b, ta, tb, cp, n, 0.4)
    out, gamma = vo.loads(h[None, :], speed)
    g_bf = -(speed * n[0]) / (_horseshoe_brute(cp, h, 1.0) @ n)
    assert gamma[0] == pytest.approx(g_bf, rel=2e-4)
    m = (a + b) / 2
    vel = np.array([speed, 0.0, 0.0]) + _horseshoe_brute(m, h, gamma[0])
    F = RHO * gamma[0] * np.cross(vel, b - a)
    q = 0.5 * RHO * speed ** 2
    assert abs(out[0]) > 1 and abs(F[0]) > 1
    assert out[0] == pytest.approx(-F[1], rel=2e-4)
    assert out[1] - 0.018 * q * 0.4 == pytest.approx(F[0], rel=2e-3)
    assert out[2] == pytest.approx(m[0], rel=1e-12)
    assert out[3] == pytest.approx(m[1], rel=1e-12)


def test_solve_matches_numpy(vo):
    rng = np.random.default_rng(11)
    for n in (1, 2, 7, 50, 120):
        A = rng.normal(size=(n, n)) + n ** 0.5 * np.eye(n)
        b = rng.normal(size=n)
        assert np.allclose(vo.solve(A, b), np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    # zero diagonal: only row swaps make it solvable
    P = np.eye(6)[[3, 0, 5, 1, 2, 4]] * np.arange(1, 7)[:, None]
    b = rng.normal(size=6)
    assert np.allclose(vo.solve(P, b), np.linalg.solve(P, b), rtol=1e-13)


def test_solve_zero_pivot_rule(vo):
    """`d || 1e-12` (vlm.ts:123, :135): a zero pivot is replaced by 1e-12."""
    b = np.array([1.0, -2.0, 3.0])
    assert np.array_equal(vo.solve(np.zeros((3, 3)), b), b / 1e-12)


def flat_wing(b, c, alpha, H, camber=0.0):
    """A flat (or cambered) rectangular wing of span b and chord c, leading edge at
    x = 0, pitched by alpha (nose up for alpha > 0) at height H."""
    st = [{"z": 0.0, "le": [0.0, 0.0], "mid": [0.5 * c, camber], "te": [c, 0.0]},
          {"z": b / 2, "le": [0.0, 0.0], "mid": [0.5 * c, camber], "te": [c, 0.0]}]
    ca, sa = math.cos(alpha), math.sin(alpha)
    pose = V.WingPose(lambda e, z, p: (p[0] * ca + p[1] * sa, -p[0] * sa + p[1] * ca),
                      lambda p: (p[0], p[1] + H, p[2]))
    return V.build_lattice([st], pose, (0.0, b / 2))


def test_lattice_geometry():
    hs = flat_wing(6.0, 1.0, 0.0, 1.0)
    assert hs.shape == (80, 20)
    assert hs[:, 18].sum() == pytest.approx(6.0, rel=1e-12)      # planform area
    # n = (p11 - p00) x (p01 - p10): -y on both halves of a flat wing
    assert np.allclose(hs[:, 15:18], [0.0, -1.0, 0.0])
    # bound vortex at the panel quarter chord, collocation point at three quarters
    x = hs[:, 0]
    assert set(np.round(x, 12)) == {0.0625, 0.3125, 0.5625, 0.8125}
    assert set(np.round(hs[:, 12], 12)) == {0.1875, 0.4375, 0.6875, 0.9375}
    # trailing-edge points at x = c; the two halves mirror in z with a and b swapped,
    # so every bound vortex runs towards +z
    assert np.allclose(hs[:, 6], 1.0) and np.allclose(hs[:, 9], 1.0)
    assert np.array_equal(hs[:40, 2], -hs[40:, 5]) and np.array_equal(hs[:40, 5], -hs[40:, 2])


@pytest.mark.parametrize("aspect", [4.0, 6.0, 8.0])
def test_flat_wing_helmbold(vo, aspect):
    """Far from the ground a flat rectangular wing's lift slope is close to Helmbold's
    2 pi A / (2 + sqrt(A^2 + 4)), its centre of pressure near the quarter chord, its
    induced drag near CL^2 / (pi A), and its drag acts at the wing's height."""
    speed, alpha, H = 50.0, 2 * V.DEG, 50.0
    q = 0.5 * RHO * speed ** 2
    hs = flat_wing(aspect, 1.0, alpha, H)
    out, gamma = vo.loads(hs, speed)
    S = hs[:, 18].sum()
    CL = -out[0] / (q * S)
    helm = 2 * math.pi * aspect / (2 + math.sqrt(aspect * aspect + 4))
    assert CL / alpha / helm == pytest.approx(0.965, abs=0.03)
    assert out[2] == pytest.approx(0.25, abs=0.02)
    CDi = (out[1] - 0.018 * q * S) / (q * S)
    assert CDi * math.pi * aspect / CL ** 2 == pytest.approx(1.0, abs=0.1)
    assert H - math.sin(alpha) <= out[3] <= H
    # symmetric loading across the span
    assert np.allclose(gamma[:40], gamma[40:], rtol=1e-9)


def test_flat_wing_linear_in_alpha_and_speed_squared(vo):
    a1 = vo.loads(flat_wing(6.0, 1.0, 1 * V.DEG, 50.0), 40.0)[0]
    a2 = vo.loads(flat_wing(6.0, 1.0, 2 * V.DEG, 50.0), 40.0)[0]
    s2 = vo.loads(flat_wing(6.0, 1.0, 2 * V.DEG, 50.0), 80.0)[0]
    assert a2[0] / a1[0] == pytest.approx(2.0, rel=2e-3)
    assert s2[0] / a2[0] == pytest.approx(4.0, rel=1e-9)


def test_ground_effect(vo):
    """An inverted wing makes more downforce the closer it runs to the ground."""
    downs = [vo.loads(flat_wing(6.0, 1.0, -4 * V.DEG, H), 50.0)[0][0] for H in (50.0, 2.0, 0.5, 0.25)]
    assert downs[0] > 0
    assert all(b > a for a, b in zip(downs, downs[1:]))
    assert downs[-1] / downs[0] > 1.5


def test_w17_gurney_gain(vo, rig):
    """README.md:47: the Gurney flap's effective height is calibrated to the usual
    +10-15 % downforce of the rear wing."""
    speed = 250 / 3.6
    _, plain = V.w17_lattices(rig, V.W17WingSetup())
    _, gur = V.w17_lattices(rig, V.W17WingSetup(gurney=True))
    assert plain.shape == (240, 20) and gur.shape == (320, 20)
    d0 = vo.loads(plain, speed)[0][0]
    d1 = vo.loads(gur, speed)[0][0]
    assert 0.10 <= d1 / d0 - 1 <= 0.15


def test_w17_setups_move_loads_the_right_way(vo, rig):
    """More flap angle -> more downforce and drag; lower front ride height -> more
    front downforce (ground effect); the rear wing's centre of pressure sits behind the
    rear axle and the front's ahead of the front axle."""
    speed = 250 / 3.6
    base = V.W17WingSetup()
    fl, rl = V.w17_lattices(rig, base)
    f0, r0 = vo.loads(fl, speed)[0], vo.loads(rl, speed)[0]
    assert f0[0] > 0 and r0[0] > 0
    assert f0[2] < V.FRONT_AXLE_X and r0[2] > V.REAR_AXLE_X
    fh, rh = V.w17_lattices(rig, V.W17WingSetup(frontDeg=7.0, rearFlapDeg=9.0, rearMainDeg=3.0))
    assert vo.loads(fh, speed)[0][0] > f0[0]
    hi = vo.loads(rh, speed)[0]
    assert hi[0] > r0[0] and hi[1] > r0[1]
    low, _ = V.w17_lattices(rig, V.W17WingSetup(rideH=0.08, rearRideH=0.14))
    assert vo.loads(low, speed)[0][0] > f0[0]
    share = V.front_axle_share({"front": V.WingLoads(*f0), "rear": V.WingLoads(*r0)})
    assert 0.0 < share < 1.0


def test_setup_from_shape():
    from aerolbm.engine import ShapeParams

    p = ShapeParams()
    p.frontWing, p.rearWing, p.frontFlapDeg, p.rearFlapDeg, p.rearMainDeg = 1, 2, 1.5, -2.0, 0.5
    p.rearShape, p.thinWings, p.gurney = 1, True, True
    s = V.setup_from_shape(p)
    assert (s.frontDeg, s.rearFlapDeg, s.rearMainDeg) == (8.5, -15.0, -3.5)
    assert s.rearTwistDeg == -16.0 and s.flapScale == 0.72 and s.gurney


def test_abi_argument_checks():
    """aero_vlm_loads rejects bad arguments before it touches the device."""
    from aerolbm import _lib

    lib = _lib.load()
    hs = np.zeros((4, 20))
    out = np.zeros(8)
    sp = np.array([10.0, 10.0])

    def call(off, count, speeds=sp, h=hs):
        off = np.ascontiguousarray(off, np.uint32)
        return lib.aero_vlm_loads(h.ctypes.data, off.ctypes.data, count, speeds.ctypes.data, out.ctypes.data, None)

    assert call([0, 2, 4], 0) == _lib.AERO_E_INVALID
    assert call([1, 2, 4], 2) == _lib.AERO_E_INVALID                   # offsets[0] != 0
    assert call([0, 3, 2], 2) == _lib.AERO_E_INVALID                   # decreasing
    assert call([0, 2, 1030], 2) == _lib.AERO_E_INVALID                # above the maximum
    assert call([0, 2, 4], 2, np.array([10.0, np.inf])) == _lib.AERO_E_INVALID
    assert b"speed" in lib.aero_global_error()
    assert lib.aero_vlm_loads(None, None, 1, None, None, None) == _lib.AERO_E_INVALID
