"""Pins the FP32 voxelizer oracle (oracle/sdf_oracle.c) against geometry computed
independently: closed-form cell sets for the cube and sphere, the convex-hull
distance of the wing section, clamp-projection distances of the rounded box and
wheel, and an analytic circle through the custom-profile path."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

SDF = orc.SDFOracle()


def shape(kind, **kw):
    s = orc.OShape()
    s.kind = kind
    s.pw, s.ph, s.pScale = 256, 128, 0.02
    s.rideH, s.rearRide = kw.get("rideH", 0.10), kw.get("rearRide", 0.12)
    s.aoa = kw.get("aoa", 6 * math.pi / 180)
    for k, v in kw.items():
        if hasattr(s, k) and k not in ("rideH", "rearRide", "aoa"):
            setattr(s, k, v)
    return s


def centres(g):
    nx, ny, nz = g.nx, g.ny, g.nz
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = np.float32
    px = f(g.ox) + (x.astype(f) + f(0.5)) * f(g.h)
    py = f(g.oy) + (y.astype(f) + f(0.5)) * f(g.h)
    pz = f(g.oz) + (z.astype(f) + f(0.5)) * f(g.h)
    return x.reshape(-1), px.reshape(-1).astype(np.float64), py.reshape(-1).astype(np.float64), pz.reshape(-1).astype(np.float64)


def test_cube_workload_is_22_cells_per_side():
    h = float(np.float32(1 / 22))
    ox = float(np.float32(2.6 - 42 / 22))
    g = orc.grid((128, 64, 64), h, (ox, 0.0, 0.0))
    solid, emit, _ = SDF.voxelize(g, shape(2))
    s3 = solid.reshape(64, 64, 128)
    zs, ys, xs = np.nonzero(s3)
    assert (xs.min(), xs.max(), ys.min(), ys.max(), zs.min(), zs.max()) == (31, 52, 11, 32, 0, 10)
    assert solid.sum() == 22 * 22 * 11 and set(np.unique(solid)) == {0, 1}
    # closed form: exactly the cells whose centre lies inside the 1 m box
    _, px, py, pz = centres(g)
    inside = (np.abs(px - 2.6) < 0.5) & (np.abs(py - 1.0) < 0.5) & (np.abs(pz) < 0.5)
    assert np.array_equal(solid != 0, inside)
    assert not emit.any()


def test_sphere_matches_closed_form_outside_band():
    g = orc.grid((96, 48, 40), 0.05, (0.0, 0.0, 0.0))
    solid, _, _ = SDF.voxelize(g, shape(1))
    x, px, py, pz = centres(g)
    d = np.sqrt((px - 2.6) ** 2 + (py - 1.0) ** 2 + pz ** 2) - 0.6
    thr = 0.3 * 0.05
    want = (d < thr) & (x > 1) & (x < 96 - 2)
    band = np.abs(d - thr) < 1e-5
    assert np.array_equal((solid != 0)[~band], want[~band])
    assert band.sum() < 10


def hull_dist(p2, c1, c2, r1, r2, nt=20001):
    """Distance from p2 (N,2) to the convex hull of two discs = min_t |p - c(t)| - r(t)."""
    t = np.linspace(0, 1, nt)
    c = c1[None] + (c2 - c1)[None] * t[:, None]
    r = r1 + (r2 - r1) * t
    d = np.sqrt(((p2[:, None, :] - c[None]) ** 2).sum(-1)) - r[None]
    return d.min(1)


@pytest.mark.parametrize("ang", [math.radians(4), math.radians(11), math.radians(-6)])
def test_wing_section_matches_convex_hull_distance(ang):
    le = np.array([1.6, 1.2])
    chord, thick, z0, z1 = 0.9, 0.06, -0.5, 0.5
    r1 = thick * 0.5
    r2 = max(thick * 0.12, 0.006)
    hh = max(chord - r1 - r2, 0.01)
    dirv = np.array([math.cos(ang), math.sin(ang)])
    c1 = le + r1 * dirv
    c2 = le + (r1 + hh) * dirv
    rng = np.random.default_rng(5)
    pts = np.column_stack([rng.uniform(1.4, 2.7, 400), rng.uniform(0.9, 1.6, 400), rng.uniform(-0.3, 0.3, 400)])
    d2 = hull_dist(pts[:, :2], c1, c2, r1, r2)
    dz = np.abs(pts[:, 2]) - 0.5
    exact = np.minimum(np.maximum(d2, dz), 0) + np.sqrt(np.maximum(d2, 0) ** 2 + np.maximum(dz, 0) ** 2)
    got = np.array([SDF.lib.orc_sd_wing(*p, le[0], le[1], chord, ang, thick, z0, z1) for p in pts])
    out = exact > 1e-3
    assert out.sum() > 100
    assert np.abs(got[out] - exact[out]).max() < 2e-5
    # inside points are negative
    assert np.all(got[exact < -2e-3] < 0)
    # positive angle raises the trailing edge (downforce sense)
    te = le + chord * dirv
    assert (te[1] > le[1]) == (ang > 0)


def test_roundbox_and_wheel_match_clamp_projection():
    rng = np.random.default_rng(6)
    b = np.array([0.5, 0.3, 0.4])
    r = 0.05
    pts = rng.uniform(-1, 1, (500, 3))
    inner = b - r
    q = np.clip(pts, -inner, inner)
    exact = np.sqrt(((pts - q) ** 2).sum(1)) - r
    got = np.array([SDF.lib.orc_sd_roundbox(*p, *b, r) for p in pts])
    out = exact > 0
    assert np.abs(got[out] - exact[out]).max() < 2e-6
    # wheel: rounded cylinder of radius R, half width hw, edge radius 0.045
    R, hw, rr = 0.36, 0.15, 0.045
    c = np.array([1.0, 0.36, 1.12])
    pts = c + rng.uniform(-0.7, 0.7, (500, 3))
    rho = np.hypot(pts[:, 0] - c[0], pts[:, 1] - c[1])
    az = np.abs(pts[:, 2] - c[2])
    qr = np.clip(rho, 0, R - rr)
    qz = np.clip(az, 0, hw - rr)
    exact = np.hypot(rho - qr, az - qz) - rr
    got = np.array([SDF.lib.orc_sd_wheel(*p, *c, R, hw) for p in pts])
    out = exact > 0
    assert np.abs(got[out] - exact[out]).max() < 2e-6


def test_material_groups_and_x_margin():
    """Kind 16 scene: materials 7 and 8 map to groups 2 and 3, others to 1; cells with
    x <= 1 or x >= nx - 2 are never solid (wgsl.ts:316)."""
    P = orc.OPrim
    prims = (P * 3)()
    for k, (mat, x0) in enumerate([(0, 1.0), (7, 2.0), (8, 3.0)]):
        prims[k].type, prims[k].material, prims[k].frame = 4, mat, 0
        prims[k].a[0], prims[k].a[1], prims[k].a[2], prims[k].a[3] = x0, 0.5, 0.0, 0.3
    prims2 = [prims[i] for i in range(3)]
    extra = (P * 1)()
    extra[0].type, extra[0].material, extra[0].frame = 4, 0, 0
    extra[0].a[0], extra[0].a[1], extra[0].a[2], extra[0].a[3] = 0.0, 0.5, 0.0, 0.3
    g = orc.grid((80, 24, 16), 0.05, (-0.2, 0.0, 0.0))
    solid, _, _ = SDF.voxelize(g, shape(16), prims=prims2 + [extra[0]])
    s3 = solid.reshape(16, 24, 80)
    assert s3[0, 10, int((1.0 + 0.2) / 0.05)] == 1
    assert s3[0, 10, int((2.0 + 0.2) / 0.05)] == 2
    assert s3[0, 10, int((3.0 + 0.2) / 0.05)] == 3
    assert not s3[:, :, :2].any() and not s3[:, :, -2:].any()
    assert s3[:, :, 2].any()  # the sphere at x = 0 does reach x index 2


def test_custom_profile_follows_circle():
    R, cx0, cy0, cell = 0.35, 2.0, 0.8, 0.02
    pw, ph = 256, 128
    gx = 0.0 + (np.arange(pw) + 0.5) * cell
    gy = 0.0 + (np.arange(ph) + 0.5) * cell
    X, Y = np.meshgrid(gx, gy)
    prof = (np.hypot(X - cx0, Y - cy0) - R).astype(np.float32).reshape(-1)
    s = shape(4, width=0.3)
    s.pScale = cell
    g = orc.grid((100, 50, 20), 0.04, (0.0, 0.0, 0.0))
    solid, _, _ = SDF.voxelize(g, s, profile=prof)
    x, px, py, pz = centres(g)
    d2 = np.hypot(px - cx0, py - cy0) - R
    dz = np.abs(pz) - 0.3
    exact = np.minimum(np.maximum(d2, dz), 0) + np.sqrt(np.maximum(d2, 0) ** 2 + np.maximum(dz, 0) ** 2)
    thr = 0.3 * 0.04
    band = np.abs(exact - thr) < 2e-3  # bilinear interpolation error of the 2 cm profile
    want = (exact < thr) & (x > 1) & (x < 98)
    assert np.array_equal((solid != 0)[~band], want[~band])


def f1_shape(**kw):
    s = shape(0, rideH=kw.pop("rideH", 0.10), rearRide=kw.pop("rearRide", 0.12))
    s.frontFlap = 14 * math.pi / 180
    s.rearMain = 9 * math.pi / 180
    s.rearFlap = 32 * math.pi / 180
    s.heatBrakes, s.heatEngine, s.heatRad = 520.0, 620.0, 75.0
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_f1_car_inside_reference_bounds_and_groups():
    """Kind 0 (f1.ts sdF1): every solid cell centre lies inside the box that the
    reference's shapeBounds gives for the F1 (engine.ts:505, plus the 0.3h band);
    front-wing cells sit ahead of the front axle, rear-wing cells behind the rear
    axle and above the body; brake heat sources surround the wheel hubs."""
    h = 10.4 / 232
    g = orc.grid((232, 68, 52), h, (-2.4, 0.0, 0.0))
    s = f1_shape()
    solid, emit, _ = SDF.voxelize(g, s)
    x, px, py, pz = centres(g)
    m = solid != 0
    tol = 0.3 * h
    assert px[m].min() >= -0.08 - tol and px[m].max() <= 5.46 + tol
    assert py[m].max() <= 0.10 + 1.34 + tol and pz[m].max() <= 0.98 + tol
    fw, rw = solid == 2, solid == 3
    assert fw.sum() > 100 and rw.sum() > 100
    # the rear-wing group includes its pylon, which starts at b + 0.36 (f1.ts:279)
    assert px[fw].max() < 1.0 and px[rw].min() > 4.4 and py[rw].min() > 0.10 + 0.36 - tol
    hot = emit > 0
    assert np.isin(np.unique(emit[hot]), [520.0, 416.0, 620.0, 75.0, 105.0]).all()
    near_front_hub = hot & (np.hypot(px - 1.0, py - 0.36) < 0.25)
    assert near_front_hub.sum() > 0


def test_f1_rake_lifts_the_rear():
    h = 10.4 / 232
    g = orc.grid((232, 68, 52), h, (-2.4, 0.0, 0.0))
    x, px, py, pz = centres(g)
    flat, _, _ = SDF.voxelize(g, f1_shape(rideH=0.10, rearRide=0.10))
    raked, _, _ = SDF.voxelize(g, f1_shape(rideH=0.10, rearRide=0.25))
    rear = px > 4.6
    assert py[(raked != 0) & rear].mean() > py[(flat != 0) & rear].mean() + 0.02
    front = (px > 0.2) & (px < 1.5)
    assert abs(py[(raked != 0) & front].mean() - py[(flat != 0) & front].mean()) < 0.03


# ---- W17 baked car (kind 5) ----
import json
from pathlib import Path

from aerolbm import w17 as W17
from aerolbm._lib import Rig

RIG = json.loads((Path(__file__).resolve().parent / "golden" / "w17_rig.json").read_text())
REF_MODELS = Path("/root/reference/public/models")


def test_half_decoder_is_exact_for_every_word():
    words = np.arange(65536, dtype=np.uint32)
    got = np.array([SDF.lib.orc_half_to_float(int(w)) for w in words], np.float32)
    want = words.astype(np.uint16).view(np.float16).astype(np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def test_parts_atlas_packing():
    body, parts = W17.synthetic_car(RIG)
    dims, atlas = W17.parts_atlas(RIG, parts)
    grids = RIG["partGrids"]
    assert dims == (max(g["dims"][0] for g in grids), max(g["dims"][1] for g in grids),
                    sum(g["dims"][2] + 1 for g in grids))
    a = atlas.reshape(dims[2], dims[1], dims[0])
    z0 = 0
    for g in grids:
        gx, gy, gz = g["dims"]
        src = parts[g["offset"]:g["offset"] + gx * gy * gz].reshape(gz, gy, gx)
        assert np.array_equal(a[z0:z0 + gz, :gy, :gx], src)
        assert (a[z0 + gz] == W17.ATLAS_FAR).all()        # padding slice
        assert (a[z0:z0 + gz, :, gx:] == W17.ATLAS_FAR).all()
        z0 += gz + 1


def w17_shape(ride=0.092, rear=0.092, duct=1.0):
    s = shape(5, rideH=ride, rearRide=rear)
    s.heatBrakes, s.heatEngine, s.heatRad, s.duct = 520.0, 620.0, 75.0, duct
    return s


def test_w17_body_is_the_trilinear_grid():
    """With no rake (ride = base ride 0.092 front and rear) and no pod squeeze the
    body frame is the identity, and away from wheels and parts the W17 distance is
    0.95 x the trilinear interpolation of the body grid (f1.ts:334-340,357)."""
    body, parts = W17.synthetic_car(RIG)
    dims, atlas = W17.parts_atlas(RIG, parts)
    SDF.set_car(body, atlas, dims, W17.rig_uniform(RIG, [0, 0, 0, 0]))
    grid = body.view(np.float16).astype(np.float64).reshape(67, 84, 368)
    rng = np.random.default_rng(9)
    pts = np.column_stack([rng.uniform(1.6, 3.4, 300), rng.uniform(0.8, 1.15, 300), rng.uniform(-0.9, 0.9, 300)])
    s = w17_shape(duct=0.0)
    checked = 0
    for x, y, z in pts:
        u, v, w = (x + 0.3) / 0.015, y / 0.015, abs(z) / 0.015
        i, j, k = int(np.floor(u)), int(np.floor(v)), int(np.floor(w))
        a, b, c = u - i, v - j, w - k
        cl = lambda t, n: min(max(t, 0), n - 1)
        val = 0.0
        for di, wa in ((0, 1 - a), (1, a)):
            for dj, wb in ((0, 1 - b), (1, b)):
                for dk, wc in ((0, 1 - c), (1, c)):
                    val += wa * wb * wc * grid[cl(k + dk, 67), cl(j + dj, 84), cl(i + di, 368)]
        d, m = SDF.sdf_point(s, float(x), float(y), float(z))
        if m == 12.0:
            assert abs(d - 0.95 * val) < 2e-5
            checked += 1
    assert checked > 250


@pytest.mark.skipif(not REF_MODELS.exists(), reason="the reference's model files are not present")
def test_w17_reference_model_oracle():
    """The shipped W17 grids at the app's medium tunnel: the car sits between the
    wheel axles' extremes, front- and rear-wing groups come from the articulated parts
    at the car's ends, the sidepod ducts are carved (emit = -1 in fluid duct cells)
    and brake heat sits at both axles."""
    body, parts, rig = W17.load_reference_car(REF_MODELS)
    dims, atlas = W17.parts_atlas(rig, parts)
    SDF.set_car(body, atlas, dims, W17.rig_uniform(rig, [0, 0, 0, 0]))
    h = 10.4 / 232
    g = orc.grid((232, 68, 52), h, (-2.4, 0.0, 0.0))
    solid, emit, _ = SDF.voxelize(g, w17_shape(ride=0.12, rear=0.14))
    x, px, py, pz = centres(g)
    m = solid != 0
    assert m.sum() > 10000
    assert px[m].min() > -0.4 and px[m].max() < 5.4 and py[m].max() < 1.3 and pz[m].max() < 1.05
    fw, rw = solid == 2, solid == 3
    assert fw.sum() > 50 and rw.sum() > 50
    assert px[fw].max() < 0.9 and px[rw].min() > 4.3
    assert (emit == -1.0).sum() > 20 and ((emit == -1.0) & m).sum() == 0
    assert np.isin([520.0, 416.0], np.unique(emit[m])).all()


def test_sincos_r64_against_libm():
    """The per-point sin/cos of the voxelizer (FP64 Cody-Waite + Taylor, one rounding to
    f32) equals the f32 rounding of the C library's double sin/cos on every tested angle,
    quadrant signs and exact zeros included."""
    import math

    import numpy as np

    rng = np.random.default_rng(17)
    angles = np.concatenate([rng.uniform(-7.0, 7.0, 20000), rng.normal(0, 0.3, 5000),
                             np.array([0.0, -0.0, math.pi / 2, -math.pi / 2, math.pi, 3 * math.pi / 4, 1e-30,
                                       0.7853981633974483, 100.0, -250.5])]).astype(np.float32)
    bad = 0
    for a in angles:
        s, c = SDF.sincos(float(a))
        ws, wc = float(np.float32(math.sin(float(a)))), float(np.float32(math.cos(float(a))))
        bad += (s != ws) + (c != wc)
    assert bad == 0
    assert SDF.sincos(0.0) == (0.0, 1.0)
    s, c = SDF.sincos(float("nan"))
    assert s != s and c != c
