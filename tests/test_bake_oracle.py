"""Pins of the distance-grid baker oracle (oracle/bake_oracle.c, bakeGrid of
scripts/bake-car.ts:201-316): the point-triangle distance against an independent
closed form, the truncating f16 encode, the reference's own baked part grids
(committed fixture), and, where the reference is present, its shipped w17.sdf.bin and
w17.parts.bin reproduced bit for bit from w17.glb and the part tags of w17.rig.bin."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc

BO = orc.BakeOracle()
GOLD = Path(__file__).resolve().parent / "golden"
REF = Path("/root/reference/public/models")


def seg_dist(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return np.linalg.norm(p - (a + t * ab))


def tri_dist(p, a, b, c):
    """Plane projection with barycentric coordinates from a 2 x 2 solve; outside the
    triangle, the nearest of the three edges."""
    e0, e1 = b - a, c - a
    m = np.array([[e0 @ e0, e0 @ e1], [e0 @ e1, e1 @ e1]])
    s, t = np.linalg.solve(m, [e0 @ (p - a), e1 @ (p - a)])
    if s >= 0 and t >= 0 and s + t <= 1:
        return np.linalg.norm(p - (a + s * e0 + t * e1))
    return min(seg_dist(p, a, b), seg_dist(p, b, c), seg_dist(p, c, a))


def test_point_triangle_distance():
    rng = np.random.default_rng(0)
    for _ in range(3000):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2
        got = BO.closest(p, np.concatenate([a, b, c]))
        assert got == pytest.approx(tri_dist(p, a, b, c), rel=1e-9, abs=1e-12)


def test_to_half_truncates():
    rng = np.random.default_rng(1)
    for v in np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-6, 5, 2000), [0.0, -0.0, 7e4, -7e4]]):
        h = BO.to_half(v)
        f = float(np.float32(v))
        back = float(np.array([h], np.uint16).view(np.float16)[0])
        if abs(f) >= 65536.0:
            assert math.isinf(back) and math.copysign(1, back) == math.copysign(1, f)
        elif abs(f) < 2.0 ** -14:
            assert back == 0.0 and math.copysign(1, back) == math.copysign(1, f)
        else:
            assert abs(back) <= abs(f) and math.copysign(1, back) == math.copysign(1, f)
            up = float(np.array([h + 1], np.uint16).view(np.float16)[0])
            assert abs(up) > abs(f)  # truncation: the next f16 away from zero is past the value


def test_oracle_reproduces_the_reference_part_grids():
    with np.load(GOLD / "w17_part_grids.npz", allow_pickle=False) as d:
        for part in (1, 3):
            lay = {"origin": d[f"p{part}_origin"], "h": float(d[f"p{part}_h"]), "dims": d[f"p{part}_dims"]}
            band, closing, shell = d[f"p{part}_params"]
            got = BO.bake(d[f"p{part}_tris"], lay, band, closing, shell)
            assert np.array_equal(got, d[f"p{part}_expected"]), part


def sphere_mesh(r, centre, n=24):
    """A UV sphere, (m, 9) corners."""
    th = np.linspace(0, np.pi, n + 1)
    ph = np.linspace(0, 2 * np.pi, 2 * n + 1)
    P = np.array([[[r * np.sin(t) * np.cos(p), r * np.cos(t), r * np.sin(t) * np.sin(p)] for p in ph] for t in th])
    P += centre
    tris = []
    for i in range(n):
        for j in range(2 * n):
            a, b, c, e = P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]
            tris += [np.concatenate([a, b, c]), np.concatenate([a, c, e])]
    return np.array(tris)


def test_sphere_closed_forms():
    """A closed sphere mesh on the symmetry plane.  Outside, the value is |p - c| - r -
    shell within the mesh faceting and f16 rounding, and beyond the band exactly the
    f16 of f32(band) - shell.  The closing regrowth (ceil(closing / h) 26-neighbour steps
    from the flood front, which stops `closing` outside the surface) reaches at most
    closing + 2 sqrt(3) h - closing = 0.07 m past the front, i.e. 0.04 m inside; every
    cell deeper than that is negative."""
    r, c = 0.2, np.array([0.5, 0.4, 0.0])
    tris = sphere_mesh(r, c)
    lay = {"origin": (0.0, 0.0, 0.0), "h": 0.02, "dims": (50, 40, 20)}
    band, closing, shell = 0.1, 0.03, 0.004
    words = BO.bake(tris, lay, band, closing, shell).reshape(20, 40, 50)
    out = words.view(np.float16).astype(np.float64)
    z, y, x = np.meshgrid(np.arange(20), np.arange(40), np.arange(50), indexing="ij")
    p = np.stack([x * 0.02, y * 0.02, z * 0.02], -1)
    dist = np.linalg.norm(p - c, axis=-1)
    deep = dist < r - 0.045
    assert deep.sum() > 1000 and np.all(out[deep] < 0)
    near_out = (dist > r + 0.01) & (dist < r + 0.08)
    np.testing.assert_allclose(out[near_out], dist[near_out] - r - shell, atol=2e-3)
    far = dist > r + 0.12
    assert far.sum() > 1000 and np.all(words[far] == BO.to_half(float(np.float32(band)) - shell))


@pytest.mark.skipif(not (REF / "w17.glb").exists(), reason="the reference's model files are absent")
def test_oracle_reproduces_the_shipped_w17_grids():
    """w17.sdf.bin (368 x 84 x 67) and all six grids of w17.parts.bin, as the reference's
    baker wrote them, from w17.glb and the per-vertex part tags of w17.rig.bin."""
    from aerolbm import bake as B

    tags = np.fromfile(REF / "w17.rig.bin", dtype=np.uint8)
    body, parts, grids = B.bake_car(REF / "w17.glb", tags, bake=BO.bake)
    assert np.array_equal(body, np.fromfile(REF / "w17.sdf.bin", dtype="<u2"))
    assert np.array_equal(parts, np.fromfile(REF / "w17.parts.bin", dtype="<u2"))
    ref = json.loads((REF / "w17.rig.json").read_text())["partGrids"]
    for g, r in zip(grids, ref):
        assert g["part"] == r["part"] and g["dims"] == r["dims"] and g["offset"] == r["offset"]
        assert g["origin"] == r["origin"] and g["h"] == r["h"]
