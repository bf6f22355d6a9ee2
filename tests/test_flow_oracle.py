"""Pins of the field-consumer oracle (oracle/flow_oracle.c) by what the mathematics
fixes: exactness of trilinear sampling on uniform and linear fields, the solid
renormalisation, the mirror and clamp rules, the vorticity of a rigid rotation,
the inverse of the PCG permutation, and closed-form particle / streamline motion
in a uniform flow."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as orc

F = orc.FlowOracle()
DIMS = (12, 10, 8)
H = 0.5
ORIGIN = (-1.0, 0.2, 0.0)
G = orc.grid(DIMS, H, ORIGIN)


class View(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("u_in", "v_real", "cp_min", "cp_max", "t_min", "t_max", "speed_max",
                                         "vort_max")]


class PSim(C.Structure):
    _fields_ = [("count", C.c_uint32), ("spawn_mode", C.c_uint32), ("color_mode", C.c_uint32),
                ("frame", C.c_uint32), ("dt", C.c_float), ("life", C.c_float), ("scatter", C.c_uint32),
                ("rake_rows", C.c_uint32), ("spawn_min", C.c_float * 3), ("rake_cols", C.c_uint32),
                ("spawn_max", C.c_float * 3), ("pad", C.c_float)]


class Stream(C.Structure):
    _fields_ = [("lines", C.c_uint32), ("steps", C.c_uint32), ("seed_mode", C.c_uint32), ("color_mode", C.c_uint32),
                ("seed_min", C.c_float * 3), ("rows", C.c_uint32), ("seed_max", C.c_float * 3), ("cols", C.c_uint32),
                ("step_len", C.c_float), ("time", C.c_float), ("per_emitter", C.c_uint32), ("emitters", C.c_uint32)]


VIEW = View(0.05, 69.4, -1.1, 0.7, -3.0, 2.5, 1.5, 0.6)


def centres():
    nx, ny, nz = DIMS
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (ORIGIN[0] + (x + 0.5) * H, ORIGIN[1] + (y + 0.5) * H, ORIGIN[2] + (z + 0.5) * H)


def field_of(ux, uy, uz, rho):
    n = G.n
    f = np.empty((n, 4), np.float32)
    for k, a in enumerate((ux, uy, uz, rho)):
        f[:, k] = np.broadcast_to(np.asarray(a, np.float32), DIMS[::-1]).reshape(-1)
    return f


def interior_points(rng, k=200):
    lo = np.array(ORIGIN) + 1.0 * H
    hi = np.array(ORIGIN) + (np.array(DIMS) - 1.0) * H
    return rng.uniform(lo, hi, (k, 3)).astype(np.float32)


# ---- PCG hash (wgsl.ts:537-541) ----

def pcg_inverse(out: int) -> int:
    """Inverse of the three invertible steps of the hash, derived independently of it."""
    M = 1 << 32
    w = out ^ (out >> 22)                       # (w >> 22) ^ w, shift >= 16
    t = (w * pow(277803737, -1, M)) % M         # odd multiplier
    r = (t >> 28) + 4                           # top 4 bits survive the xorshift (r >= 4)
    s = t
    for _ in range(8):
        s = t ^ (s >> r)
    return ((s - 2891336453) * pow(747796405, -1, M)) % M


def test_pcg_is_the_published_permutation():
    rng = np.random.default_rng(0)
    for v in list(rng.integers(0, 1 << 32, 4000, dtype=np.uint64)) + [0, 1, 0xFFFFFFFF]:
        v = int(v)
        assert pcg_inverse(F.pcg(v)) == v
    outs = {F.pcg(v) for v in range(20000)}
    assert len(outs) == 20000


# ---- sampleFlow (wgsl.ts:331-355) ----

def test_uniform_field_sampled_everywhere():
    rng = np.random.default_rng(1)
    f = field_of(0.03, -0.01, 0.02, 1.01)
    t = np.full(G.n, 300.0, np.float32)
    pts = np.concatenate([interior_points(rng), [[-50.0, 0.1, 0.2], [40.0, 80.0, 9.0], [0.3, 0.4, 0.01]]]).astype(np.float32)
    out = F.sample(G, f, t, VIEW, 0, pts)
    np.testing.assert_allclose(out[:, 0:6], np.broadcast_to([0.03, -0.01, 0.02, 1.01, 300.0, 1.0], out[:, :6].shape),
                               rtol=2e-6, atol=1e-8)
    # the mirror image of the lower half: u_z flips sign, nothing else changes
    neg = pts.copy()
    neg[:, 2] = -neg[:, 2]
    m = F.sample(G, f, t, VIEW, 0, neg)
    assert np.array_equal(m[:, [0, 1, 3, 4, 5]], out[:, [0, 1, 3, 4, 5]])
    assert np.array_equal(m[:, 2], -out[:, 2])


def test_linear_field_is_reproduced_at_interior_points():
    rng = np.random.default_rng(2)
    X, Y, Z = centres()
    ux = 0.01 + 0.002 * X - 0.003 * Y + 0.001 * Z
    uy = -0.02 + 0.004 * Y
    uz = 0.005 * X + 0.002 * Z
    rho = 1.0 + 0.01 * X
    f = field_of(ux, uy, uz, rho)
    t = (20.0 + 3.0 * X + 2.0 * Y - Z).astype(np.float32).reshape(-1)
    p = interior_points(rng)
    out = F.sample(G, f, t, VIEW, 0, p)
    x, y, z = p[:, 0].astype(np.float64), p[:, 1].astype(np.float64), p[:, 2].astype(np.float64)
    want = np.stack([0.01 + 0.002 * x - 0.003 * y + 0.001 * z, -0.02 + 0.004 * y, 0.005 * x + 0.002 * z,
                     1.0 + 0.01 * x, 20.0 + 3.0 * x + 2.0 * y - z], 1)
    np.testing.assert_allclose(out[:, :5], want, rtol=0, atol=2e-6 * np.abs(want).max())
    assert np.all(out[:, 5] == out[:, 5]) and np.allclose(out[:, 5], 1.0, atol=1e-6)


def test_near_plane_corners_mirror_into_plane_zero():
    """At z = h/4 the lower corner row lies at cell z = -1, which mirrors to z = 0, so
    the sample equals the bilinear sample of plane 0 (taken at its centre height)."""
    X, Y, Z = centres()
    f = field_of(0.01 * X, 0.02 * Y, 0.03 * Z + 0.001, 1.0 + 0.0 * X)
    t = (5.0 * Z).astype(np.float32).reshape(-1)
    p = np.array([[0.3, 1.1, 0.25 * H], [1.7, 2.2, 0.25 * H]], np.float32)
    q = p.copy()
    q[:, 2] = 0.5 * H
    a = F.sample(G, f, t, VIEW, 0, p)
    b = F.sample(G, f, t, VIEW, 0, q)
    np.testing.assert_allclose(a[:, :6], b[:, :6], rtol=1e-6, atol=1e-9)


def test_solid_corners_are_dropped_and_weights_renormalised():
    rng = np.random.default_rng(3)
    X, Y, Z = centres()
    vals = rng.uniform(-0.05, 0.05, (4,) + DIMS[::-1]).astype(np.float32)
    vals[3] = 1.0 + np.abs(vals[3])
    solid = rng.random(DIMS[::-1]) < 0.3
    vals[3][solid] = 0.0
    f = field_of(*vals)
    t = rng.uniform(0, 100, G.n).astype(np.float32)
    p = interior_points(rng, 300)
    out = F.sample(G, f, t, VIEW, 0, p)
    # float64 renormalised trilinear over fluid corners
    g = (p.astype(np.float64) - np.array(ORIGIN)) / H - 0.5
    i0 = np.floor(g).astype(int)
    fr = g - i0
    acc = np.zeros((len(p), 5))
    ws = np.zeros(len(p))
    for k in range(8):
        o = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
        c = i0 + o
        w = np.prod(np.where(o == 1, fr, 1 - fr), axis=1)
        idx = c[:, 0] + DIMS[0] * (c[:, 1] + DIMS[1] * c[:, 2])
        v = f[idx]
        ok = v[:, 3] > 0
        acc[ok, :4] += v[ok] * w[ok, None]
        acc[ok, 4] += t[idx][ok] * w[ok]
        ws[ok] += w[ok]
    live = ws >= 1e-5
    want = acc[live] / ws[live, None]
    np.testing.assert_allclose(out[live, :5], want, rtol=0, atol=2e-5 * max(1.0, np.abs(want).max()))
    np.testing.assert_allclose(out[live, 5], ws[live], atol=1e-6)
    dead = out[~live]
    assert np.all(dead[:, :6] == [0, 0, 0, 1, 0, 0])


def test_all_solid_and_outside_samples():
    f = field_of(0.01, 0.0, 0.0, 0.0)  # every cell solid
    t = np.ones(G.n, np.float32)
    out = F.sample(G, f, t, VIEW, 0, np.array([[0.5, 0.5, 0.5]], np.float32))
    assert np.array_equal(out[0, :6], [0, 0, 0, 1, 0, 0])
    # far outside clamps to the boundary cells: the sample at a boundary cell centre
    X, Y, Z = centres()
    f = field_of(0.01 * X + 0.02 * Y, 0.0 * X, 0.0 * X, 1.0 + 0.0 * X)
    p = np.array([[-100.0, ORIGIN[1] + 3.5 * H, ORIGIN[2] + 2.5 * H]], np.float32)
    q = np.array([[ORIGIN[0] + 0.5 * H, ORIGIN[1] + 3.5 * H, ORIGIN[2] + 2.5 * H]], np.float32)
    assert np.allclose(F.sample(G, f, t, VIEW, 0, p)[0, :6], F.sample(G, f, t, VIEW, 0, q)[0, :6], atol=1e-7)


# ---- vorticityAt and fieldValue (wgsl.ts:356-363,430-457) ----

def test_rigid_rotation_vorticity():
    """u = w x (r - c) about z: curl = 2 w, and vorticityAt = |curl| h / 2 * 2 = 2 w h."""
    rng = np.random.default_rng(4)
    X, Y, Z = centres()
    w = 0.004
    xc, yc = 1.5, 2.0
    f = field_of(-w * (Y - yc), w * (X - xc), 0.0 * X, 1.0 + 0.0 * X)
    t = np.zeros(G.n, np.float32)
    lo = np.array(ORIGIN) + 2.0 * H
    hi = np.array(ORIGIN) + (np.array(DIMS) - 2.0) * H
    p = rng.uniform(lo, hi, (100, 3)).astype(np.float32)
    out = F.sample(G, f, t, VIEW, 3, p)
    np.testing.assert_allclose(out[:, 6], 2 * w * H, rtol=2e-4)
    np.testing.assert_allclose(out[:, 7], out[:, 6] / (VIEW.u_in * VIEW.vort_max), rtol=1e-6)


def test_field_values_on_a_uniform_flow():
    f = field_of(0.03, 0.04, 0.0, 1.0)
    t = np.full(G.n, 12.0, np.float32)
    p = np.array([[0.5, 1.0, 1.0]], np.float32)
    v = VIEW
    speed = F.sample(G, f, t, v, 0, p)[0, 7]
    assert speed == pytest.approx(0.05 / (v.u_in * v.speed_max), rel=1e-6)
    cp = F.sample(G, f, t, v, 1, p)[0, 7]   # rho equals the reference density: cp = 0
    assert cp == pytest.approx((0 - v.cp_min) / (v.cp_max - v.cp_min), rel=1e-6)
    temp = F.sample(G, f, t, v, 2, p)[0, 7]
    assert temp == pytest.approx((12.0 - v.t_min) / (v.t_max - v.t_min), rel=1e-6)
    wake = F.sample(G, f, t, v, 4, p)[0, 7]  # u_x / u_in = 0.6
    assert wake == pytest.approx((1 - 0.6) / 0.9, rel=1e-5)


# ---- particles (wgsl.ts:547-594) ----

def psim(count, **kw):
    ps = PSim()
    ps.count = count
    ps.spawn_mode = kw.get("spawn_mode", 1)
    ps.color_mode = kw.get("color_mode", 0)
    ps.frame = kw.get("frame", 1)
    ps.dt = kw.get("dt", 2.0)
    ps.life = kw.get("life", 900.0)
    ps.scatter = kw.get("scatter", 0)
    ps.rake_rows, ps.rake_cols = 12, 7
    ps.spawn_min[:] = kw.get("lo", (ORIGIN[0] + 2 * H, 0.3, 0.0))
    ps.spawn_max[:] = kw.get("hi", (0.0, 3.5, 2.0))
    return ps


def test_particles_advect_by_the_closed_form_step():
    U = 0.05
    f = field_of(U, 0.0, 0.0, 1.0)
    t = np.zeros(G.n, np.float32)
    rng = np.random.default_rng(5)
    n = 64
    parts = np.zeros((n, 8), np.float32)
    parts[:, :3] = rng.uniform((-0.5, 0.8, 0.2), (2.0, 3.0, 2.5), (n, 3))
    parts[:, 3] = 10.0
    out = F.particles_step(G, f, t, psim(n, dt=4.0), VIEW, parts)
    d = U * H * 4.0
    np.testing.assert_allclose(out[:, 0], parts[:, 0] + d, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(out[:, 1:3], parts[:, 1:3])
    assert np.all(out[:, 3] == 11.0)
    np.testing.assert_allclose(out[:, 4], d, rtol=1e-6)
    np.testing.assert_allclose(out[:, 7], U / (VIEW.u_in * VIEW.speed_max), rtol=1e-6)
    # a step longer than 2h is capped at 2h (2.5h and 5h long here)
    for dt in (50.0, 100.0):
        out = F.particles_step(G, f, t, psim(n, dt=dt), VIEW, parts)
        np.testing.assert_allclose(out[:, 4], 2 * H, rtol=1e-6)


def test_particles_respawn_at_the_outlet_and_when_old():
    f = field_of(0.05, 0.0, 0.0, 1.0)
    t = np.zeros(G.n, np.float32)
    hx = ORIGIN[0] + DIMS[0] * H
    parts = np.zeros((4, 8), np.float32)
    parts[:, :3] = [[hx - 1.5 * H - 0.01, 1.0, 1.0]] * 4   # crosses hx - 1.5h this step
    parts[:, 3] = 3.0
    parts[3, :3] = [0.0, 1.0, 1.0]
    parts[3, 3] = 5000.0                                     # beyond any maxLife: respawned, then advected
    ps = psim(4, dt=4.0)
    out = F.particles_step(G, f, t, ps, VIEW, parts)
    for i in range(3):
        assert out[i, 3] == 0.0 and np.array_equal(out[i, 4:], [0, 0, 0, -1])
        lane = i % 84
        row, col = lane // 7, lane % 7
        y0 = ps.spawn_min[1] * (1 - (row + 0.5) / 12) + ps.spawn_max[1] * ((row + 0.5) / 12)
        assert ps.spawn_min[0] <= out[i, 0] <= ps.spawn_min[0] + 0.12
        assert abs(out[i, 1] - y0) <= 0.0061
    assert out[3, 3] == 1.0 and out[3, 4] == pytest.approx(0.05 * H * 4.0, rel=1e-6)


def test_scatter_and_fog_spawn_regions():
    f = field_of(0.0, 0.0, 0.0, 1.0)  # still air: positions stay where spawned
    t = np.zeros(G.n, np.float32)
    parts = np.zeros((500, 8), np.float32)
    hx = ORIGIN[0] + DIMS[0] * H
    out = F.particles_step(G, f, t, psim(500, scatter=1, spawn_mode=0), VIEW, parts)
    lo = (ORIGIN[0] + 2 * H, 0.3, 0.0)
    assert np.all(out[:, 0] >= lo[0] - 1e-6) and np.all(out[:, 0] <= hx - 2 * H + 1e-6)
    assert np.all((out[:, 1] >= 0.3) & (out[:, 1] <= 3.5)) and np.all((out[:, 2] >= 0.0) & (out[:, 2] <= 2.0))
    assert np.all(out[:, 3] == 1.0)
    assert np.ptp(out[:, 0]) > 0.5 * (hx - 2 * H - lo[0])


# ---- streamlines (render-wgsl.ts:536-575) ----

def stream(lines, steps, **kw):
    st = Stream()
    st.lines, st.steps = lines, steps
    st.seed_mode = kw.get("seed_mode", 0)
    st.color_mode = kw.get("color_mode", 0)
    st.seed_min[:] = kw.get("lo", (ORIGIN[0] + 1.1 * H, 1.0, 0.6))
    st.seed_max[:] = kw.get("hi", (0.0, 3.0, 2.4))
    st.rows, st.cols = kw.get("rows", 3), kw.get("cols", 2)
    st.step_len = kw.get("step_len", 0.13)
    st.time = kw.get("time", 0.0)
    st.per_emitter, st.emitters = kw.get("per", 1), kw.get("emitters", 1)
    return st


def test_streamlines_in_a_uniform_flow():
    f = field_of(0.05, 0.0, 0.0, 1.0)
    t = np.zeros(G.n, np.float32)
    st = stream(6, 80)
    pts = F.streamlines(G, f, t, st, VIEW)
    hx = ORIGIN[0] + DIMS[0] * H
    x0 = st.seed_min[0]
    live_want = math.floor((hx - 3 * H - x0) / st.step_len) + 1
    for line in pts:
        live = line[:, 3] >= 0
        assert live.sum() == live_want
        k = np.arange(live_want)
        np.testing.assert_allclose(line[:live_want, 0], x0 + k * st.step_len, atol=1e-5)
        assert np.all(line[live_want:, 3] == -1.0)
        assert np.all(line[live_want:, 0] == line[live_want, 0])
    # rake seeds (render-wgsl.ts:545-549)
    fy = (np.arange(3) + 0.5) / 3
    np.testing.assert_allclose(pts[::2, 0, 1], st.seed_min[1] * (1 - fy) + st.seed_max[1] * fy, rtol=1e-6)


def test_emitter_ring_seeds_and_stagnant_flow():
    f = field_of(0.0, 0.0, 0.0, 1.0)  # no flow: every line dies at its seed
    t = np.zeros(G.n, np.float32)
    em = np.array([[1.0, 2.0, 1.5, 0.2], [2.0, 1.5, 1.0, 0.1]], np.float32)
    st = stream(12, 5, seed_mode=1, per=6, emitters=2, time=0.7)
    pts = F.streamlines(G, f, t, st, VIEW, em)
    assert np.all(pts[:, :, 3] == -1.0)
    for line in range(12):
        e = em[min(line // 6, 1)]
        k = line % 6
        ring = 0.35 + 0.65 * (k % 3) / 2.0
        r = math.hypot(pts[line, 0, 1] - e[1], pts[line, 0, 2] - e[2])
        assert pts[line, 0, 0] == e[0]
        assert r == pytest.approx(e[3] * ring, rel=1e-5)
        a = ((k + 0.5) / 6) * 6.2831853 + 0.7 * 0.6
        assert math.atan2(pts[line, 0, 2] - e[2], pts[line, 0, 1] - e[1]) == pytest.approx(
            math.atan2(math.sin(a), math.cos(a)), abs=1e-4)
