"""GPU parity of the field consumers (csrc/flow.hip, through the C ABI) with the
f32 oracle (oracle/flow_oracle.c) on the same fields and temperatures: probes,
particle dispatches, rake and emitter-ring streamlines bit-identical; closed-form
motion in a uniform flow at the F1 grid size; argument errors."""
import numpy as np
import pytest

import aerolbm
from aerolbm import flow as FL
from aerolbm import workloads as W
from aerolbm._lib import AeroError, PSim, Stream, View
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FO = orc.FlowOracle()


def scene_engine(dims=(64, 24, 16), steps=40):
    """A small tunnel with a body, a front and a rear wing and a heat source, run a
    few dozen LBM and heat steps so that fields and temperatures are non-trivial."""
    e = aerolbm.HipAeroEngine(dims=dims, h=0.06, origin=(-0.4, 0.0, 0.0), u_lb=0.06, tau0=0.51, smag=0.1,
                              storage="f32", math="fast")
    s = W.base_shape(aerolbm.ShapeKind.Scene, ride_h=0.03, rear_ride=0.06)
    P = W.prim
    prims = [P(W.PRIM_ROUNDBOX, W.MAT_REST, 1, 1.2, 0.35, 0.0, 0.6, 0.2, 0.35, 0.05),
             P(W.PRIM_WING, W.MAT_FRONT_WING, 1, 0.2, 0.12, 0.3, 8 * W.DEG, 0.04, -0.4, 0.4),
             P(W.PRIM_WING, W.MAT_REAR_WING, 1, 1.9, 0.7, 0.3, 9 * W.DEG, 0.05, -0.3, 0.3),
             P(W.PRIM_EMITTER, W.MAT_REST, 1, 1.0, 0.55, 0.2, 0.1, 300.0)]
    e.set_shape(s, prims=prims)
    e.reset()
    for _ in range(steps // 8):
        e.step(8)
        e.heat_step(2)
    return e


def ogrid(e):
    return orc.grid(e.dims, e.h, e.origin)


def view_for(e):
    return View(e.u_lb, 69.4, -1.1, 0.7, -3.0, 2.5, 1.5, 0.6)


def probe_points(e, rng, k=3000):
    lo = np.array(e.origin) - 2 * e.h
    hi = np.array(e.origin) + (np.array(e.dims) + 2) * e.h
    p = rng.uniform(lo, hi, (k, 3)).astype(np.float32)
    p[: k // 4, 2] *= -1.0  # mirror half
    return p


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_probes_match_oracle_bit_for_bit(mode):
    print(f"flow probes mode={mode}", flush=True)
    rng = np.random.default_rng(20 + mode)
    e = scene_engine()
    f, t = e.fields(), e.temps()
    assert (f[:, 3] <= 0).any() and np.abs(f[:, 1]).max() > 0 and t.max() > 0
    pts = probe_points(e, rng)
    v = view_for(e)
    got = e.sample_flow(pts, v, mode)
    want = FO.sample(ogrid(e), f, t, v, mode, pts)
    assert np.array_equal(got, want, equal_nan=True)
    e.close()


@pytest.mark.parametrize("spawn_mode,color", [(1, 0), (0, 1), (1, 2), (0, 4), (1, 3)])
def test_particle_frames_match_oracle_bit_for_bit(spawn_mode, color):
    print(f"particles spawn={spawn_mode} colour={color}", flush=True)
    rng = np.random.default_rng(30 + color)
    e = scene_engine()
    f, t = e.fields(), e.temps()
    n = 4000
    e.particles_alloc(n + 17)
    parts = np.zeros((n + 17, 8), np.float32)
    lo = np.array(e.origin)
    hi = lo + np.array(e.dims) * e.h
    parts[:, :3] = rng.uniform(lo - 0.1, hi + 0.1, (n + 17, 3))
    parts[:, 3] = rng.uniform(0, 1200, n + 17)  # some beyond maxLife
    e.write_particles(parts)
    v = view_for(e)
    g = ogrid(e)
    state = parts
    for frame in range(1, 5):
        ps = PSim()
        ps.count, ps.spawn_mode, ps.color_mode, ps.frame = n, spawn_mode, color, frame
        ps.dt, ps.life, ps.scatter = 6.0, 900.0, 1 if frame == 1 else 0
        ps.rake_rows, ps.rake_cols = 12, 7
        ps.spawn_min[:] = (e.origin[0] + 2 * e.h, 0.04, 0.0)
        ps.spawn_max[:] = (0.0, 0.9, 0.6)
        e.particles_step(ps, v)
        state = FO.particles_step(g, f, t, ps, v, state)
        got = e.particles()
        assert np.array_equal(got, state, equal_nan=True), frame
        assert np.array_equal(got[n:], parts[n:])  # beyond count: untouched
    e.close()


def test_rake_streamlines_match_oracle_bit_for_bit():
    print("streamlines rake", flush=True)
    e = scene_engine()
    f, t = e.fields(), e.temps()
    v = view_for(e)
    for color in (0, 1, 3):
        st = Stream()
        st.lines, st.steps, st.seed_mode, st.color_mode = 90, 360, 0, color
        st.seed_min[:] = (-0.35, 0.02, 0.03)
        st.seed_max[:] = (0.0, 1.2, 0.8)
        st.rows, st.cols, st.step_len, st.per_emitter, st.emitters = 18, 5, 0.03, 1, 1
        got = e.streamlines(st, v)
        want = FO.streamlines(ogrid(e), f, t, st, v)
        assert np.array_equal(got, want, equal_nan=True), color
        assert (got[:, :, 3] >= 0).sum() > 90 * 20  # lines do travel
    e.close()


def test_emitter_streamlines_match_oracle():
    """Seeds on emitter rings (the ring angle's sin / cos evaluated in FP64 on both
    sides): the lines are bit-identical to the oracle's."""
    print("streamlines emitters", flush=True)
    e = scene_engine()
    f, t = e.fields(), e.temps()
    v = view_for(e)
    em = np.array([[0.3, 0.25, 0.3, 0.06], [1.0, 0.5, 0.35, 0.05], [2.0, 0.7, 0.2, 0.06]], np.float32)
    st = Stream()
    st.lines, st.steps, st.seed_mode, st.color_mode = 159, 360, 1, 0
    st.per_emitter, st.emitters, st.step_len, st.time, st.rows, st.cols = 53, 3, 0.022, 1.7, 1, 1
    got = e.streamlines(st, v, em)
    want = FO.streamlines(ogrid(e), f, t, st, v, em)
    assert np.array_equal(got, want, equal_nan=True)
    assert (got[:, :, 3] >= 0).sum() > 159 * 20  # lines do travel
    e.close()


def test_uniform_flow_at_f1_size():
    """Size-independent check on the 320x96x96 grid: in uniform flow every particle
    advances by u h dt along x (or respawns past the outlet band) and every rake line
    marches at stepLen until 3h before the outlet."""
    print("flow uniform F1 size", flush=True)
    w = W.f1()
    e = aerolbm.HipAeroEngine(storage="f32", math="fast", **w.engine_kwargs())
    U = 0.05
    fl = np.zeros((e.n, 4), np.float32)
    fl[:, 0] = U
    fl[:, 3] = 1.0
    e.write_fields(fl)
    e.write_temps(np.zeros(e.n, np.float32))
    p = aerolbm.ShapeParams()
    n = FL.PARTICLES["high"]
    e.particles_alloc(n)
    rng = np.random.default_rng(40)
    lo = np.array(e.origin)
    hi = lo + np.array(e.dims) * e.h
    parts = np.zeros((n, 8), np.float32)
    parts[:, :3] = rng.uniform(lo + [0.1, 0.1, 0.0], hi - [1.0, 0.1, 0.1], (n, 3))
    parts[:, 3] = 1.0
    e.write_particles(parts)
    ps = FL.psim_uniform(n, FL.FlowMode.Smoke, 0, 5, 8, e.h, p)
    e.particles_step(ps, View(U, 69.4, -1.1, 0.7, -3, 2.5, 1.5, 0.6))
    out = e.particles()
    d = np.float32(U) * np.float32(e.h) * np.float32(ps.dt)
    np.testing.assert_allclose(out[:, 0], parts[:, 0] + d, rtol=0, atol=2e-6)
    assert np.all(out[:, 3] == 2.0)
    st, _ = FL.stream_uniform(p, FL.FlowMode.Streamlines, 0, 0.0)
    pts = e.streamlines(st, View(U, 69.4, -1.1, 0.7, -3, 2.5, 1.5, 0.6))
    hx = e.origin[0] + e.dims[0] * e.h
    live = min(st.steps, int(np.floor((hx - 3 * e.h - st.seed_min[0]) / st.step_len)) + 1)
    assert all(abs(k - live) <= 1 for k in FL.lines_alive(pts))
    np.testing.assert_allclose(pts[:, 1:, 0] - pts[:, :-1, 0], st.step_len, rtol=0, atol=1e-5)
    e.close()


def test_flow_argument_errors():
    e = aerolbm.HipAeroEngine(dims=(32, 16, 8), h=0.1, origin=(0.0, 0.0, 0.0))
    v = View(0.05, 69.4, -1.1, 0.7, -3, 2.5, 1.5, 0.6)
    e.particles_alloc(10)
    ps = PSim()
    ps.count, ps.spawn_mode, ps.rake_rows, ps.rake_cols = 11, 0, 1, 1
    with pytest.raises(AeroError):
        e.particles_step(ps, v)  # beyond the allocation
    ps.count, ps.spawn_mode, ps.rake_rows = 10, 1, 0
    with pytest.raises(AeroError):
        e.particles_step(ps, v)  # empty rake
    st = Stream()
    st.lines, st.steps, st.seed_mode, st.emitters, st.per_emitter = 4, 10, 1, 0, 1
    with pytest.raises(AeroError):
        e.streamlines(st, v, np.zeros((8, 4), np.float32))
    st.seed_mode, st.rows, st.cols = 0, 0, 1
    with pytest.raises(AeroError):
        e.streamlines(st, v)
    e.close()
