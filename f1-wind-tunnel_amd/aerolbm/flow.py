"""Host side of the field consumers: the uniforms the reference's AeroEngine fills
for its particle and streamline dispatches, and one frame's flow work.

Reference: src/engine/engine.ts:14-28 (particle counts, MAX_LINES, STREAM_STEPS),
:120-141 (f1Anchors, w17Place), :391-393 (visualFlowRate), :474-488 (shapeBounds),
:503-519 (viewUniform), :526-572 (emitterPoints, streamConfig), :588-642 (the
flow part of frame) and :731-737 (spawnRegion).  The kernels are in
csrc/flow.hip behind aero_particles_step / aero_streamlines.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from ._lib import PSim, Stream, View
from .engine import ORIGIN_X, ShapeKind, ShapeParams

# engine.ts:14-28
PARTICLES = {"low": 40_000, "medium": 80_000, "high": 140_000}
MAX_LINES = 160
STREAM_STEPS = 360
PARTICLE_LIFE = 900.0


class FlowMode:
    """engine.ts:34."""
    Off, Streamlines, Smoke, Fog, Vortices = 0, 1, 2, 3, 4


class ColorField:
    """fieldValue ids (wgsl.ts:441-457); Vorticity colours particles by speed
    (engine.ts:588)."""
    Speed, Pressure, Temperature, Vorticity, Wake = 0, 1, 2, 3, 4


# car-model.ts:15-20
FRONT_AXLE_X = 1.0
BASE_RIDE_H = 0.092
REAR_AXLE_X = 4.4


def is_race_car(kind: int) -> bool:
    return kind in (ShapeKind.F1, ShapeKind.W17)


def visual_flow_rate(kmh: float = 250.0) -> float:
    """engine.ts:391-393."""
    return min(1.6, max(0.35, kmh / 250.0))


def view_uniform(u_lb: float, kmh: float, p: Optional[ShapeParams] = None) -> View:
    """viewUniform (engine.ts:503-519)."""
    v_real = kmh / 3.6
    stag = 0.5 * v_real * v_real / 1005
    hot = p is not None and is_race_car(p.kind) and bool(p.heatBrakes or p.heatEngine or p.heatRad)
    return View(u_lb, v_real, -1.1, 0.7, 0.0 if hot else -stag * 1.4, 220.0 if hot else stag * 1.05, 1.5, 0.6)


def shape_bounds(p: ShapeParams) -> Tuple[Tuple[float, float, float], Tuple[float, float, float]]:
    """shapeBounds (engine.ts:474-488) for the built-in kinds; custom profiles use
    unit bounds (the profile metadata is the caller's)."""
    k = p.kind
    if k == ShapeKind.F1:
        return (-0.08, 0.0, -0.98), (5.46, p.rideH + 1.34, 0.98)
    if k == ShapeKind.W17:
        return (-0.3, 0.0, -0.98), (5.2, p.rideH + 1.16, 0.98)
    if k == ShapeKind.Empty:
        return (2.6, 0.5, -0.05), (2.7, 0.6, 0.05)
    if k == ShapeKind.Sphere:
        return (1.95, 0.35, -0.65), (3.25, 1.65, 0.65)
    if k == ShapeKind.Cube:
        return (2.05, 0.45, -0.55), (3.15, 1.55, 0.55)
    if k == ShapeKind.Wing:
        return (1.3, 0.2, -1.3), (3.6, 2.2, 1.3)
    zr = p.customWidth
    return (-0.05, 0.0, -zr - 0.05), (1.05, 1.05, zr + 0.05)


def spawn_region(p: ShapeParams, h: float) -> dict:
    """spawnRegion (engine.ts:731-737)."""
    lo, hi = shape_bounds(p)
    z_max = max(0.3, hi[2] + 0.25)
    y_max = max(0.5, hi[1] + 0.25)
    y_min = 0.04 if is_race_car(p.kind) else max(0.04, lo[1] - 0.25)
    return {"min": (ORIGIN_X + h * 2, y_min, 0.0), "max": (0.0, y_max, z_max), "rows": 12, "cols": 7}


def psim_uniform(count: int, mode: int, color: int, frame: int, steps: int, h: float, p: ShapeParams,
                 kmh: float = 250.0, scatter: bool = False) -> PSim:
    """The PSim of frame() (engine.ts:595-611): dt = evenSteps * visualFlowRate."""
    even = max(2, steps - steps % 2)
    sp = spawn_region(p, h)
    color_mode = 0 if color == ColorField.Vorticity else color
    ps = PSim()
    ps.count = count
    ps.spawn_mode = 0 if mode == FlowMode.Fog else 1
    ps.color_mode = color_mode
    ps.frame = frame
    ps.dt = even * visual_flow_rate(kmh)
    ps.life = PARTICLE_LIFE
    ps.scatter = 1 if scatter else 0
    ps.rake_rows, ps.rake_cols = sp["rows"], sp["cols"]
    ps.spawn_min[:] = sp["min"]
    ps.spawn_max[:] = sp["max"]
    return ps


def f1_anchor_heights(ride_h: float) -> Tuple[float, float]:
    """yw, yr of f1Anchors (engine.ts:120-133)."""
    return 0.19 + ride_h * 0.45, ride_h + 0.8


def w17_place(p: ShapeParams, x: float, y: float, z: float) -> Tuple[float, float, float]:
    """w17Place (engine.ts:136-141): a body-frame point moved with ride height and rake."""
    rake = math.atan((p.rearRideH - p.rideH) / (REAR_AXLE_X - FRONT_AXLE_X))
    c, s = math.cos(rake), math.sin(rake)
    bx, by = x - FRONT_AXLE_X, y - BASE_RIDE_H
    return FRONT_AXLE_X + c * bx - s * by, p.rideH + s * bx + c * by, z


def emitter_points(p: ShapeParams) -> Tuple[np.ndarray, int]:
    """emitterPoints (engine.ts:526-556): 8 x (x, y, z, r), and how many are used."""
    if p.kind == ShapeKind.W17:
        def place(x, y, z, r):
            return [*w17_place(p, x, y, z), r]
        pts = []
        if p.frontWingOn:
            pts += [place(0.62, 0.44, 0.69, 0.05), place(0.5, 0.27, 0.2, 0.04)]
        if p.rearWingOn:
            pts += [place(5.13, 0.96, 0.57, 0.05)]
        pts += [place(2.4, 0.16, 0.86, 0.05)]
    elif is_race_car(p.kind):
        yw, yr = f1_anchor_heights(p.rideH)
        pts = [[0.52, yw + 0.2, 0.926, 0.06], [0.52, yw + 0.07, 0.21, 0.05], [5.36, yr + 0.26, 0.494, 0.06],
               [2.22, p.rideH + 0.06, 0.79, 0.05]]
    else:
        lo, hi = shape_bounds(p)
        pts = [[hi[0], hi[1] - 0.05, hi[2] - 0.05, 0.08], [hi[0], lo[1] + 0.05, hi[2] - 0.05, 0.08],
               [hi[0], (lo[1] + hi[1]) / 2, 0.05, 0.1]]
    e = np.zeros((8, 4), np.float32)
    e[:len(pts)] = np.asarray(pts, np.float32)
    return e, len(pts)


def stream_uniform(p: ShapeParams, mode: int, color: int, time: float) -> Tuple[Stream, np.ndarray]:
    """streamConfig + the Stream of frame() (engine.ts:558-572,633-641)."""
    st = Stream()
    st.steps = STREAM_STEPS
    st.color_mode = 0 if color == ColorField.Vorticity else color
    st.time = time
    em, count = emitter_points(p)
    if mode == FlowMode.Vortices:
        per = MAX_LINES // count
        st.lines, st.seed_mode, st.per_emitter, st.emitters = per * count, 1, per, count
        st.step_len, st.rows, st.cols = 0.022, 1, 1
        return st, em
    lo, hi = shape_bounds(p)
    f1 = is_race_car(p.kind)
    rows, cols = 18, 5
    st.lines, st.seed_mode, st.per_emitter, st.emitters = rows * cols, 0, 1, 1
    st.step_len, st.rows, st.cols = 0.03, rows, cols
    st.seed_min[:] = (-1.2, 0.05 if f1 else max(0.04, lo[1] - 0.2), 0.03)
    st.seed_max[:] = (0.0, 1.3 if f1 else hi[1] + 0.2, 0.95 if f1 else hi[2] + 0.2)
    return st, em


@dataclass
class FlowFrame:
    """The flow part of frame() (engine.ts:588-642,713-714) on one engine: particles
    for Smoke / Fog, streamlines for Streamlines / Vortices."""
    engine: object
    shape: ShapeParams
    quality: str = "high"
    kmh: float = 250.0
    frame_no: int = 0
    flow_time: float = 0.0
    scatter_frames: int = 2

    def __post_init__(self):
        self.count = PARTICLES[self.quality]
        self.engine.particles_alloc(self.count)

    def run(self, mode: int, color: int, steps: int, dt: float = 1 / 60) -> Optional[np.ndarray]:
        """One frame's flow dispatch after `steps` LBM steps; returns the streamline
        points for the streamline modes."""
        self.frame_no += 1
        self.flow_time += dt * visual_flow_rate(self.kmh)
        view = view_uniform(self.engine.u_lb, self.kmh, self.shape)
        out = None
        if mode in (FlowMode.Smoke, FlowMode.Fog):
            ps = psim_uniform(self.count, mode, color, self.frame_no, steps, self.engine.h, self.shape, self.kmh,
                              scatter=self.scatter_frames > 0)
            self.engine.particles_step(ps, view)
            if self.scatter_frames > 0:
                self.scatter_frames -= 1
        if mode in (FlowMode.Streamlines, FlowMode.Vortices):
            st, em = stream_uniform(self.shape, mode, color, self.flow_time)
            out = self.engine.streamlines(st, view, em)
        return out


def lines_alive(pts: np.ndarray) -> List[int]:
    """Number of live points per streamline (value >= 0)."""
    return [int((line[:, 3] >= 0).sum()) for line in pts]
