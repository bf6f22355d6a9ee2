"""Named tunnel workloads (grids, lattice parameters and geometry).

  cube        1 m cube (22 cells per side) in a 128x64x64 half tunnel, h = 1/22 m,
              u = 0.05, tau0 = 0.51, Cs = 0.1; Cd averaged over steps 1001-2000.
  f1          generic F1 body at 4.5 cm cells in a 320x96x96 half tunnel: rounded-box
              chassis 5.0 x 1.9 x 0.95 m, wheels r = 0.36 m w = 0.30 m, two-element
              front and rear wings at seeded angles in [4, 14] deg, rake 30 mm front /
              80 mm rear, six seeded heat sources.
  aeromap     the cube of "cube" raked over a 5x5 ride-height grid (front 20-60 mm x
              rear 40-120 mm), plus 8 uneven tunnels from 160x48x48 to 256x80x80.
  large       the generic F1 body at 2 cm cells, 768x224x192, tau0 = 0.5005, Cs = 0.12.

The cube sits where the reference's cube kind puts it (sceneSdf kind 2,
src/engine/wgsl.ts:105: centre (2.6, 1.0, 0), half size 0.5 m, corner radius 0.03);
the grid origin places its centre at cell x = 42.  The ride-height map applies the
reference's rake rotation about the front axle (src/engine/f1.ts:37-46) to the cube,
whose base sits at the front ride height.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from ._lib import Prim, Shape
from .engine import PROFILE_H, PROFILE_W, ShapeKind

DEG = math.pi / 180

PRIM_ROUNDBOX, PRIM_WHEEL, PRIM_WING, PRIM_SPHERE, PRIM_CAPSULE, PRIM_CYLZ, PRIM_EMITTER = 1, 2, 3, 4, 5, 6, 7
MAT_REST, MAT_FRONT_WING, MAT_REAR_WING = 0, 7, 8


def prim(type_: int, material: int, frame: int, *a: float) -> Prim:
    p = Prim()
    p.type, p.material, p.frame = type_, material, frame
    for i, v in enumerate(a):
        p.a[i] = float(v)
    return p


def base_shape(kind: int, ride_h: float = 0.10, rear_ride: float = 0.12) -> Shape:
    s = Shape()
    s.kind = kind
    s.pw, s.ph = PROFILE_W, PROFILE_H
    s.p_scale = 0.02
    s.ride_h = ride_h
    s.rear_ride = rear_ride
    s.aoa = 6 * DEG
    s.duct = 1.0
    return s


@dataclass
class Workload:
    name: str
    dims: Tuple[int, int, int]
    h: float
    origin: Tuple[float, float, float]
    u_lb: float
    tau0: float
    smag: float
    steps: int
    shape: Shape
    prims: List[Prim] = field(default_factory=list)
    avg_from: int = 0          # force averaging starts after this many steps
    frontal_area: float = 1.0  # full-body reference area for Cd (m^2)

    @property
    def cells(self) -> int:
        return self.dims[0] * self.dims[1] * self.dims[2]

    def engine_kwargs(self) -> dict:
        return dict(dims=self.dims, h=self.h, origin=self.origin, u_lb=self.u_lb, tau0=self.tau0, smag=self.smag)


def cube(steps: int = 2000) -> Workload:
    h = np.float32(1.0 / 22.0)
    ox = float(np.float32(2.6 - 42.0 / 22.0))
    return Workload("cube", (128, 64, 64), float(h), (ox, 0.0, 0.0), 0.05, 0.51, 0.1, steps,
                    base_shape(ShapeKind.Cube), avg_from=steps // 2, frontal_area=1.0)


def generic_f1_prims(seed: int = 0, ride_front: float = 0.03) -> List[Prim]:
    """Generic F1 body in the raked body frame (rake applied from the Shape uniform)."""
    rng = np.random.default_rng(seed)
    a_fm, a_ff, a_rm, a_rf = rng.uniform(4.0, 14.0, 4) * DEG
    b = ride_front
    P: List[Prim] = []
    # chassis: 5.0 (x) x 0.95 (y) x 1.9 (z, full width), base at the ride height
    P.append(prim(PRIM_ROUNDBOX, MAT_REST, 1, 2.95, b + 0.475, 0.0, 2.5, 0.475, 0.95, 0.08))
    # wheels r = 0.36, w = 0.30, standing on the ground, outboard of the chassis
    P.append(prim(PRIM_WHEEL, MAT_REST, 0, 1.0, 0.36, 1.12, 0.36, 0.15))
    P.append(prim(PRIM_WHEEL, MAT_REST, 0, 4.4, 0.36, 1.12, 0.36, 0.15))
    # front wing: main plane + flap (thin NACA-like plates), span |z| <= 0.9
    le_m = (-0.40, b + 0.09)
    c_m = 0.32
    P.append(prim(PRIM_WING, MAT_FRONT_WING, 1, le_m[0], le_m[1], c_m, a_fm, 0.028, -0.9, 0.9))
    te = (le_m[0] + math.cos(a_fm) * c_m, le_m[1] + math.sin(a_fm) * c_m)
    P.append(prim(PRIM_WING, MAT_FRONT_WING, 1, te[0] - 0.03, te[1] + 0.025, 0.20, a_fm + a_ff, 0.02, -0.9, 0.9))
    # rear wing: main plane + flap above the rear of the chassis, span |z| <= 0.5
    le_r = (5.0, b + 1.12)
    c_r = 0.36
    P.append(prim(PRIM_WING, MAT_REAR_WING, 1, le_r[0], le_r[1], c_r, a_rm, 0.04, -0.5, 0.5))
    te = (le_r[0] + math.cos(a_rm) * c_r, le_r[1] + math.sin(a_rm) * c_r)
    P.append(prim(PRIM_WING, MAT_REAR_WING, 1, te[0] - 0.03, te[1] + 0.03, 0.24, a_rm + a_rf, 0.03, -0.5, 0.5))
    # rear-wing pillar (body)
    P.append(prim(PRIM_ROUNDBOX, MAT_REST, 1, 5.15, b + 1.05, 0.0, 0.06, 0.12, 0.02, 0.01))
    # six seeded heat sources on the chassis skin, values in the reference's range
    for k in range(6):
        x = rng.uniform(0.8, 5.2)
        on_top = k % 2 == 0
        y = b + 0.95 if on_top else rng.uniform(b + 0.2, b + 0.8)
        z = rng.uniform(0.0, 0.8) if on_top else 0.95
        P.append(prim(PRIM_EMITTER, MAT_REST, 1, x, y, z, 0.15, rng.uniform(75.0, 620.0)))
    return P


def f1(steps: int = 10_000, seed: int = 0) -> Workload:
    s = base_shape(ShapeKind.Scene, ride_h=0.03, rear_ride=0.08)
    h = float(np.float32(0.045))
    return Workload("f1", (320, 96, 96), h, (-2.4, 0.0, 0.0), 0.05, 0.51, 0.1, steps, s,
                    generic_f1_prims(seed, 0.03), avg_from=steps // 2, frontal_area=1.9 * 0.95)


def large(steps: int = 50_000, seed: int = 0) -> Workload:
    s = base_shape(ShapeKind.Scene, ride_h=0.03, rear_ride=0.08)
    h = float(np.float32(0.02))
    return Workload("large", (768, 224, 192), h, (-2.4, 0.0, 0.0), 0.05, 0.5005, 0.12, steps, s,
                    generic_f1_prims(seed, 0.03), avg_from=steps // 2, frontal_area=1.9 * 0.95)


AEROMAP_FRONT = (0.020, 0.030, 0.040, 0.050, 0.060)
AEROMAP_REAR = (0.040, 0.060, 0.080, 0.100, 0.120)
UNEVEN_DIMS = ((160, 48, 48), (176, 52, 52), (184, 56, 56), (200, 60, 60),
               (216, 68, 68), (232, 72, 72), (240, 76, 76), (256, 80, 80))


def raked_cube_prims(ride_front: float) -> List[Prim]:
    # The cube of sceneSdf kind 2 in the raked body frame, base at the front ride height.
    return [prim(PRIM_ROUNDBOX, MAT_REST, 1, 2.6, ride_front + 0.5, 0.0, 0.5, 0.5, 0.5, 0.03)]


def aeromap(steps: int = 2000, uneven: bool = True) -> List[Workload]:
    base = cube(steps)
    out: List[Workload] = []
    for fr in AEROMAP_FRONT:
        for rr in AEROMAP_REAR:
            s = base_shape(ShapeKind.Scene, ride_h=fr, rear_ride=rr)
            out.append(Workload(f"aeromap_{int(fr*1000)}_{int(rr*1000)}", base.dims, base.h, base.origin,
                                base.u_lb, base.tau0, base.smag, steps, s, raked_cube_prims(fr),
                                avg_from=steps // 2, frontal_area=1.0))
    if uneven:
        for d in UNEVEN_DIMS:
            out.append(Workload(f"uneven_{d[0]}x{d[1]}x{d[2]}", d, base.h, base.origin, base.u_lb, base.tau0,
                                base.smag, steps, base_shape(ShapeKind.Cube), avg_from=steps // 2))
    return out


def cd_from_forces(sum_fx_lattice: float, steps: int, h: float, u_lb: float, area: float) -> float:
    """Cd from the half-body lattice x-force sum: Cd*A = 4 F h^2 / u^2 with F the mean
    per-step force (engine.ts:752 divided by q = rho v^2 / 2)."""
    f = sum_fx_lattice / steps
    return 4.0 * f * h * h / (u_lb * u_lb) / area
