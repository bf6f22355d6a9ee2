"""aerolbm — host side of the MI355X wind-tunnel solver (libaero_lbm.so).

Mirrors the simulation surface of the reference's AeroEngine
(src/engine/engine.ts:159-769) over the C ABI declared in include/aero_lbm.h.
"""
from ._lib import AeroError, Grid, Heat, Prim, Rig, Shape, Sim, load  # noqa: F401
from .engine import CustomProfile, Forces, ForcePair, HipAeroEngine, ShapeKind, ShapeParams, shape_uniform  # noqa: F401
from . import workloads  # noqa: F401
