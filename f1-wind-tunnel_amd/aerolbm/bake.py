"""Host side of the W17 distance-grid baker: the mesh preparation of
scripts/bake-car.ts, with the distance grids computed on the GPU (aero_bake_sdf).

Reference: src/engine/glb.ts:33-130 (parseGlb: node walk, world transform stored as
f32), src/engine/car-model.ts:13-33 (CAR_LAYOUT, WHEEL_MATERIALS, toCarFrame) and
:36-46 (RIG), scripts/bake-car.ts:16-22 (primitives in the car frame), :237-316
(bakeGrid, here on the GPU), :318-358 (the body grid and the six part grids).

The per-vertex part tags are the baker's connected-component analysis of the wing
elements (bake-car.ts:24-160), which the reference ships as w17.rig.bin; they are an
input here.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import struct
from dataclasses import dataclass
from pathlib import Path
from typing import Callable, List, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import check, ptr

# car-model.ts:13-33
MESH_FRONT_AXLE_Z = 1.615
MESH_GROUND_Y = -0.04
FRONT_AXLE_X = 1.0
BODY_LAYOUT = {"origin": (-0.3, 0.0, 0.0), "h": 0.015, "dims": (368, 84, 67)}
WHEEL_MATERIALS = {"tread_medium.007", "sidewall.007", "mercedes_wheel_hub.007", "discs.007"}
# car-model.ts:37-46 (RIG)
PART_MASK = 7
ASSEMBLY_SHIFT = 3
PARTS = {"frontFlapA": 1, "frontFlapB": 2, "rearFlap": 3, "rearMain": 4, "frontStatic": 5, "rearStatic": 6}
ASSEMBLY = {"frontWing": 1, "rearWing": 2}
PART_GRID_H = 0.01
PART_MARGIN = 0.04

# bake-car.ts:318-336: (part, tag selector, closing radius)
PART_SPECS: List[Tuple[int, Callable[[int], bool], float]] = [
    (PARTS["frontFlapA"], lambda t: (t & PART_MASK) == PARTS["frontFlapA"], 0.012),
    (PARTS["frontFlapB"], lambda t: (t & PART_MASK) == PARTS["frontFlapB"], 0.012),
    (PARTS["rearFlap"], lambda t: (t & PART_MASK) == PARTS["rearFlap"], 0.012),
    (PARTS["rearMain"], lambda t: (t & PART_MASK) == PARTS["rearMain"], 0.012),
    (PARTS["frontStatic"], lambda t: (t & PART_MASK) == 0 and (t >> ASSEMBLY_SHIFT) == ASSEMBLY["frontWing"], 0.03),
    (PARTS["rearStatic"], lambda t: (t & PART_MASK) == 0 and (t >> ASSEMBLY_SHIFT) == ASSEMBLY["rearWing"], 0.03),
]
BODY_BAND, BODY_CLOSING, BODY_SHELL = 0.075, 0.06, 0.008
PART_SHELL = 0.004


@dataclass
class Prim:
    material: str
    verts: np.ndarray    # (n, 3) float64, car frame
    indices: np.ndarray  # (3m,) uint32
    wheel: bool


def _node_matrix(n: dict) -> np.ndarray:
    """nodeMatrix (glb.ts:35-49), column-major float64."""
    m = np.zeros(16, np.float64)
    if "matrix" in n:
        m[:] = n["matrix"]
        return m
    tx, ty, tz = n.get("translation", (0, 0, 0))
    x, y, z, w = n.get("rotation", (0, 0, 0, 1))
    sx, sy, sz = n.get("scale", (1, 1, 1))
    m[0] = (1 - 2 * (y * y + z * z)) * sx; m[1] = 2 * (x * y + z * w) * sx; m[2] = 2 * (x * z - y * w) * sx
    m[4] = 2 * (x * y - z * w) * sy; m[5] = (1 - 2 * (x * x + z * z)) * sy; m[6] = 2 * (y * z + x * w) * sy
    m[8] = 2 * (x * z + y * w) * sz; m[9] = 2 * (y * z - x * w) * sz; m[10] = (1 - 2 * (x * x + y * y)) * sz
    m[12] = tx; m[13] = ty; m[14] = tz; m[15] = 1
    return m


def _mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """mul (glb.ts:51-59): column-major a * b, each entry summed k = 0..3 in order."""
    o = np.zeros(16, np.float64)
    for c in range(4):
        for r in range(4):
            s = 0.0
            for k in range(4):
                s += float(a[k * 4 + r]) * float(b[c * 4 + k])
            o[c * 4 + r] = s
    return o


def parse_glb(path: Path):
    """(primitives in node-walk order, material names): each primitive's world
    positions stored as float32 and its indices (glb.ts:61-130)."""
    b = Path(path).read_bytes()
    magic, _, _ = struct.unpack("<III", b[:12])
    if magic != 0x46546C67:
        raise ValueError("not a GLB file")
    jlen = struct.unpack("<I", b[12:16])[0]
    js = json.loads(b[20:20 + jlen])
    bin_off = 20 + jlen + 8
    sizes = {5126: 4, 5125: 4, 5123: 2, 5121: 1}
    dtypes = {5126: "<f4", 5125: "<u4", 5123: "<u2", 5121: "u1"}
    ncomp = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}

    def accessor(i):
        a = js["accessors"][i]
        bv = js["bufferViews"][a["bufferView"]]
        n = ncomp[a["type"]]
        size = sizes[a["componentType"]]
        stride = bv.get("byteStride", n * size)
        base = bin_off + bv.get("byteOffset", 0) + a.get("byteOffset", 0)
        raw = np.frombuffer(b, dtype=np.uint8, count=(a["count"] - 1) * stride + n * size, offset=base)
        rows = np.lib.stride_tricks.as_strided(raw, shape=(a["count"], n * size), strides=(stride, 1))
        return np.ascontiguousarray(rows).view(dtypes[a["componentType"]]).reshape(a["count"], n)

    prims = []

    def walk(index, parent):
        node = js["nodes"][index]
        world = _mul(parent, _node_matrix(node))
        if "mesh" in node:
            for p in js["meshes"][node["mesh"]]["primitives"]:
                pos = accessor(p["attributes"]["POSITION"]).astype(np.float64)
                x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
                wp = np.stack([((world[0] * x + world[4] * y) + world[8] * z) + world[12],
                               ((world[1] * x + world[5] * y) + world[9] * z) + world[13],
                               ((world[2] * x + world[6] * y) + world[10] * z) + world[14]], 1).astype(np.float32)
                if "indices" in p:
                    idx = accessor(p["indices"]).reshape(-1).astype(np.uint32)
                else:
                    idx = np.arange(len(wp), dtype=np.uint32)
                prims.append((p.get("material", 0), wp, idx))
        for c in node.get("children", []):
            walk(c, world)

    ident = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1], np.float64)
    for r in js["scenes"][js.get("scene", 0)]["nodes"]:
        walk(r, ident)
    names = [m.get("name", "") for m in js.get("materials", [])]
    return prims, names


def to_car_frame(p32: np.ndarray) -> np.ndarray:
    """toCarFrame (car-model.ts:27-29) in float64: (1 + (1.615 - z), y - (-0.04), x)."""
    p = p32.astype(np.float64)
    return np.stack([FRONT_AXLE_X + (MESH_FRONT_AXLE_Z - p[:, 2]), p[:, 1] - MESH_GROUND_Y, p[:, 0]], 1)


def load_prims(glb: Path) -> List[Prim]:
    """bake-car.ts:16-22."""
    raw, names = parse_glb(glb)
    return [Prim(names[m], to_car_frame(wp), idx, names[m] in WHEEL_MATERIALS) for m, wp, idx in raw]


def select_triangles(prims: Sequence[Prim], tags: np.ndarray, want: Callable[[int], bool]) -> np.ndarray:
    """(m, 9) float64 corners of the non-wheel triangles whose first vertex's tag is
    wanted (bake-car.ts:244-253), in the reference's order."""
    out = []
    lut = np.array([bool(want(t)) for t in range(256)])
    vbase = 0
    for p in prims:
        if not p.wheel:
            f = p.indices.reshape(-1, 3)
            keep = lut[tags[vbase + f[:, 0]]]
            tri = p.verts[f[keep]].reshape(-1, 9)
            out.append(tri)
        vbase += len(p.verts)
    return np.ascontiguousarray(np.concatenate(out) if out else np.zeros((0, 9)), dtype=np.float64)


def part_layout(prims: Sequence[Prim], tags: np.ndarray, want: Callable[[int], bool]) -> dict:
    """The grid of one part (bake-car.ts:339-348): the bounding box of its vertices
    (|z|) plus a 4 cm margin, z from the symmetry plane, 1 cm cells."""
    lut = np.array([bool(want(t)) for t in range(256)])
    pts = []
    vbase = 0
    for p in prims:
        if not p.wheel:
            sel = lut[tags[vbase:vbase + len(p.verts)]]
            v = p.verts[sel]
            pts.append(np.stack([v[:, 0], v[:, 1], np.abs(v[:, 2])], 1))
        vbase += len(p.verts)
    pts = np.concatenate(pts)
    mn, mx = pts.min(0), pts.max(0)
    h = PART_GRID_H
    origin = (float(mn[0]) - PART_MARGIN, float(mn[1]) - PART_MARGIN, 0.0)
    dims = tuple(int(math.ceil((float(mx[k]) + PART_MARGIN - origin[k]) / h)) + 1 for k in range(3))
    return {"origin": origin, "h": h, "dims": dims}


class SdfGrid(C.Structure):
    """aero_sdf_grid."""
    _fields_ = [("origin", C.c_double * 3), ("h", C.c_double), ("dims", C.c_uint32 * 3), ("pad", C.c_uint32)]


def bake_grid(tris: np.ndarray, layout: dict, band: float, closing: float, shell: float) -> np.ndarray:
    """aero_bake_sdf: the f16 words of one distance grid (z, y, x order, x fastest)."""
    lib = _lib.load()
    t = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
    g = SdfGrid()
    g.origin[:] = layout["origin"]
    g.h = layout["h"]
    g.dims[:] = layout["dims"]
    nx, ny, nz = layout["dims"]
    out = np.empty(nx * ny * nz, np.uint16)
    check(lib.aero_bake_sdf(ptr(t) if len(t) else None, len(t), C.byref(g), band, closing, shell, ptr(out)), None,
          "aero_bake_sdf")
    return out


def bake_car(glb: Path, tags: np.ndarray, bake=bake_grid):
    """The body grid and the six part grids of bake-car.ts:318-358.  Returns
    (body u16, parts u16 concatenated, partGrids layout list)."""
    prims = load_prims(glb)
    body = bake(select_triangles(prims, tags, lambda t: t == 0), BODY_LAYOUT, BODY_BAND, BODY_CLOSING, BODY_SHELL)
    grids, chunks, offset = [], [], 0
    for part, want, closing in PART_SPECS:
        lay = part_layout(prims, tags, want)
        g = bake(select_triangles(prims, tags, want), lay, max(0.03, closing + 0.02), closing, PART_SHELL)
        grids.append({"part": part, "origin": list(lay["origin"]), "h": lay["h"], "dims": list(lay["dims"]),
                      "offset": offset})
        chunks.append(g)
        offset += len(g)
    return body, np.concatenate(chunks), grids
