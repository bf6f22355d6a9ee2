"""Host side of the W17 baked car (shape kind 5).

Reference: src/engine/car-model.ts:15-23 (layout), :151-171 (rigUniform), :174-190
(partsAtlas), src/engine/engine.ts:410-422 (setCarModel) and :451-462 (the rig that
setShape derives from the setup).  The grids are the f16 words of
public/models/w17.sdf.bin (368 x 84 x 67, x fastest) and w17.parts.bin (one grid per
articulated part, located by w17.rig.json's partGrids).
"""
from __future__ import annotations

import json
import math
from pathlib import Path
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from ._lib import Rig

DEG = math.pi / 180
BODY_DIMS = (368, 84, 67)
BODY_ORIGIN = (-0.3, 0.0, 0.0)
BODY_H = 0.015
ATLAS_FAR = 0x2400  # f16 0.015625, the atlas padding of car-model.ts:178

# car-model.ts:138-149, engine.ts:111-113
FRONT_LEVEL_FLAP = (0.0, 7.0, -9.0)
REAR_LEVEL_MAIN = (0.0, 3.0, -4.0)
REAR_LEVEL_FLAP = (0.0, 9.0, -13.0)
REAR_SHAPE_TWIST = {0: 0.0, 1: -16.0, 2: 6.0}
THIN_FLAP_SCALE = 0.72
COMPACT_POD_SQUEEZE = 0.32


def parts_atlas(rig: Dict, parts: np.ndarray) -> Tuple[Tuple[int, int, int], np.ndarray]:
    """partsAtlas (car-model.ts:174-190): the part grids stacked along z, one padding
    slice between them, padded with ATLAS_FAR.  Returns (dims (nx, ny, nz), u16 data)."""
    grids = rig["partGrids"]
    nx = max(g["dims"][0] for g in grids)
    ny = max(g["dims"][1] for g in grids)
    nz = sum(g["dims"][2] + 1 for g in grids)
    out = np.full((nz, ny, nx), ATLAS_FAR, np.uint16)
    z0 = 0
    for g in grids:
        gx, gy, gz = g["dims"]
        src = parts[g["offset"]:g["offset"] + gx * gy * gz].reshape(gz, gy, gx)
        out[z0:z0 + gz, :gy, :gx] = src
        z0 += gz + 1
    return (nx, ny, nz), out.reshape(-1)


def rig_uniform(rig: Dict, angles: Sequence[float], rear_twist: float = 0.0, flap_scale: float = 1.0,
                pod_squeeze: float = 0.0, fly: Sequence[float] = (0.0, 0.0, 1.0, 1.0)) -> Rig:
    """rigUniform (car-model.ts:151-171) packed as the Rig struct."""
    r = Rig()
    fh0, fh1 = rig["frontHinges"]
    for i in range(71):
        r.front[i][0], r.front[i][1] = fh0[i]
        r.front[i][2], r.front[i][3] = fh1[i]
    r.hinges[0], r.hinges[1] = rig["rearMainHinge"]
    r.hinges[2], r.hinges[3] = rig["rearFlapHinge"]
    for k in range(4):
        r.angles[k] = angles[k]
        r.fly[k] = fly[k]
    r.extra[0], r.extra[1], r.extra[2], r.extra[3] = rear_twist, flap_scale, pod_squeeze, 0.0
    z = 0
    grids = rig["partGrids"]
    for i in range(12):
        r.grids[i][0] = r.grids[i][1] = r.grids[i][2] = 0.0
        r.grids[i][3] = 1.0
    for i, g in enumerate(grids):
        r.grids[2 * i][0], r.grids[2 * i][1], r.grids[2 * i][2] = g["origin"]
        r.grids[2 * i][3] = g["h"]
        r.grids[2 * i + 1][0], r.grids[2 * i + 1][1], r.grids[2 * i + 1][2] = g["dims"]
        r.grids[2 * i + 1][3] = z
        z += g["dims"][2] + 1
    return r


def rig_for_setup(rig: Dict, p) -> Rig:
    """The rig setShape derives from a setup (engine.ts:451-462)."""
    front = FRONT_LEVEL_FLAP[int(p.frontWing)] if 0 <= int(p.frontWing) < 3 else FRONT_LEVEL_FLAP[0]
    rw = int(p.rearWing) if 0 <= int(p.rearWing) < 3 else 0
    front_deg = p.frontFlapDeg + front
    angles = [front_deg * 0.8 * DEG, front_deg * DEG, (p.rearFlapDeg + REAR_LEVEL_FLAP[rw]) * DEG,
              (REAR_LEVEL_MAIN[rw] + p.rearMainDeg) * DEG]
    return rig_uniform(rig, angles, rear_twist=REAR_SHAPE_TWIST.get(int(p.rearShape), 0.0) * DEG,
                       flap_scale=THIN_FLAP_SCALE if p.thinWings else 1.0,
                       pod_squeeze=COMPACT_POD_SQUEEZE if int(p.sidepods) == 1 else 0.0,
                       fly=(0.0, 0.0, 1.0 if p.frontWingOn else 0.0, 1.0 if p.rearWingOn else 0.0))


def load_reference_car(models: Path):
    """(body u16, parts u16, rig dict) from a directory holding w17.sdf.bin,
    w17.parts.bin and w17.rig.json."""
    models = Path(models)
    body = np.fromfile(models / "w17.sdf.bin", dtype="<u2")
    parts = np.fromfile(models / "w17.parts.bin", dtype="<u2")
    rig = json.loads((models / "w17.rig.json").read_text())
    assert body.size == BODY_DIMS[0] * BODY_DIMS[1] * BODY_DIMS[2]
    return body, parts, rig


def _sd_roundbox(px, py, pz, c, b, r):
    qx, qy, qz = np.abs(px - c[0]) - (b[0] - r), np.abs(py - c[1]) - (b[1] - r), np.abs(pz - c[2]) - (b[2] - r)
    out = np.sqrt(np.maximum(qx, 0) ** 2 + np.maximum(qy, 0) ** 2 + np.maximum(qz, 0) ** 2)
    return out + np.minimum(np.maximum(qx, np.maximum(qy, qz)), 0) - r


def synthetic_car(rig: Dict) -> Tuple[np.ndarray, np.ndarray]:
    """Grids in the shipped format with analytic content, for tests that must run
    where the model files are absent: a rounded-box tub with an airbox in the body
    grid, and a thin rounded plate across each part grid (its chord along x)."""
    nx, ny, nz = BODY_DIMS
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    px, py, pz = BODY_ORIGIN[0] + x * BODY_H, y * BODY_H, z * BODY_H
    d = np.minimum(_sd_roundbox(px, py, pz, (2.4, 0.42, 0.0), (2.3, 0.28, 0.42), 0.1),
                   np.sqrt((px - 3.3) ** 2 + (py - 0.85) ** 2 + pz ** 2) - 0.2)
    body = d.astype(np.float16).view(np.uint16).reshape(-1)
    total = sum(g["dims"][0] * g["dims"][1] * g["dims"][2] for g in rig["partGrids"])
    parts = np.empty(total, np.uint16)
    for g in rig["partGrids"]:
        gx, gy, gz = g["dims"]
        h = g["h"]
        kz, ky, kx = np.meshgrid(np.arange(gz), np.arange(gy), np.arange(gx), indexing="ij")
        ext = (np.array([gx, gy, gz]) - 1) * h
        c = (ext[0] / 2, ext[1] / 2, ext[2] / 2)
        plate = _sd_roundbox(kx * h, ky * h, kz * h, c, (ext[0] * 0.4, 0.012, ext[2] * 0.45), 0.005)
        parts[g["offset"]:g["offset"] + gx * gy * gz] = plate.astype(np.float16).view(np.uint16).reshape(-1)
    return body, parts
