"""GPU distance-grid baker (aero_bake_sdf, csrc/bake.hip): the reference's own baked
part grids of the W17 (front flap A and rear flap, from public/models/w17.parts.bin)
reproduced bit for bit from their triangles; a sphere on the symmetry plane at the
body grid's size against the FP64 oracle; an empty mesh; argument errors."""
import ctypes as C
import time
from pathlib import Path

import numpy as np
import pytest

from aerolbm import _lib
from aerolbm import bake as B
from aerolbm._lib import AeroError
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
BO = orc.BakeOracle()


def test_reference_part_grids_bit_for_bit():
    with np.load(GOLD / "w17_part_grids.npz", allow_pickle=False) as d:
        for part in (1, 3):
            print(f"bake part {part}", flush=True)
            lay = {"origin": tuple(d[f"p{part}_origin"]), "h": float(d[f"p{part}_h"]),
                   "dims": tuple(int(v) for v in d[f"p{part}_dims"])}
            band, closing, shell = (float(v) for v in d[f"p{part}_params"])
            got = B.bake_grid(d[f"p{part}_tris"], lay, band, closing, shell)
            assert np.array_equal(got, d[f"p{part}_expected"]), part


def sphere(r, centre, n=48):
    th = np.linspace(0, np.pi, n + 1)
    ph = np.linspace(0, 2 * np.pi, 2 * n + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    X = np.stack([r * np.sin(T) * np.cos(P), r * np.cos(T), r * np.sin(T) * np.sin(P)], -1) + centre
    a, b, c, e = X[:-1, :-1], X[1:, :-1], X[1:, 1:], X[:-1, 1:]
    t1 = np.concatenate([a, b, c], -1).reshape(-1, 9)
    t2 = np.concatenate([a, c, e], -1).reshape(-1, 9)
    return np.ascontiguousarray(np.concatenate([t1, t2]))


def test_body_size_sphere_matches_oracle():
    """368 x 84 x 67 cells at 1.5 cm (the W17 body grid) around two UV spheres (11,520
    triangles), one cut by the symmetry plane, with the body grid's band, closing and
    shell."""
    print("bake sphere at body size", flush=True)
    tris = np.concatenate([sphere(0.4, np.array([2.4, 0.6, 0.0])), sphere(0.15, np.array([4.6, 0.9, 0.3]), 24)])
    lay = B.BODY_LAYOUT
    t0 = time.perf_counter()
    got = B.bake_grid(tris, lay, B.BODY_BAND, B.BODY_CLOSING, B.BODY_SHELL)
    print(f"GPU bake {len(tris)} triangles over {np.prod(lay['dims'])} cells: "
          f"{(time.perf_counter() - t0) * 1e3:.1f} ms", flush=True)
    want = BO.bake(tris, lay, B.BODY_BAND, B.BODY_CLOSING, B.BODY_SHELL)
    assert np.array_equal(got, want)
    v = got.view(np.float16)
    assert (v < 0).sum() > 10000 and (v > 0).sum() > 1000000


def test_empty_mesh_is_all_outside():
    lay = {"origin": (0.0, 0.0, 0.0), "h": 0.01, "dims": (20, 10, 5)}
    got = B.bake_grid(np.zeros((0, 9)), lay, 0.05, 0.02, 0.004)
    assert np.all(got == BO.to_half(float(np.float32(0.05)) - 0.004))


def test_bake_argument_errors():
    lib = _lib.load()
    g = B.SdfGrid()
    g.h = 0.01
    g.dims[:] = (0, 4, 4)
    out = np.zeros(16, np.uint16)
    assert lib.aero_bake_sdf(None, 0, C.byref(g), 0.05, 0.02, 0.004, _lib.ptr(out)) == -1
    assert b"grid" in lib.aero_global_error()
    g.dims[:] = (4, 2, 2)
    assert lib.aero_bake_sdf(None, 0, C.byref(g), -1.0, 0.02, 0.004, _lib.ptr(out)) == -1
    assert lib.aero_bake_sdf(None, 3, C.byref(g), 0.05, 0.02, 0.004, _lib.ptr(out)) == -1  # triangles missing
    with pytest.raises(AeroError):
        B.bake_grid(np.zeros((0, 9)), {"origin": (0, 0, 0), "h": 0.0, "dims": (4, 4, 4)}, 0.05, 0.02, 0.004)
