"""Force post-processing between the tunnel's force readback and what the app shows:
the W17's panel-method wings and calibrated body, the telemetry averaging, and the W17
ride-height map cells.

Reference: src/main.ts:16-28 (constants), :586-616 (withPanelWings: body coefficients
averaged with a sample-count-dependent rate, normalised by dynamic pressure, calibrated
and offset by the sidepod / cooling package; wing forces from the panel method),
:618-625 (frontAxleShare, in vlm.py), :627-645 and :669-689 (onForces: warm-up gate,
3 s window mean, chart smoothing), src/ui/aeromap.ts:90-101 (a map cell from the mean of
the samples taken at one pair of ride heights).
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

from .engine import ForcePair, Forces
from . import vlm as V

KGF = 9.81
AIR_DENSITY = 1.225
# main.ts:20-28: ΔCdA / ΔClA in m² per cooling and sidepod package; the body
# calibration of the W17 (drag scale, missing floor load).
COOLING_COEF = ((-0.03, 0.01), (0.0, 0.0), (0.055, -0.02))
SIDEPOD_COEF = ((0.0, 0.0), (-0.05, 0.035))
BODY_CDA_SCALE = 0.69
BODY_CLA_FLOOR = 0.31
# main.ts:634-645
AVG_WINDOW_MS = 3000.0
CHART_POINTS = 220
WARMUP_MS = 2500.0
FORCE_KEYS = ("drag", "down", "fd", "fdd", "rd", "rdd", "bd", "bdd")


def _pick(table, i: int, default: int):
    """`TABLE[i] ?? TABLE[default]` for a small integer index."""
    return table[i] if 0 <= int(i) < len(table) else table[default]


@dataclass
class PanelWings:
    """withPanelWings (main.ts:586-616) with its body-coefficient state (bodyCoef,
    reset by resetBodyCoef at :593 whenever the chart resets)."""
    cd: float = 0.0
    cl: float = 0.0
    n: int = 0

    def reset(self) -> None:
        self.n = 0

    def apply(self, f: Forces, wing_loads: Optional[Dict[str, V.WingLoads]], kmh: float, sidepods: int = 0,
              cooling: int = 1, w17: bool = True) -> Forces:
        """The forces the app shows: unchanged unless the car is the W17 and its wing
        loads are known; then body = calibrated long-run coefficients x q, wings = the
        panel method, totals their sums."""
        if not w17 or wing_loads is None:
            return f
        q = 0.5 * AIR_DENSITY * (kmh / 3.6) ** 2
        a = 0.15 if self.n < 20 else (0.04 if self.n < 60 else 0.011)
        self.cd += (f.body.drag / q - self.cd) * (1 if self.n == 0 else a)
        self.cl += (f.body.down / q - self.cl) * (1 if self.n == 0 else a)
        self.n += 1
        pods = _pick(SIDEPOD_COEF, sidepods, 0)
        cool = _pick(COOLING_COEF, cooling, 1)
        fw = ForcePair(drag=wing_loads["front"].dragN, down=wing_loads["front"].downforceN)
        rw = ForcePair(drag=wing_loads["rear"].dragN, down=wing_loads["rear"].downforceN)
        body = ForcePair(drag=(self.cd * BODY_CDA_SCALE + pods[0] + cool[0]) * q,
                         down=(self.cl + BODY_CLA_FLOOR + pods[1] + cool[1]) * q)
        return Forces(dragN=body.drag + fw.drag + rw.drag, downforceN=body.down + fw.down + rw.down, body=body,
                      frontWing=fw, rearWing=rw)


def sample_of(t_ms: float, f: Forces) -> Dict[str, float]:
    """The telemetry Sample of one reading (main.ts:673-676)."""
    return {"t": t_ms, "drag": f.dragN, "down": f.downforceN, "fd": f.frontWing.drag, "fdd": f.frontWing.down,
            "rd": f.rearWing.drag, "rdd": f.rearWing.down, "bd": f.body.drag, "bdd": f.body.down}


@dataclass
class ForceTelemetry:
    """engine.onForces (main.ts:669-689) without the drawing: readings before the
    warm-up ends are dropped; `smoothed` is the mean of the readings of the last 3 s;
    `history` holds up to 220 chart points, each moved 35 % of the way from the
    previous point to the new reading."""
    warmup_until: float = WARMUP_MS
    recent: deque = field(default_factory=deque)
    smoothed: Dict[str, float] = field(default_factory=lambda: {k: 0.0 for k in ("t",) + FORCE_KEYS})
    history: List[Dict[str, float]] = field(default_factory=list)

    def reset(self, now_ms: float, panel: Optional[PanelWings] = None) -> None:
        """resetChart (main.ts:647-651): the chart, and the body coefficients."""
        self.history.clear()
        self.recent.clear()
        if panel is not None:
            panel.reset()
        self.warmup_until = now_ms + WARMUP_MS

    def on_forces(self, now_ms: float, f: Forces,
                  post: Optional[Callable[[Forces], Forces]] = None) -> Optional[Dict[str, float]]:
        """One reading; `post` (withPanelWings) runs only after the warm-up gate, as in
        the reference.  Returns the raw sample, or None while warming up."""
        if now_ms < self.warmup_until:
            return None
        if post is not None:
            f = post(f)
        raw = sample_of(now_ms, f)
        self.recent.append(raw)
        while len(self.recent) > 1 and now_ms - self.recent[0]["t"] > AVG_WINDOW_MS:
            self.recent.popleft()
        for k in FORCE_KEYS:
            acc = 0.0
            for r in self.recent:
                acc = acc + r[k]
            self.smoothed[k] = acc / len(self.recent)
        point = dict(raw)
        if self.history:
            last = self.history[-1]
            for k in FORCE_KEYS:
                point[k] = last[k] + (raw[k] - last[k]) * 0.35
        self.history.append(point)
        if len(self.history) > CHART_POINTS:
            self.history.pop(0)
        return raw


def cell_of(samples: Sequence[Dict[str, float]], balance: Optional[float]) -> Optional[dict]:
    """A map cell from the samples taken at one pair of ride heights
    (aeromap.ts:96-101): mean downforce in kg, downforce / max(drag, 1), and the
    balance hook's value or 0.5; None without samples."""
    if not samples:
        return None
    down = 0.0
    drag = 0.0
    for s in samples:
        down = down + s["down"]
    for s in samples:
        drag = drag + s["drag"]
    down /= len(samples)
    drag /= len(samples)
    return {"down": down / KGF, "eff": down / max(drag, 1.0), "bal": 0.5 if balance is None else balance}


def w17_map_cells(forces: Sequence[Forces], pairs, rig: dict, setup: V.W17WingSetup, kmh: float,
                  sidepods: int = 0, cooling: int = 1) -> List[dict]:
    """The W17 ride-height map as the reference fills it (main.ts:669-675 feeding
    aeromap.ts:96-101), from one measured Forces per (front, rear) pair of a batched
    run: each point's forces go through withPanelWings with the panel-method wing
    loads at that pair's ride heights, and `bal` is frontAxleShare of those loads.
    Every pair's front and rear lattices are solved in one GPU call.

    The body term starts from a fresh coefficient state per point, so a cell's body
    load is that point's own coefficient.  In the reference the state runs on across
    the sweep (its rate falls to 0.011 per reading, a 12 s average), so its cells mix
    the body loads of the points measured before; a batched map has no such order."""
    setups = [V.W17WingSetup(**{**setup.__dict__, "rideH": float(fr), "rearRideH": float(rr)}) for fr, rr in pairs]
    loads = V.w17_wing_loads_many(rig, setups, [kmh / 3.6] * len(setups))
    cells = []
    for f, ld, (fr, rr) in zip(forces, loads, pairs):
        shown = PanelWings().apply(f, ld, kmh, sidepods, cooling, w17=True)
        c = cell_of([sample_of(0.0, shown)], V.front_axle_share(ld))
        c.update({"front": fr, "rear": rr, "forces": shown})
        cells.append(c)
    return cells
