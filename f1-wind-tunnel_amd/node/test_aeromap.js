"use strict";
// Node-side batched ride-height map (run on a machine with the GPU): the reference's
// 5 x 5 front/rear sweep (src/ui/aeromap.ts:15-19) of the generic F1 car on a coarse
// tunnel, settled and measured in one batch.  Prints the cells as one JSON line for the
// Python host to compare.
const { aeroMap, ShapeKind, Storage } = require("./hip-engine");

const out = aeroMap({
  grid: { dims: [96, 32, 32], h: 0.12, origin: [-2.4, 0, 0] },
  shape: (front, rear) => ({ kind: ShapeKind.F1, rideH: front, rearRideH: rear, aoaDeg: 6, cooling: 1,
                             heatBrakes: false, heatEngine: false, heatRad: false }),
  settle: 40, measure: 40, stepsPerGraph: 20, storage: Storage.F16S,
});
const ok = out.cells.length === 25 && out.cells.every((c) => Number.isFinite(c.down) && Number.isFinite(c.eff));
console.log(JSON.stringify({ ok, deviceMs: out.deviceMs,
  cells: out.cells.map((c) => ({ front: c.front, rear: c.rear, down: c.down, eff: c.eff, bal: c.bal,
                                 dragN: c.forces.dragN, downforceN: c.forces.downforceN })) }));
process.exit(ok ? 0 : 1);
