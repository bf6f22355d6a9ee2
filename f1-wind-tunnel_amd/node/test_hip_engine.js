"use strict";
// Node-side checks of HipAeroEngine over the N-API addon (run on a machine with the
// GPU): closed-form voxel count of the 1 m cube, the uniform-flow fixed point of an
// empty tunnel, force scaling and readback, error mapping.  Prints one JSON line with
// the figures a caller can cross-check against the Python host.
const { HipAeroEngine, ShapeKind, Storage, FlowMode, ColorField } = require("./hip-engine");

const f32 = (x) => Math.fround(x);
const cubeGrid = { dims: [128, 64, 64], h: f32(1 / 22), origin: [f32(2.6 - 42 / 22), 0, 0] };
const out = {};
let failed = 0;
const check = (name, ok, detail) => {
  out[name] = { ok, detail };
  if (!ok) failed++;
};

(async () => {
  // 1) voxel count of the cube: 22 x 22 x 11 cells in the half tunnel
  const e = new HipAeroEngine(cubeGrid, { uLb: 0.05, tau0: 0.51, smag: 0.1, storage: Storage.F32 | Storage.MATH_EXACT });
  e.setShape({ kind: ShapeKind.Cube, rideH: 0.1, rearRideH: 0.12, aoaDeg: 6, cooling: 1 });
  e.reset();
  let dbg = await e.debugRead();
  let solid = 0;
  for (let i = 0; i < e.n; i++) if (dbg.solid[i]) solid++;
  check("cube_cells", solid === 22 * 22 * 11, solid);

  // 2) forces after 200 steps: drag positive, fixed-point and FP64 sums agree
  e.step(200);
  const fd = e.readForces(false, false);
  const fq = e.readForces(true, true);
  check("cube_drag", fd.dragN > 0 && Math.abs(fq.dragN - fd.dragN) <= 1e-3 * Math.abs(fd.dragN),
        { fp64: fd.dragN, quantized: fq.dragN });
  const raw = require("./hip-engine").addon.readForces(e.ctx, false);
  check("reset_zeroes", raw.steps === 0, raw.steps);
  out.cube_forces = fd;

  // 3) flow visualisation on the settled cube flow: three smoke frames, a pressure-
  // coloured fog frame, a streamline and a vortex frame; sums for the Python cross-check
  // checksum: the float32 words summed as unsigned integers modulo 2^32
  const sum = (a) => {
    const w = new Uint32Array(a.buffer, a.byteOffset, a.length);
    let s = 0;
    for (let i = 0; i < w.length; i++) s = (s + w[i]) >>> 0;
    return s;
  };
  for (let k = 0; k < 3; k++) e.flowFrame(FlowMode.Smoke, ColorField.Speed, 8);
  e.flowFrame(FlowMode.Fog, ColorField.Pressure, 8);
  const parts = e.particles();
  const lines = e.flowFrame(FlowMode.Streamlines, ColorField.Speed, 8);
  const vort = e.flowFrame(FlowMode.Vortices, ColorField.Wake, 8);
  let live = 0;
  for (let i = 3; i < lines.length; i += 4) if (lines[i] >= 0) live++;
  check("particles", parts.length === 140000 * 8 && parts.some((v, i) => i % 8 === 3 && v > 0), parts.length);
  check("streamlines", lines.length === 90 * 360 * 4 && live > 90 * 10, live);
  out.flow = { particles: sum(parts), lines: sum(lines), vortices: sum(vort), live };
  e.destroy();

  // 4) an empty tunnel at uniform inflow is a fixed point of the whole step
  const g = new HipAeroEngine({ dims: [40, 16, 12], h: 0.1, origin: [0, 0, 0] },
                              { uLb: 0.05, tau0: 0.51, smag: 0.1, storage: Storage.F32 | Storage.MATH_EXACT });
  g.setShape({ kind: ShapeKind.Empty, rideH: 0.1, rearRideH: 0.12 });
  g.reset();
  g.step(20);
  dbg = await g.debugRead();
  let dev = 0;
  for (let i = 0; i < g.n; i++) {
    dev = Math.max(dev, Math.abs(dbg.fields[4 * i] - f32(0.05)), Math.abs(dbg.fields[4 * i + 1]),
                   Math.abs(dbg.fields[4 * i + 2]), Math.abs(dbg.fields[4 * i + 3] - 1));
  }
  check("uniform_fixed_point", dev <= 1e-6, dev);
  // frame(): 8 frames of 8 steps deliver one force reading
  let got = null;
  g.onForces = (f) => { got = f; };
  for (let k = 0; k < 8; k++) g.frame(8);
  check("frame_forces", got !== null && Math.abs(got.dragN) < 1e-3, got && got.dragN);
  g.destroy();

  // 5) errors surface as thrown Errors with the library's message
  let msg = "";
  try {
    new HipAeroEngine({ dims: [16, 8, 8], h: 0.1 }, { tau0: 0.5 });
  } catch (err) {
    msg = String(err.message);
  }
  check("error_mapping", /aero_create failed/.test(msg), msg);

  console.log(JSON.stringify(out));
  process.exit(failed ? 1 : 0);
})().catch((err) => {
  console.error(err);
  process.exit(2);
});
