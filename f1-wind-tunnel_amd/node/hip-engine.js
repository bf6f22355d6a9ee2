"use strict";
// HipAeroEngine: the simulation surface of the reference's AeroEngine
// (src/engine/engine.ts:159-769) on an MI355X, over the N-API addon aero_napi.node and
// libaero_lbm.so.  Typed for TypeScript hosts by hip-engine.d.ts; plain CommonJS so
// that it runs on the Node of this image without a compile step.
//
// Constants and packing follow engine.ts: quality presets (:14-28), domain length and
// origin (:110, :249, :263), lattice constants, the Shape uniform (:429-448), the
// frame's evenSteps LBM + 2 heat steps (:585, :708-712), force scaling (:739-761).

const path = require("path");

const addon = require(process.env.AERO_NAPI_ADDON || path.join(__dirname, "..", "lib", "aero_napi.node"));

const QUALITIES = { low: [168, 50, 38], medium: [232, 68, 52], high: [296, 88, 66] };
const DOMAIN_LENGTH = 10.4;
const ORIGIN_X = -2.4;
const FORCE_SCALE = 10000;
const AIR_DENSITY = 1.225;
const PROFILE_W = 256, PROFILE_H = 128;
const LATTICE_U = 0.0925;
const TAU0 = 0.508;
const SMAG = 0.17;
const HEAT_COOL = 0.004;
const HEAT_DIFFUSE = 0.04;
const DEG = Math.PI / 180;
const COOLING_HEAT = [1.55, 1, 0.6];
const COOLING_DUCT = [0, 1, 1.3];

const ShapeKind = { F1: 0, Sphere: 1, Cube: 2, Wing: 3, Custom: 4, W17: 5, Empty: 6, Scene: 16 };
const Storage = { F32: 0, F16S: 1, BF16S: 2, MATH_EXACT: 0x100 };

// The app's initial setup (src/main.ts:99-100) fills fields a caller leaves out.
const SHAPE_DEFAULTS = {
  kind: ShapeKind.F1, frontWing: 0, rearWing: 0, sidepod: 0, cover: 0, nose: 0, rearRideH: 0.14, rearMainDeg: 0,
  frontFlapDeg: 0, rearFlapDeg: 0, rideH: 0.12, aoaDeg: 6, customMode: 0, customWidth: 0.6, customAxisY: 1,
  heatBrakes: true, heatEngine: true, heatRad: true, cooling: 1,
};

/** Packs ShapeParams into the 96-byte Shape uniform exactly as engine.ts:429-448. */
function packShape(params, profileMeta) {
  const p = Object.assign({}, SHAPE_DEFAULTS, params);
  const buf = new ArrayBuffer(96);
  const u = new Uint32Array(buf);
  const f = new Float32Array(buf);
  const cool = p.cooling === null ? 1 : p.cooling;
  u[0] = p.kind;
  u[1] = p.customMode;
  u[2] = PROFILE_W;
  u[3] = PROFILE_H;
  f[4] = (p.frontFlapDeg + 14) * DEG;
  f[5] = 9 * DEG;
  f[6] = (p.rearFlapDeg + 32) * DEG;
  f[7] = p.rideH;
  f[8] = p.aoaDeg * DEG;
  f[9] = p.customWidth;
  f[10] = p.customAxisY;
  f[11] = profileMeta ? profileMeta.cell : 0.02;
  f[12] = profileMeta ? profileMeta.originX : 0;
  f[13] = profileMeta ? profileMeta.originY : 0;
  f[14] = p.heatBrakes ? 520 : 0;
  f[15] = p.heatEngine ? 620 * COOLING_HEAT[cool] : 0;
  f[16] = p.heatRad ? 75 * COOLING_HEAT[cool] : 0;
  f[17] = p.frontWing;
  f[18] = p.rearWing;
  f[19] = p.sidepod;
  f[20] = p.cover;
  f[21] = p.nose;
  f[22] = p.rearRideH;
  f[23] = COOLING_DUCT[cool];
  return new Uint8Array(buf);
}

class HipAeroEngine {
  /**
   * @param {string|{dims:number[],h:number,origin:number[]}} quality preset name, or an explicit grid
   * @param {{uLb?:number,tau0?:number,smag?:number,storage?:number,kmh?:number}} [opts]
   */
  constructor(quality, opts) {
    opts = opts || {};
    let dims, h, origin;
    if (typeof quality === "string" || quality === undefined) {
      dims = QUALITIES[quality || "medium"];
      if (!dims) throw new Error(`unknown quality ${quality}`);
      h = DOMAIN_LENGTH / dims[0];
      origin = [ORIGIN_X, 0, 0];
    } else {
      dims = quality.dims;
      h = quality.h;
      origin = quality.origin || [ORIGIN_X, 0, 0];
    }
    this.dims = dims.slice();
    this.n = dims[0] * dims[1] * dims[2];
    this.h = h;
    this.origin = origin.slice();
    this.uLb = opts.uLb !== undefined ? opts.uLb : LATTICE_U;
    this.tau0 = opts.tau0 !== undefined ? opts.tau0 : TAU0;
    this.smag = opts.smag !== undefined ? opts.smag : SMAG;
    this.kmh = opts.kmh !== undefined ? opts.kmh : 250;
    this.frameNo = 0;
    this.stepsSinceRead = 0;
    this.onForces = undefined;
    this.profileMeta = undefined;
    this.profile = null;
    this.ctx = addon.create(new Float64Array([dims[0], dims[1], dims[2], this.n, origin[0], origin[1], origin[2], h]),
                            new Float64Array([this.uLb, this.tau0, this.smag]),
                            opts.storage !== undefined ? opts.storage : Storage.F32);
  }

  destroy() {
    if (this.ctx) {
      addon.destroy(this.ctx);
      this.ctx = null;
    }
  }

  /** engine.ts:371-381 */
  reset() {
    addon.init(this.ctx);
    this.stepsSinceRead = 0;
  }

  /** engine.ts:67-73 custom profile: sdf Float32Array of 256 x 128 plus its placement */
  setProfile(sdf, originX, originY, cell) {
    this.profile = sdf;
    this.profileMeta = { originX, originY, cell };
  }

  /** engine.ts:426-470; prims (Uint8Array of 64-byte aero_prim records) for kind 16 scenes */
  setShape(p, prims) {
    const needsProfile = p.kind === ShapeKind.Custom;
    let prof = null;
    if (needsProfile) prof = this.profile || new Float32Array(PROFILE_W * PROFILE_H).fill(10);
    addon.voxelize(this.ctx, packShape(p, this.profileMeta), prof, prims || null);
  }

  /** nsteps stream-collide steps */
  step(nsteps) {
    addon.step(this.ctx, nsteps);
    this.stepsSinceRead += nsteps;
  }

  heatStep(nsteps, dt) {
    addon.heatStep(this.ctx, new Float32Array([dt, HEAT_COOL, HEAT_DIFFUSE, 0]), nsteps);
  }

  /** The compute part of engine.ts:574-729: evenSteps LBM steps, two heat steps, and a
   *  force readback every 8th frame delivered to onForces. */
  frame(steps) {
    this.frameNo++;
    const even = Math.max(2, steps - (steps % 2));
    this.step(even);
    this.heatStep(2, even / 2);
    if (this.frameNo % 8 === 0) {
      const f = this.readForces(true);
      if (f && this.onForces) this.onForces(f);
    }
  }

  /** engine.ts:739-761; quantized = the reference's i32 fixed-point sums */
  readForces(reset, quantized) {
    const r = addon.readForces(this.ctx, reset !== false);
    if (reset !== false) this.stepsSinceRead = 0;
    if (r.steps === 0) return null;
    const vReal = this.kmh / 3.6;
    const raw = quantized ? r.quantized.map((v) => v / FORCE_SCALE) : r.sums;
    const scale = (1 / r.steps) * AIR_DENSITY * (vReal / this.uLb) ** 2 * this.h * this.h * 2;
    const pair = (g) => ({ drag: raw[g * 3] * scale, down: -raw[g * 3 + 1] * scale });
    const body = pair(0), frontWing = pair(1), rearWing = pair(2);
    return {
      dragN: body.drag + frontWing.drag + rearWing.drag,
      downforceN: body.down + frontWing.down + rearWing.down,
      body, frontWing, rearWing,
    };
  }

  /** engine.ts:763-768 */
  async debugRead() {
    const flags = addon.readFlags(this.ctx);
    return { fields: addon.readFields(this.ctx), solid: flags.solid, temps: addon.readTemps(this.ctx) };
  }

  populations() { return addon.readPopulations(this.ctx); }
  guardCount() { return addon.guardCount(this.ctx); }
  sync() { addon.sync(this.ctx); }
}

module.exports = { HipAeroEngine, ShapeKind, Storage, packShape, addon, QUALITIES, LATTICE_U, TAU0, SMAG };
