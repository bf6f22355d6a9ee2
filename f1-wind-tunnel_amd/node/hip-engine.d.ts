// Type declarations of hip-engine.js for TypeScript hosts: the simulation surface of
// the reference's AeroEngine (src/engine/engine.ts:159-769), with ShapeParams and
// Forces as declared there (:36-65, :100-107).

export declare const ShapeKind: {
  readonly F1: 0; readonly Sphere: 1; readonly Cube: 2; readonly Wing: 3; readonly Custom: 4;
  readonly W17: 5; readonly Empty: 6; readonly Scene: 16;
};
export declare const Storage: { readonly F32: 0; readonly F16S: 1; readonly BF16S: 2; readonly MATH_EXACT: 0x100 };

export interface ShapeParams {
  kind: number;
  frontWing?: number; rearWing?: number; sidepod?: number; cover?: number; nose?: number;
  rearRideH: number; rideH: number;
  rearMainDeg?: number; frontFlapDeg?: number; rearFlapDeg?: number; aoaDeg?: number;
  customMode?: number; customWidth?: number; customAxisY?: number;
  heatBrakes?: boolean; heatEngine?: boolean; heatRad?: boolean; cooling?: number;
}

export interface ForcePair { drag: number; down: number }
export interface Forces { dragN: number; downforceN: number; body: ForcePair; frontWing: ForcePair; rearWing: ForcePair }

export interface GridSpec { dims: [number, number, number]; h: number; origin?: [number, number, number] }
export interface EngineOptions { uLb?: number; tau0?: number; smag?: number; storage?: number; kmh?: number }

export declare function packShape(p: ShapeParams, profileMeta?: { originX: number; originY: number; cell: number }): Uint8Array;

export declare class HipAeroEngine {
  constructor(quality?: "low" | "medium" | "high" | GridSpec, opts?: EngineOptions);
  readonly dims: [number, number, number];
  readonly n: number;
  readonly h: number;
  readonly origin: [number, number, number];
  uLb: number;
  kmh: number;
  onForces?: (f: Forces) => void;
  destroy(): void;
  reset(): void;
  setProfile(sdf: Float32Array, originX: number, originY: number, cell: number): void;
  setShape(p: ShapeParams, prims?: Uint8Array | null): void;
  step(nsteps: number): void;
  heatStep(nsteps: number, dt: number): void;
  frame(steps: number): void;
  readForces(reset?: boolean, quantized?: boolean): Forces | null;
  debugRead(): Promise<{ fields: Float32Array; solid: Uint32Array; temps: Float32Array }>;
  populations(): Float32Array;
  guardCount(): number;
  sync(): void;
}
