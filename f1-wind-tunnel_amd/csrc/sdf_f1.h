// Analytic F1 car (shape kind 0) for the voxelizer.
// Reference definitions: src/engine/f1.ts:20-320 (sdF1 and its parts) and
// src/engine/f1.ts:397-407 (f1Emission).  f32 arithmetic in the reference's order.
// The rake angle and the rear-wing main-plane angle depend on the shape only; their
// cos/sin arrive precomputed in F1Ctx.  The per-point angles of sdWingZ use
// sincos_r64 (sdf.h), which the CPU oracle evaluates to the same bits.
#pragma once
#include "sdf.h"

namespace aero {
namespace f1 {

using namespace sdf;

struct F1Ctx {
    const aero_shape* sh;
    float rake_c, rake_s;   // cos/sin(rakeAngle())                     f1.ts:37
    float rw_cos, rw_sin;   // cos/sin of the rear main-plane angle angM  f1.ts:268
    float rw_angm;          // angM
};

constexpr float DEG = 0.017453292f;
constexpr float FA = 1.0f, RA = 4.6f, WR = 0.36f, FWZ = 0.84f, FHW = 0.152f, RWZ = 0.80f, RHW = 0.20f;
constexpr float RA26 = 4.4f, WR26 = 0.355f, FHW26 = 0.14f, RHW26 = 0.1875f;
constexpr float BODY_X = 0.9444444f, BODY_Z = 0.96f, BODY_LIP = 0.96f;

struct R2 { float d, m; };
__device__ __forceinline__ R2 opU2(R2 a, R2 b) { return (b.d < a.d) ? b : a; }

__device__ __forceinline__ v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }

__device__ __forceinline__ v3 toBodyFrame(const F1Ctx& c, v3 pw) {
    const float dx = pw.x - FA, dy = pw.y - c.sh->ride_h;
    const v3 p = mk3(FA + dx * c.rake_c + dy * c.rake_s, c.sh->ride_h - dx * c.rake_s + dy * c.rake_c, pw.z);
    float x = p.x;
    if (x > FA) x = (x > RA26) ? x + (RA - RA26) : FA + (x - FA) / BODY_X;
    return mk3(x, p.y, p.z / BODY_Z);
}
__device__ __forceinline__ float smin(float a, float b, float k) {
    const float h = fmax_(k - fabs_(a - b), 0.0f) / k;
    return fmin_(a, b) - h * h * k * 0.25f;
}
__device__ __forceinline__ float smax(float a, float b, float k) { return -smin(-a, -b, k); }
__device__ __forceinline__ float sdEllipsoid(v3 p, v3 r) {
    const float k0 = length(mk3(p.x / r.x, p.y / r.y, p.z / r.z));
    const float k1 = length(mk3(p.x / (r.x * r.x), p.y / (r.y * r.y), p.z / (r.z * r.z)));
    return k0 * (k0 - 1.0f) / fmax_(k1, 1e-5f);
}
__device__ __forceinline__ float sdWingZ(v3 p, v2 le, float chord, float ang, float thick, float z0, float z1,
                                         float rise, float taper, float curl) {
    const float s = smoothstep(curl, z1, p.z);
    const float a = ang + rise * s * 0.6f;
    float sa, ca;
    sincos_r64(a, &sa, &ca);
    return sdWingCS(p, mk2(le.x, le.y + rise * s), chord * (1.0f - taper * s), ca, sa, thick, z0, z1) * 0.85f;
}
__device__ __forceinline__ float tunnelRoof(float x, float b) {
    if (x < 2.5f) return mix(b + 0.12f, b + 0.05f, smoothstep(1.55f, 2.5f, x));
    if (x < 3.8f) return b + 0.05f;
    return mix(b + 0.05f, b + 0.25f, smoothstep(3.8f, 4.98f, x));
}
__device__ float sdFloor(v3 p, float b) {
    const float roof = tunnelRoof(p.x, b);
    const float zLim = mix(0.80f, 0.58f, smoothstep(3.6f, 4.15f, p.x));
    const float top = (fabs_(p.z - 0.40f) < 0.28f) ? roof + 0.03f : b + 0.05f;
    const float slab = fmax_(fmax_(b - p.y, p.y - top), fmax_(fabs_(p.x - 3.265f) - 1.715f, p.z - zLim));
    const float tunnel = fmax_(p.y - roof, fabs_(p.z - 0.40f) - 0.255f);
    float d = fmax_(slab, -tunnel) * 0.8f;
    for (int i = 0; i < 2; i++) {
        const float fz = 0.26f + (float)i * 0.16f;
        const float fence = sdBox(p - mk3(1.84f, b + 0.08f, fz), mk3(0.28f, 0.08f, 0.012f));
        d = fmin_(d, fmax_(fence, p.y - roof - 0.02f));
    }
    for (int i = 0; i < 2; i++) {
        const float sz = 0.30f + (float)i * 0.20f;
        const float strake = sdBox(p - mk3(4.62f, b + 0.17f, sz), mk3(0.34f, 0.17f, 0.008f));
        d = fmin_(d, fmax_(strake, p.y - roof));
    }
    d = fmin_(d, sdCapsule(p, mk3(2.25f, b + 0.065f, 0.79f), mk3(3.85f, b + 0.05f, 0.78f), 0.02f));
    return d;
}
__device__ __forceinline__ float sdTyre(v3 p, v3 c, float hw) {
    const v3 q = p - c;
    return sdRoundRect(mk2(length(mk2(q.x, q.y)) - 0.2925f, q.z), mk2(0.0625f, hw), 0.06f);
}
__device__ __forceinline__ float sdWheelCover(v3 p, v3 c, float hw) {
    const v3 q = p - c;
    float d = sdCylZ(q, 0.234f, hw - 0.028f);
    d = fmin_(d, sdCylZ(q - mk3(0.0f, 0.0f, hw - 0.02f), 0.045f, 0.02f));
    return d;
}
__device__ R2 sdWheelSet(v3 p, R2 res0, v3 c, float hw, float inboard) {
    R2 res = res0;
    if (sdBox(p - c, mk3(WR26 + 0.02f, WR26 + 0.02f, hw + 0.2f)) > res.d) return res;
    res = opU2(res, R2{sdTyre(p, c, hw), 2.0f});
    res = opU2(res, R2{sdWheelCover(p, c, hw), 9.0f});
    const float duct = sdCylZ(p - mk3(c.x, c.y, c.z - hw - inboard * 0.5f), 0.17f, inboard * 0.5f);
    const float ductWing = sdRoundBox(p - mk3(c.x - 0.04f, c.y + 0.21f, c.z - hw - 0.05f), mk3(0.16f, 0.012f, 0.05f), 0.006f);
    res = opU2(res, R2{fmin_(duct, ductWing), 10.0f});
    return res;
}
__device__ float sdSuspension(v3 p, float b) {
    const v3 fo = mk3(FA + 0.02f, WR + 0.16f, FWZ - FHW - 0.05f);
    float d = sdCapsule(p, mk3(1.10f, b + 0.43f, 0.17f), fo, 0.017f);
    d = fmin_(d, sdCapsule(p, mk3(1.42f, b + 0.40f, 0.17f), fo, 0.017f));
    const v3 fl = mk3(FA, WR - 0.13f, FWZ - FHW - 0.05f);
    d = fmin_(d, sdCapsule(p, mk3(0.98f, b + 0.19f, 0.15f), fl, 0.017f));
    d = fmin_(d, sdCapsule(p, mk3(1.38f, b + 0.17f, 0.15f), fl, 0.017f));
    d = fmin_(d, sdCapsule(p, mk3(FA + 0.03f, WR - 0.10f, FWZ - FHW - 0.07f), mk3(1.20f, b + 0.47f, 0.16f), 0.014f));
    d = fmin_(d, sdCapsule(p, mk3(0.86f, b + 0.29f, 0.16f), mk3(0.88f, WR + 0.02f, FWZ - FHW - 0.05f), 0.013f));
    const v3 ro = mk3(RA, WR + 0.17f, RWZ - RHW - 0.05f);
    d = fmin_(d, sdCapsule(p, mk3(4.30f, b + 0.42f, 0.20f), ro, 0.017f));
    d = fmin_(d, sdCapsule(p, mk3(4.78f, b + 0.40f, 0.16f), ro, 0.017f));
    const v3 rl = mk3(RA, WR - 0.14f, RWZ - RHW - 0.05f);
    d = fmin_(d, sdCapsule(p, mk3(4.28f, b + 0.16f, 0.24f), rl, 0.017f));
    d = fmin_(d, sdCapsule(p, mk3(4.82f, b + 0.15f, 0.18f), rl, 0.017f));
    d = fmin_(d, sdCapsule(p, mk3(RA - 0.03f, WR + 0.14f, RWZ - RHW - 0.07f), mk3(4.40f, b + 0.18f, 0.22f), 0.014f));
    return d;
}
__device__ __forceinline__ uint32_t variant(float v) { return (uint32_t)(v + 0.5f); }
__device__ __forceinline__ float noseTipY(const F1Ctx& c, float b) {
    return b + (variant(c.sh->nose) == 2u ? 0.25f : 0.17f);
}
__device__ __forceinline__ float noseStartX(const F1Ctx& c) { return variant(c.sh->nose) == 1u ? 0.34f : 0.10f; }
__device__ float sdNose(const F1Ctx& c, v3 p, float b) {
    const uint32_t v = variant(c.sh->nose);
    const float x0 = noseStartX(c);
    const float y0 = noseTipY(c, b);
    v2 s0 = mk2(0.055f, 0.11f);
    if (v == 1u) s0 = mk2(0.075f, 0.15f);
    if (v == 2u) s0 = mk2(0.045f, 0.075f);
    float d = sdTaper(p, x0, 1.26f, mk2(y0, 0.0f), mk2(b + 0.36f, 0.0f), s0, mk2(0.15f, 0.17f), 0.04f, 0.12f);
    d = smin(d, sdEllipsoid(p - mk3(x0 + 0.02f, y0, 0.0f), mk3(0.09f, s0.x, s0.y)), 0.04f);
    if (v != 0u) {
        const float yw = 0.19f + b * 0.45f;
        const float pylon = sdRoundBox(p - mk3(x0 + 0.06f, 0.5f * (yw + y0), 0.09f),
                                       mk3(0.09f, 0.5f * (y0 - yw) + 0.02f, 0.006f), 0.004f);
        d = fmin_(d, pylon);
    }
    return d;
}
__device__ float sdEngineCover(const F1Ctx& c, v3 p, float b) {
    const uint32_t v = variant(c.sh->cover);
    if (v == 1u)
        return sdTaper(p, 2.96f, 4.86f, mk2(b + 0.60f, 0.0f), mk2(b + 0.40f, 0.0f), mk2(0.32f, 0.31f), mk2(0.12f, 0.13f), 0.15f, 0.08f);
    if (v == 2u)
        return sdTaper(p, 2.96f, 4.70f, mk2(b + 0.57f, 0.0f), mk2(b + 0.34f, 0.0f), mk2(0.29f, 0.19f), mk2(0.06f, 0.05f), 0.12f, 0.04f);
    return sdTaper(p, 2.96f, 4.86f, mk2(b + 0.58f, 0.0f), mk2(b + 0.36f, 0.0f), mk2(0.30f, 0.24f), mk2(0.08f, 0.07f), 0.14f, 0.06f);
}
__device__ float sdSidepods(const F1Ctx& c, v3 p, float b) {
    const uint32_t v = variant(c.sh->sidepod);
    float pod;
    if (v == 1u) {
        pod = sdTaper(p, 2.30f, 4.05f, mk2(b + 0.40f, 0.36f), mk2(b + 0.24f, 0.24f), mk2(0.16f, 0.08f), mk2(0.07f, 0.05f), 0.06f, 0.04f);
        pod = smax(pod, -sdRoundBox(p - mk3(2.27f, b + 0.42f, 0.38f), mk3(0.06f, 0.12f, 0.05f), 0.03f), 0.02f);
        return pod;
    }
    if (v == 2u) {
        pod = sdTaper(p, 2.30f, 4.25f, mk2(b + 0.41f, 0.52f), mk2(b + 0.24f, 0.30f), mk2(0.18f, 0.28f), mk2(0.08f, 0.11f), 0.10f, 0.05f);
        pod = smax(pod, -sdCapsule(p, mk3(2.30f, b + 0.13f, 0.84f), mk3(3.70f, b + 0.08f, 0.58f), 0.15f), 0.05f);
        pod = smax(pod, -sdCapsule(p, mk3(2.55f, b + 0.66f, 0.52f), mk3(4.00f, b + 0.40f, 0.36f), 0.11f), 0.04f);
        pod = smax(pod, -sdRoundBox(p - mk3(2.28f, b + 0.47f, 0.57f), mk3(0.06f, 0.10f, 0.17f), 0.05f), 0.02f);
        return pod;
    }
    if (v == 3u) {
        pod = sdTaper(p, 2.30f, 4.25f, mk2(b + 0.47f, 0.52f), mk2(b + 0.26f, 0.28f), mk2(0.15f, 0.25f), mk2(0.07f, 0.10f), 0.12f, 0.05f);
        pod = smax(pod, -sdCapsule(p, mk3(2.25f, b + 0.17f, 0.86f), mk3(3.70f, b + 0.10f, 0.56f), 0.22f), 0.05f);
        pod = smax(pod, -sdRoundBox(p - mk3(2.28f, b + 0.52f, 0.55f), mk3(0.06f, 0.07f, 0.17f), 0.04f), 0.02f);
        return pod;
    }
    pod = sdTaper(p, 2.30f, 4.25f, mk2(b + 0.41f, 0.52f), mk2(b + 0.22f, 0.28f), mk2(0.18f, 0.25f), mk2(0.07f, 0.10f), 0.12f, 0.05f);
    pod = smax(pod, -sdCapsule(p, mk3(2.30f, b + 0.13f, 0.82f), mk3(3.70f, b + 0.08f, 0.56f), 0.16f), 0.05f);
    pod = smax(pod, -sdRoundBox(p - mk3(2.28f, b + 0.47f, 0.57f), mk3(0.06f, 0.10f, 0.17f), 0.05f), 0.02f);
    return pod;
}
__device__ float sdBodywork(const F1Ctx& c, v3 p, float b) {
    float d = sdNose(c, p, b);
    d = smin(d, sdTaper(p, 1.20f, 2.24f, mk2(b + 0.36f, 0.0f), mk2(b + 0.42f, 0.0f), mk2(0.15f, 0.17f), mk2(0.25f, 0.29f), 0.12f, 0.16f), 0.08f);
    d = smin(d, sdTaper(p, 2.20f, 3.06f, mk2(b + 0.42f, 0.0f), mk2(b + 0.42f, 0.0f), mk2(0.25f, 0.30f), mk2(0.25f, 0.34f), 0.16f, 0.16f), 0.06f);
    d = smin(d, sdEngineCover(c, p, b), 0.10f);
    const float keel = sdTaper(p, 1.30f, 4.70f, mk2(b + 0.22f, 0.0f), mk2(b + 0.20f, 0.0f), mk2(0.19f, 0.19f), mk2(0.17f, 0.14f), 0.06f, 0.05f);
    d = smin(d, keel, 0.08f);
    d = smin(d, sdSidepods(c, p, b), 0.12f);
    float airbox = sdRoundBox(p - mk3(3.02f, b + 0.84f, 0.0f), mk3(0.17f, 0.11f, 0.12f), 0.08f);
    airbox = smax(airbox, -sdEllipsoid(p - mk3(2.85f, b + 0.86f, 0.0f), mk3(0.09f, 0.066f, 0.078f)), 0.015f);
    d = smin(d, airbox, 0.10f);
    d = smax(d, -sdRoundBox(p - mk3(2.62f, b + 0.72f, 0.0f), mk3(0.42f, 0.13f, 0.215f), 0.10f), 0.03f);
    return d;
}
__device__ float sdHalo(v3 p, float b) {
    float d = sdCapsule(p, mk3(2.17f, b + 0.62f, 0.0f), mk3(2.36f, b + 0.90f, 0.0f), 0.034f);
    d = smin(d, sdCapsule(p, mk3(2.36f, b + 0.90f, 0.0f), mk3(2.52f, b + 0.92f, 0.20f), 0.031f), 0.03f);
    d = smin(d, sdCapsule(p, mk3(2.52f, b + 0.92f, 0.20f), mk3(2.82f, b + 0.88f, 0.26f), 0.031f), 0.03f);
    d = smin(d, sdCapsule(p, mk3(2.82f, b + 0.88f, 0.26f), mk3(3.02f, b + 0.74f, 0.25f), 0.033f), 0.03f);
    return d;
}
__device__ float sdFrontWing(const F1Ctx& c, v3 p, float yw) {
    const float ff = c.sh->front_flap;
    const uint32_t v = variant(c.sh->front_wing);
    float d = 1e3f;
    float ep = 0.0f;
    if (v == 1u) {
        d = sdWingZ(p, mk2(-0.03f, yw), 0.31f, 2.0f * DEG + ff * 0.1f, 0.034f, 0.0f, 0.94f, 0.03f, 0.2f, 0.5f);
        d = fmin_(d, sdWingZ(p, mk2(0.21f, yw + 0.036f), 0.17f, 9.0f * DEG + ff * 0.45f, 0.024f, 0.0f, 0.94f, 0.15f, 0.35f, 0.3f));
        d = fmin_(d, sdWingZ(p, mk2(0.335f, yw + 0.075f + ff * 0.03f), 0.14f, 17.0f * DEG + ff * 0.75f, 0.020f, 0.10f, 0.94f, 0.19f, 0.35f, 0.25f));
        d = fmin_(d, sdWingZ(p, mk2(0.425f, yw + 0.118f + ff * 0.07f), 0.12f, 26.0f * DEG + ff, 0.018f, 0.12f, 0.94f, 0.22f, 0.4f, 0.25f));
        ep = sdRoundRect(mk2(p.x - 0.25f, p.y - (yw + 0.17f)), mk2(0.29f, 0.20f), 0.08f);
    } else if (v == 2u) {
        d = sdWingZ(p, mk2(-0.03f, yw), 0.31f, 3.0f * DEG + ff * 0.1f, 0.034f, 0.0f, 0.94f, -0.02f, 0.15f, 0.55f);
        d = fmin_(d, sdWingZ(p, mk2(0.21f, yw + 0.04f), 0.19f, 13.0f * DEG + ff * 0.45f, 0.024f, 0.0f, 0.94f, -0.04f, 0.35f, 0.35f));
        d = fmin_(d, sdWingZ(p, mk2(0.35f, yw + 0.085f + ff * 0.03f), 0.16f, 22.0f * DEG + ff * 0.75f, 0.020f, 0.06f, 0.94f, -0.06f, 0.4f, 0.35f));
        d = fmin_(d, sdWingZ(p, mk2(0.45f, yw + 0.13f + ff * 0.07f), 0.13f, 31.0f * DEG + ff, 0.018f, 0.08f, 0.94f, -0.08f, 0.45f, 0.35f));
        ep = sdRoundRect(mk2(p.x - 0.25f, p.y - (yw + 0.06f)), mk2(0.29f, 0.10f), 0.06f);
    } else if (v == 3u) {
        d = sdWingZ(p, mk2(-0.03f, yw), 0.30f, 1.0f * DEG + ff * 0.05f, 0.030f, 0.0f, 0.94f, 0.0f, 0.1f, 0.6f);
        d = fmin_(d, sdWingZ(p, mk2(0.20f, yw + 0.03f), 0.16f, 4.0f * DEG + ff * 0.3f, 0.022f, 0.0f, 0.94f, 0.02f, 0.2f, 0.5f));
        ep = sdRoundRect(mk2(p.x - 0.22f, p.y - (yw + 0.05f)), mk2(0.26f, 0.085f), 0.05f);
    } else {
        d = sdWingZ(p, mk2(-0.03f, yw), 0.31f, 2.0f * DEG + ff * 0.1f, 0.034f, 0.0f, 0.94f, 0.06f, 0.2f, 0.5f);
        d = fmin_(d, sdWingZ(p, mk2(0.21f, yw + 0.036f), 0.17f, 9.0f * DEG + ff * 0.45f, 0.024f, 0.0f, 0.94f, 0.09f, 0.25f, 0.45f));
        d = fmin_(d, sdWingZ(p, mk2(0.335f, yw + 0.075f + ff * 0.03f), 0.14f, 17.0f * DEG + ff * 0.75f, 0.020f, 0.10f, 0.94f, 0.11f, 0.25f, 0.4f));
        d = fmin_(d, sdWingZ(p, mk2(0.425f, yw + 0.118f + ff * 0.07f), 0.12f, 26.0f * DEG + ff, 0.018f, 0.12f, 0.94f, 0.13f, 0.25f, 0.4f));
        ep = sdRoundRect(mk2(p.x - 0.25f, p.y - (yw + 0.085f)), mk2(0.29f, 0.115f), 0.07f);
    }
    const float span = 0.94f;
    d = fmin_(d, fmax_(ep, fabs_(p.z - (span + 0.012f)) - 0.010f));
    d = fmin_(d, sdRoundBox(p - mk3(0.22f, yw - 0.008f, span - 0.04f), mk3(0.30f, 0.006f, 0.06f), 0.004f));
    return d;
}
// Rear-wing variant parameters (f1.ts:253-266); angM's cos/sin come from F1Ctx.
struct RearWingCfg { float chordM, angM, riseM, chordF, angF; v2 epH; };
__host__ __device__ inline RearWingCfg rear_wing_cfg(const aero_shape& s) {
    const uint32_t v = (uint32_t)(s.rear_wing + 0.5f);
    RearWingCfg r{0.36f, s.rear_main, -0.10f, 0.26f, s.rear_flap, mk2(0.33f, 0.29f)};
    if (v == 1u) {
        r.chordM = 0.42f; r.angM += 3.0f * DEG; r.riseM = -0.15f;
        r.chordF = 0.32f; r.angF += 6.0f * DEG;
        r.epH = mk2(0.36f, 0.33f);
    } else if (v == 2u) {
        r.chordM = 0.27f; r.angM -= 5.0f * DEG; r.riseM = 0.02f;
        r.chordF = 0.15f; r.angF *= 0.55f;
        r.epH = mk2(0.30f, 0.25f);
    }
    return r;
}
__device__ float sdRearWing(const F1Ctx& c, v3 p, float b, float yr) {
    const RearWingCfg w = rear_wing_cfg(*c.sh);
    float d = sdWingZ(p, mk2(4.98f, yr), w.chordM, w.angM, 0.05f, 0.0f, 0.50f, w.riseM, 0.12f, 0.22f);
    const float k = w.chordM - 0.01f;
    const v2 te = mk2(4.98f + c.rw_cos * k + -0.03f, yr + c.rw_sin * k + 0.035f);
    d = fmin_(d, sdWingZ(p, te, w.chordF, w.angF, 0.03f, 0.0f, 0.49f, w.riseM * 0.8f, 0.10f, 0.28f));
    d = fmin_(d, sdCapsule(p, mk3(5.26f, yr + 0.05f, 0.0f), mk3(5.30f, yr + 0.12f, 0.0f), 0.024f));
    const float ep = sdRoundRect(mk2(p.x - 5.25f, p.y - (yr - 0.03f)), w.epH, 0.13f);
    d = fmin_(d, fmax_(ep, fabs_(p.z - 0.513f) - 0.012f));
    d = fmin_(d, sdRoundBox(p - mk3(5.09f, 0.5f * (b + 0.36f + yr + 0.06f), 0.06f),
                            mk3(0.07f, 0.5f * (yr + 0.06f - b - 0.36f), 0.008f), 0.006f));
    return d;
}
__device__ R2 sdF1(const F1Ctx& c, v3 pw) {
    const v3 pworld = mk3(pw.x, pw.y, fabs_(pw.z));
    const v3 p = toBodyFrame(c, pworld);
    const float b = c.sh->ride_h;
    const float yw = 0.19f + b * 0.45f;
    const float yr = b + 0.80f;
    R2 res{1e3f, 0.0f};
    if (sdBox(p - mk3(2.72f, b + 0.52f, 0.0f), mk3(2.74f, 0.62f, 0.95f)) < res.d) {
        res = opU2(res, R2{sdBodywork(c, p, b), 0.0f});
        res = opU2(res, R2{sdEllipsoid(p - mk3(2.67f, b + 0.71f, 0.0f), mk3(0.15f, 0.13f, 0.12f)), 3.0f});
        res = opU2(res, R2{sdHalo(p, b), 4.0f});
        float mirror = sdRoundBox(p - mk3(2.32f, b + 0.67f, 0.60f), mk3(0.035f, 0.035f, 0.085f), 0.02f);
        mirror = fmin_(mirror, sdCapsule(p, mk3(2.36f, b + 0.66f, 0.53f), mk3(2.45f, b + 0.50f, 0.46f), 0.011f));
        res = opU2(res, R2{mirror, 0.0f});
        float carbon = sdSuspension(p, b);
        carbon = fmin_(carbon, sdTaper(p, 4.78f, 5.40f, mk2(b + 0.33f, 0.0f), mk2(b + 0.33f, 0.0f), mk2(0.07f, 0.065f),
                                       mk2(0.045f, 0.045f), 0.04f, 0.03f));
        carbon = fmin_(carbon, sdRoundBox(p - mk3(3.95f, b + 0.74f, 0.0f), mk3(0.55f, 0.10f, 0.008f), 0.006f));
        res = opU2(res, R2{carbon, 1.0f});
        res = opU2(res, R2{sdRoundBox(p - mk3(5.405f, b + 0.33f, 0.0f), mk3(0.012f, 0.028f, 0.05f), 0.006f), 11.0f});
        float exhaust = sdCapsule(p, mk3(4.95f, b + 0.44f, 0.0f), mk3(5.13f, b + 0.45f, 0.0f), 0.042f);
        exhaust = fmax_(exhaust, -sdCapsule(p, mk3(5.0f, b + 0.44f, 0.0f), mk3(5.2f, b + 0.45f, 0.0f), 0.028f));
        res = opU2(res, R2{exhaust, 10.0f});
    }
    if (sdBox(p - mk3(3.26f, b + 0.2f, 0.4f), mk3(1.75f, 0.22f, 0.42f)) < res.d) {
        res = opU2(res, R2{sdFloor(p, b), 1.0f});
    }
    res.d *= BODY_LIP * BODY_X;
    if (sdBox(p - mk3(0.24f, yw + 0.17f, 0.48f), mk3(0.36f, 0.34f, 0.50f)) < res.d / (BODY_LIP * BODY_X)) {
        res = opU2(res, R2{sdFrontWing(c, p, yw) * BODY_LIP * BODY_X, 7.0f});
    }
    if (sdBox(p - mk3(5.20f, b + 0.64f, 0.27f), mk3(0.46f, 0.56f, 0.29f)) < res.d / (BODY_LIP * BODY_X)) {
        res = opU2(res, R2{sdRearWing(c, p, b, yr) * BODY_LIP * BODY_X, 8.0f});
    }
    res = sdWheelSet(pworld, res, mk3(FA, WR26, FWZ * BODY_Z), FHW26, 0.10f);
    res = sdWheelSet(pworld, res, mk3(RA26, WR26, RWZ * BODY_Z), RHW26, 0.08f);
    return res;
}
// f1Emission, f1.ts:397-407.
__device__ float f1Emission(const F1Ctx& c, v3 pw) {
    const v3 p = toBodyFrame(c, mk3(pw.x, pw.y, fabs_(pw.z)));
    const float b = c.sh->ride_h;
    float e = 0.0f;
    if (length(mk2(p.x, p.y) - mk2(FA, WR)) < 0.21f && fabs_(p.z - (FWZ - FHW - 0.03f)) < 0.12f) e = fmax_(e, c.sh->heat_brakes);
    if (length(mk2(p.x, p.y) - mk2(RA, WR)) < 0.21f && fabs_(p.z - (RWZ - RHW - 0.03f)) < 0.12f) e = fmax_(e, c.sh->heat_brakes * 0.8f);
    if (p.x > 5.02f && sdCapsule(p, mk3(4.95f, b + 0.44f, 0.0f), mk3(5.16f, b + 0.45f, 0.0f), 0.09f) < 0.0f) e = fmax_(e, c.sh->heat_engine);
    if (p.x > 3.85f && p.x < 4.5f && p.z > 0.10f && p.y > b + 0.12f && p.y < b + 0.45f) e = fmax_(e, c.sh->heat_rad);
    if (p.x > 3.55f && p.x < 4.25f && p.z < 0.13f && p.y > b + 0.42f) e = fmax_(e, c.sh->heat_rad * 1.4f);
    return e;
}

}  // namespace f1
}  // namespace aero
