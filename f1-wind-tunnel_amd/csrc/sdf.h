// Signed-distance primitives evaluated by the voxelizer (device and host).
// Reference definitions: src/engine/wgsl.ts:20-76 and src/engine/f1.ts:48-63.
// All arithmetic is f32 in the reference's operation order; the library is built
// with -ffp-contract=off so these match the reference's WGSL expressions.
#pragma once
#include <hip/hip_runtime.h>

namespace aero {
namespace sdf {

struct v2 { float x, y; };
struct v3 { float x, y, z; };
__host__ __device__ __forceinline__ v2 mk2(float x, float y) { return {x, y}; }
__host__ __device__ __forceinline__ v3 mk3(float x, float y, float z) { return {x, y, z}; }
__host__ __device__ __forceinline__ v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ v2 operator-(v2 a, v2 b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ float fmin_(float a, float b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ float fabs_(float a) { return __builtin_fabsf(a); }
__host__ __device__ __forceinline__ float clamp_(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
__host__ __device__ __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
__host__ __device__ __forceinline__ float length(v3 a) { return sqrt_(a.x * a.x + a.y * a.y + a.z * a.z); }
__host__ __device__ __forceinline__ float length(v2 a) { return sqrt_(a.x * a.x + a.y * a.y); }
__host__ __device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ float dot(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }
__host__ __device__ __forceinline__ float mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__host__ __device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
    const float t = clamp_((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

__host__ __device__ __forceinline__ float sdBox(v3 p, v3 b) {
    const v3 q = mk3(fabs_(p.x) - b.x, fabs_(p.y) - b.y, fabs_(p.z) - b.z);
    return length(mk3(fmax_(q.x, 0.0f), fmax_(q.y, 0.0f), fmax_(q.z, 0.0f))) + fmin_(fmax_(q.x, fmax_(q.y, q.z)), 0.0f);
}
__host__ __device__ __forceinline__ float sdRoundBox(v3 p, v3 b, float r) {
    return sdBox(p, mk3(b.x - r, b.y - r, b.z - r)) - r;
}
__host__ __device__ __forceinline__ float sdCapsule(v3 p, v3 a, v3 b, float r) {
    const v3 pa = p - a, ba = b - a;
    const float hh = clamp_(dot(pa, ba) / dot(ba, ba), 0.0f, 1.0f);
    return length(mk3(pa.x - ba.x * hh, pa.y - ba.y * hh, pa.z - ba.z * hh)) - r;
}
__host__ __device__ __forceinline__ float sdRoundRect(v2 p, v2 b, float r) {
    const v2 q = mk2(fabs_(p.x) - b.x + r, fabs_(p.y) - b.y + r);
    return length(mk2(fmax_(q.x, 0.0f), fmax_(q.y, 0.0f))) + fmin_(fmax_(q.x, q.y), 0.0f) - r;
}
__host__ __device__ __forceinline__ float opExtrude(float d2, float dz) {
    return fmin_(fmax_(d2, dz), 0.0f) + length(mk2(fmax_(d2, 0.0f), fmax_(dz, 0.0f)));
}
__host__ __device__ __forceinline__ float sdWheel(v3 p, v3 c, float r, float hw) {
    const v3 q = p - c;
    const float rr = 0.045f;
    const v2 d = mk2(length(mk2(q.x, q.y)) - (r - rr), fabs_(q.z) - (hw - rr));
    return fmin_(fmax_(d.x, d.y), 0.0f) + length(mk2(fmax_(d.x, 0.0f), fmax_(d.y, 0.0f))) - rr;
}
__host__ __device__ __forceinline__ float sdTaper(v3 p, float x0, float x1, v2 c0, v2 c1, v2 s0, v2 s1, float r0,
                                                  float r1) {
    const float t = clamp_((p.x - x0) / (x1 - x0), 0.0f, 1.0f);
    const v2 c = mk2(mix(c0.x, c1.x, t), mix(c0.y, c1.y, t));
    const v2 s = mk2(mix(s0.x, s1.x, t), mix(s0.y, s1.y, t));
    const float r = fmin_(mix(r0, r1, t), fmin_(s.x, s.y));
    const float d2 = sdRoundRect(mk2(p.y, p.z) - c, s, r);
    const float dx = fabs_(p.x - 0.5f * (x0 + x1)) - 0.5f * (x1 - x0);
    return opExtrude(d2, dx) * 0.8f;
}
__host__ __device__ __forceinline__ float sdUnevenCapsule(v2 p0, float r1, float r2, float hh) {
    const v2 p = mk2(fabs_(p0.x), p0.y);
    const float b = (r1 - r2) / hh;
    const float a = sqrt_(fmax_(1.0f - b * b, 0.0f));
    const float k = dot(p, mk2(-b, a));
    if (k < 0.0f) return length(p) - r1;
    if (k > a * hh) return length(p - mk2(0.0f, hh)) - r2;
    return dot(p, mk2(a, b)) - r1;
}
// sdWing with cos(ang), sin(ang) supplied by the caller.
__host__ __device__ __forceinline__ float sdWingCS(v3 p, v2 le, float chord, float c, float s, float thick,
                                                   float z0, float z1) {
    const v2 q = mk2(p.x, p.y) - le;
    const float lx = q.x * c + q.y * s;
    const float ly = -q.x * s + q.y * c;
    const float r1 = thick * 0.5f;
    const float r2 = fmax_(thick * 0.12f, 0.006f);
    const float d2 = sdUnevenCapsule(mk2(ly, lx - r1), r1, r2, fmax_(chord - r1 - r2, 0.01f));
    return opExtrude(d2, fabs_(p.z - 0.5f * (z0 + z1)) - 0.5f * (z1 - z0));
}
__host__ __device__ __forceinline__ float sdCylZ(v3 p, float r, float hz) {
    const v2 d = mk2(length(mk2(p.x, p.y)) - r, fabs_(p.z) - hz);
    return fmin_(fmax_(d.x, d.y), 0.0f) + length(mk2(fmax_(d.x, 0.0f), fmax_(d.y, 0.0f)));
}

}  // namespace sdf
}  // namespace aero
