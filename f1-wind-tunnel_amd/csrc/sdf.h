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

// sin and cos of an f32 angle, evaluated in FP64 and rounded once to f32, so that this
// device code and the CPU oracle give the same bits where the device's cosf/sinf and a C
// library's may differ in the last one.
// Cody-Waite reduction by pi/2 with fdlibm's two-part constant (the first part has 33
// significant bits, so k * P1 is exact for |k| < 2^20), then the Taylor series of sin
// to r^19 and of cos to r^20 on |r| <= pi/4 (truncation below 2^-70), Horner order.
__host__ __device__ __forceinline__ void sincos_r64(float a, float* sn, float* cs) {
    const double x = (double)a;
    if (!(__builtin_fabs(x) <= 1e9)) {
        *sn = *cs = __builtin_nanf("");
        return;
    }
    const double kd = __builtin_rint(x * 0x1.45f306dc9c883p-1);  // x * 2/pi
    const double r = (x - kd * 0x1.921fb544p+0) - kd * 0x1.0b4611a626331p-34;
    const double r2 = r * r;
    const double ps = r2 * (-0x1.5555555555555p-3 + r2 * (0x1.1111111111111p-7 + r2 * (-0x1.a01a01a01a01ap-13 +
                      r2 * (0x1.71de3a556c734p-19 + r2 * (-0x1.ae64567f544e4p-26 + r2 * (0x1.6124613a86d09p-33 +
                      r2 * (-0x1.ae7f3e733b81fp-41 + r2 * (0x1.952c77030ad4ap-49 + r2 * -0x1.2f49b46814157p-57))))))));
    const double s = r + r * ps;
    const double pc = r2 * (-0x1.0000000000000p-1 + r2 * (0x1.5555555555555p-5 + r2 * (-0x1.6c16c16c16c17p-10 +
                      r2 * (0x1.a01a01a01a01ap-16 + r2 * (-0x1.27e4fb7789f5cp-22 + r2 * (0x1.1eed8eff8d898p-29 +
                      r2 * (-0x1.93974a8c07c9dp-37 + r2 * (0x1.ae7f3e733b81fp-45 + r2 * (-0x1.6827863b97d97p-53 +
                      r2 * 0x1.e542ba4020225p-62)))))))));
    const double c = 1.0 + pc;
    const int q = (int)((long long)kd & 3);
    const double so = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
    const double co = q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
    *sn = (float)so;
    *cs = (float)co;
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
