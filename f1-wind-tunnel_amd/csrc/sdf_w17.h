// W17 baked car (shape kind 5) for the voxelizer: the body distance grid, the
// articulated wing parts and their rig, the tyres, the sidepod cooling ducts and the
// heat sources.  Reference: src/engine/f1.ts:322-395 (w17BodyFrame, sdW17Body,
// sdW17Part, sdW17, w17Duct, w17Emission) and src/engine/car-model.ts:59-136 (rig
// hinges, angles, scale, rotation; sidepod squeeze).
//
// The grids are the shipped r16float textures (engine.ts:410-417), read as IEEE
// half words and filtered here in FP32: trilinear with clamp-to-edge addressing at
// continuous texel coordinates, the sampler of engine.ts:271.
#pragma once

#include <hip/hip_fp16.h>

#include "aero_lbm.h"
#include "sdf.h"
#include "sdf_f1.h"

namespace aero {
namespace w17 {

using namespace sdf;

constexpr float kOx = -0.3f, kH = 0.015f, kBaseRide = 0.092f;
constexpr int kNx = 368, kNy = 84, kNz = 67;

struct Ctx {
    const aero_shape* sh;
    const aero_rig* rig;       // device copy
    const uint16_t* body;      // kNx * kNy * kNz half words, x fastest
    const uint16_t* parts;     // parts atlas
    uint32_t ax, ay, az;       // atlas dims
    float rake_c, rake_s;      // cos/sin of the rake angle (f1.ts:37)
};

__device__ __forceinline__ float half_at(const uint16_t* t, int nx, int ny, int nz, int x, int y, int z) {
    x = x < 0 ? 0 : (x > nx - 1 ? nx - 1 : x);
    y = y < 0 ? 0 : (y > ny - 1 ? ny - 1 : y);
    z = z < 0 ? 0 : (z > nz - 1 ? nz - 1 : z);
    const uint16_t h = t[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z)];
    return __half2float(__ushort_as_half(h));
}
__device__ __forceinline__ float tex3(const uint16_t* t, int nx, int ny, int nz, float u, float v, float w) {
    const float fu = __builtin_floorf(u), fv = __builtin_floorf(v), fw = __builtin_floorf(w);
    const int x0 = (int)fu, y0 = (int)fv, z0 = (int)fw;
    const float a = u - fu, b = v - fv, c = w - fw;
    const float c00 = mix(half_at(t, nx, ny, nz, x0, y0, z0), half_at(t, nx, ny, nz, x0 + 1, y0, z0), a);
    const float c10 = mix(half_at(t, nx, ny, nz, x0, y0 + 1, z0), half_at(t, nx, ny, nz, x0 + 1, y0 + 1, z0), a);
    const float c01 = mix(half_at(t, nx, ny, nz, x0, y0, z0 + 1), half_at(t, nx, ny, nz, x0 + 1, y0, z0 + 1), a);
    const float c11 =
        mix(half_at(t, nx, ny, nz, x0, y0 + 1, z0 + 1), half_at(t, nx, ny, nz, x0 + 1, y0 + 1, z0 + 1), a);
    return mix(mix(c00, c10, b), mix(c01, c11, b), c);
}

// ---- rig (car-model.ts:71-113) ----
__device__ __forceinline__ float rig_weight(uint32_t part, float z) {
    if (part == 1u || part == 2u) {
        const float az = fabs_(z);
        return smoothstep(0.16f, 0.22f, az) * (1.0f - smoothstep(0.5f, 0.56f, az));
    }
    return (part == 3u || part == 4u) ? 1.0f : 0.0f;
}
__device__ __forceinline__ v2 rig_hinge(const aero_rig& r, uint32_t part, float z) {
    if (part == 3u) return mk2(r.hinges[2], r.hinges[3]);
    if (part == 4u) return mk2(r.hinges[0], r.hinges[1]);
    const float f = clamp_(fabs_(z) / 0.01f, 0.0f, 70.0f);
    const uint32_t i = (uint32_t)__builtin_floorf(f);
    const uint32_t j = i + 1u < 70u ? i + 1u : 70u;
    const float t = f - (float)i;
    const float* a = r.front[i];
    const float* b = r.front[j];
    return part == 1u ? mk2(mix(a[0], b[0], t), mix(a[1], b[1], t)) : mk2(mix(a[2], b[2], t), mix(a[3], b[3], t));
}
__device__ __forceinline__ float rig_angle(const aero_rig& r, uint32_t part, float z) {
    float a = 0.0f;
    if (part == 1u) a = r.angles[0];
    else if (part == 2u) a = r.angles[1];
    else if (part == 3u) a = r.angles[2];
    else if (part == 4u) a = r.angles[3];
    if (part == 3u || part == 4u) a += r.extra[0] * smoothstep(0.12f, 0.48f, fabs_(z));
    return a * rig_weight(part, z);
}
__device__ __forceinline__ float rig_scale(const aero_rig& r, uint32_t part) {
    if (part == 3u) return r.extra[1];
    if (part == 1u || part == 2u) return 1.0f - (1.0f - r.extra[1]) * 0.65f;
    return 1.0f;
}
// rigRotate(p, part, -1): from the posed part back into its rest grid.
__device__ __forceinline__ v3 rig_rotate_inv(const aero_rig& r, v3 p, uint32_t part) {
    if (part == 0u || part > 4u) return p;
    const float a = -rig_angle(r, part, p.z);
    const float k = rig_scale(r, part);
    const v2 h = rig_hinge(r, part, p.z);
    float s, c;
    sincos_r64(a, &s, &c);
    v2 d = mk2(p.x - h.x, p.y - h.y);
    d = mk2(c * d.x - s * d.y, s * d.x + c * d.y);
    d = mk2(d.x / k, d.y / k);
    return mk3(h.x + d.x, h.y + d.y, p.z);
}
// Sidepod squeeze (car-model.ts:122-135)
__device__ __forceinline__ float pod_weight(v3 p) {
    return smoothstep(1.85f, 2.25f, p.x) * (1.0f - smoothstep(3.55f, 4.05f, p.x)) * smoothstep(0.12f, 0.2f, p.y) *
           (1.0f - smoothstep(0.6f, 0.7f, p.y)) * smoothstep(0.26f, 0.4f, fabs_(p.z));
}
__device__ __forceinline__ v3 pod_deform_inv(const aero_rig& r, v3 p) {
    const float c = r.extra[2] * pod_weight(p);
    if (c <= 0.0f) return p;
    const float az = fabs_(p.z);
    const float kz = 1.0f / (1.0f - c);
    const float ky = 1.0f / (1.0f - 0.45f * c);
    const float nz = az > 0.26f ? 0.26f + (az - 0.26f) * kz : az;
    const float sg = p.z > 0.0f ? 1.0f : (p.z < 0.0f ? -1.0f : 0.0f);
    return mk3(p.x, 0.12f + (p.y - 0.12f) * ky, sg * nz);
}

// ---- car (f1.ts:330-395) ----
__device__ __forceinline__ v3 body_frame(const Ctx& c, v3 pw) {
    const float dx = pw.x - 1.0f, dy = pw.y - c.sh->ride_h;
    return mk3(1.0f + dx * c.rake_c + dy * c.rake_s, kBaseRide - dx * c.rake_s + dy * c.rake_c, pw.z);
}
__device__ __forceinline__ float sd_body(const Ctx& c, v3 p) {
    const v3 half = mk3(kNx * kH * 0.5f, kNy * kH * 0.5f, kNz * kH * 0.5f);
    const float box = sdBox(p - mk3(kOx + half.x, half.y, half.z), half);
    if (box > 0.0f) return box + 0.07f;
    return tex3(c.body, kNx, kNy, kNz, (p.x - kOx) / kH, p.y / kH, p.z / kH);
}
__device__ __forceinline__ float sd_part(const Ctx& c, v3 pb, uint32_t part) {
    const float* g0 = c.rig->grids[(part - 1u) * 2u];
    const float* g1 = c.rig->grids[(part - 1u) * 2u + 1u];
    const v3 q = rig_rotate_inv(*c.rig, pb, part);
    const v3 half = mk3((g1[0] - 1.0f) * g0[3] * 0.5f, (g1[1] - 1.0f) * g0[3] * 0.5f, (g1[2] - 1.0f) * g0[3] * 0.5f);
    const float box = sdBox(q - mk3(g0[0] + half.x, g0[1] + half.y, g0[2] + half.z), half);
    if (box > 0.0f) return box + 0.03f;
    const v3 cell = mk3((q.x - g0[0]) / g0[3], (q.y - g0[1]) / g0[3], (q.z - g0[2]) / g0[3]);
    return tex3(c.parts, (int)c.ax, (int)c.ay, (int)c.az, cell.x, cell.y, cell.z + g1[3]) *
           fmin_(rig_scale(*c.rig, part), 1.0f);
}

struct R2 { float d, m; };
__device__ __forceinline__ R2 opU(R2 a, R2 b) { return (b.d < a.d) ? b : a; }

__device__ inline R2 sdW17(const Ctx& c, v3 pw) {
    const v3 front = mk3(1.0f, 0.36f, 0.779f), rear = mk3(4.4f, 0.351f, 0.737f);
    const v3 pa = mk3(pw.x, pw.y, fabs_(pw.z));
    const v3 p = body_frame(c, pa);
    const aero_rig& r = *c.rig;
    const float squeeze = 1.0f - r.extra[2] * pod_weight(p);
    R2 res = {sd_body(c, pod_deform_inv(r, p)) * 0.95f * squeeze, 12.0f};
    if (r.fly[2] > 0.5f && p.x < 0.85f && p.y < 0.6f)
        res = opU(res, R2{fmin_(fmin_(sd_part(c, p, 1u), sd_part(c, p, 2u)), sd_part(c, p, 5u)) * 0.95f, 7.0f});
    if (r.fly[3] > 0.5f && p.x > 4.4f && p.y > 0.5f)
        res = opU(res, R2{fmin_(fmin_(sd_part(c, p, 3u), sd_part(c, p, 4u)), sd_part(c, p, 6u)) * 0.95f, 8.0f});
    res = opU(res, R2{fmin_(f1::sdTyre(pa, front, 0.175f), sdCylZ(pa - front, 0.234f, 0.15f)), 2.0f});
    res = opU(res, R2{fmin_(f1::sdTyre(pa, rear, 0.205f), sdCylZ(pa - rear, 0.234f, 0.18f)), 2.0f});
    return res;
}
__device__ inline float w17Duct(const Ctx& c, v3 pw) {
    if (c.sh->duct <= 0.0f) return 1e3f;
    const v3 p = pod_deform_inv(*c.rig, body_frame(c, mk3(pw.x, pw.y, fabs_(pw.z))));
    const float k = c.sh->duct;
    const float mouth = sdCapsule(p, mk3(2.05f, 0.52f, 0.58f), mk3(3.0f, 0.50f, 0.48f), 0.09f * k);
    const float mainc = sdCapsule(p, mk3(2.05f, 0.55f, 0.47f), mk3(3.25f, 0.46f, 0.43f), 0.1f * k);
    const float exitc = sdCapsule(p, mk3(3.25f, 0.46f, 0.43f), mk3(4.1f, 0.40f, 0.32f), 0.085f * k);
    return fmin_(fmin_(mouth, mainc), exitc);
}
__device__ inline float w17Emission(const Ctx& c, v3 pw) {
    const v3 front = mk3(1.0f, 0.36f, 0.779f), rear = mk3(4.4f, 0.351f, 0.737f);
    const v3 pa = mk3(pw.x, pw.y, fabs_(pw.z));
    const v3 p = body_frame(c, pa);
    const aero_shape& s = *c.sh;
    float e = 0.0f;
    if (length(mk2(pa.x - front.x, pa.y - front.y)) < 0.2f && fabs_(pa.z - (front.z - 0.215f)) < 0.1f)
        e = fmax_(e, s.heat_brakes);
    if (length(mk2(pa.x - rear.x, pa.y - rear.y)) < 0.2f && fabs_(pa.z - (rear.z - 0.245f)) < 0.1f)
        e = fmax_(e, s.heat_brakes * 0.8f);
    if (length(p - mk3(5.07f, 0.53f, 0.0f)) < 0.1f) e = fmax_(e, s.heat_engine);
    if (p.x > 3.7f && p.x < 4.3f && p.y > 0.2f && p.y < 0.5f && p.z > 0.18f && p.z < 0.48f) e = fmax_(e, s.heat_rad);
    if (p.x > 4.15f && p.x < 4.65f && p.y > 0.5f && p.z < 0.16f) e = fmax_(e, s.heat_rad * 1.4f);
    return e;
}

}  // namespace w17
}  // namespace aero
