// SDF voxelizer for gfx950 (reference behaviour: VOXELIZE, src/engine/wgsl.ts:294-322,
// scene dispatch wgsl.ts:78-116).  One thread per cell centre; writes the solid
// group and the heat emission.  Angles that depend only on the shape (the wing
// angle of kind 3, the rake, the wing elements of a primitive scene) are turned
// into cos/sin on the host once per call.
#include "aero_internal.h"
#include "sdf.h"
#include "sdf_f1.h"
#include "sdf_w17.h"

namespace aero {

using namespace sdf;

struct VoxArgs {
    aero_shape shape;
    const float* profile;
    const aero_prim* prims;
    uint32_t nprims;
    uint32_t nx, ny, nz;
    float ox, oy, oz, h;
    float rake_c, rake_s;   // cos/sin of the rake angle (f1.ts:37)
    float wing_c, wing_s;   // cos/sin(-aoa) for kind 3
    float rw_cos, rw_sin;   // cos/sin of the rear main-plane angle (kind 0)
    const uint16_t* car_body;   // W17 (kind 5): body grid, parts atlas, rig
    const uint16_t* car_parts;
    uint32_t atlas[3];
    const aero_rig* rig;
    uint32_t* solid;
    float* emit;
};

struct Res { float d, m; };
__device__ __forceinline__ Res opU(Res a, Res b) { return (b.d < a.d) ? b : a; }

// ---- custom 2-D profile (wgsl.ts:79-100) ----
__device__ float profile_at(const VoxArgs& a, int ix, int iy) {
    const int x = ix < 0 ? 0 : (ix > (int)a.shape.pw - 1 ? (int)a.shape.pw - 1 : ix);
    const int y = iy < 0 ? 0 : (iy > (int)a.shape.ph - 1 ? (int)a.shape.ph - 1 : iy);
    return a.profile[(uint32_t)y * a.shape.pw + (uint32_t)x];
}
__device__ float sd_profile2d(const VoxArgs& a, v2 q) {
    const aero_shape& s = a.shape;
    const v2 g = mk2((q.x - s.p_origin[0]) / s.p_scale - 0.5f, (q.y - s.p_origin[1]) / s.p_scale - 0.5f);
    const v2 gc = mk2(clamp_(g.x, 0.0f, (float)s.pw - 1.0f), clamp_(g.y, 0.0f, (float)s.ph - 1.0f));
    const int i0x = (int)__builtin_floorf(gc.x), i0y = (int)__builtin_floorf(gc.y);
    const v2 f = mk2(gc.x - (float)i0x, gc.y - (float)i0y);
    const float p00 = profile_at(a, i0x, i0y), p10 = profile_at(a, i0x + 1, i0y);
    const float p01 = profile_at(a, i0x, i0y + 1), p11 = profile_at(a, i0x + 1, i0y + 1);
    const float ra = mix(p00, p10, f.x), rc = mix(p01, p11, f.x);
    return mix(ra, rc, f.y) + length(g - gc) * s.p_scale;
}
__device__ float sd_custom(const VoxArgs& a, v3 p) {
    if (a.shape.profile_mode == 1u) {
        const float r = length(mk2(p.y - a.shape.axis_y, p.z));
        return sd_profile2d(a, mk2(p.x, a.shape.axis_y + r));
    }
    return opExtrude(sd_profile2d(a, mk2(p.x, p.y)), fabs_(p.z) - a.shape.width);
}

// ---- primitive scene (kind 16) ----
__device__ __forceinline__ v3 rake_apply(const VoxArgs& a, v3 pw) {
    const float FA = 1.0f;
    const float dx = pw.x - FA, dy = pw.y - a.shape.ride_h;
    return mk3(FA + dx * a.rake_c + dy * a.rake_s, a.shape.ride_h - dx * a.rake_s + dy * a.rake_c, pw.z);
}
__device__ float prim_sdf(const aero_prim& q, v3 p) {
    const float* a = q.a;
    switch (q.type) {
    case 1: return sdRoundBox(p - mk3(a[0], a[1], a[2]), mk3(a[3], a[4], a[5]), a[6]);
    case 2: return sdWheel(p, mk3(a[0], a[1], a[2]), a[3], a[4]);
    case 3: return sdWingCS(p, mk2(a[0], a[1]), a[2], a[10], a[11], a[4], a[5], a[6]);
    case 4: return length(p - mk3(a[0], a[1], a[2])) - a[3];
    case 5: return sdCapsule(p, mk3(a[0], a[1], a[2]), mk3(a[3], a[4], a[5]), a[6]);
    case 6: return sdCylZ(p - mk3(a[0], a[1], a[2]), a[3], a[4]);
    default: return 1e3f;
    }
}
__device__ Res scene_sdf(const VoxArgs& a, v3 pw, float* emission) {
    const v3 pa = mk3(pw.x, pw.y, fabs_(pw.z));
    const v3 pb = rake_apply(a, pa);
    Res r = {1e3f, 0.0f};
    float e = 0.0f;
    for (uint32_t k = 0; k < a.nprims; k++) {
        const aero_prim& q = a.prims[k];
        const v3 p = q.frame == 1u ? pb : pa;
        if (q.type == 7u) {
            if (length(p - mk3(q.a[0], q.a[1], q.a[2])) < q.a[3]) e = fmax_(e, q.a[4]);
            continue;
        }
        r = opU(r, Res{prim_sdf(q, p), (float)q.material});
    }
    *emission = e;
    return r;
}

__device__ __forceinline__ w17::Ctx w17_ctx(const VoxArgs& a) {
    return w17::Ctx{&a.shape, a.rig, a.car_body, a.car_parts, a.atlas[0], a.atlas[1], a.atlas[2], a.rake_c, a.rake_s};
}

// sceneSdf (wgsl.ts:101-111) and sceneEmission (wgsl.ts:112-116).
__device__ Res scene(const VoxArgs& a, v3 p, float* emission) {
    *emission = 0.0f;
    switch (a.shape.kind) {
    case AERO_KIND_F1: {
        const f1::F1Ctx c{&a.shape, a.rake_c, a.rake_s, a.rw_cos, a.rw_sin, 0.0f};
        const f1::R2 r = f1::sdF1(c, p);
        *emission = f1::f1Emission(c, p);
        return Res{r.d, r.m};
    }
    case AERO_KIND_SPHERE: return Res{length(p - mk3(2.6f, 1.0f, 0.0f)) - 0.6f, 6.0f};
    case AERO_KIND_CUBE: return Res{sdRoundBox(p - mk3(2.6f, 1.0f, 0.0f), mk3(0.5f, 0.5f, 0.5f), 0.03f), 6.0f};
    case AERO_KIND_WING:
        return Res{sdWingCS(p, mk2(1.6f, 1.2f), 1.8f, a.wing_c, a.wing_s, 0.24f, -1.25f, 1.25f), 6.0f};
    case AERO_KIND_EMPTY: return Res{1e3f, 0.0f};
    case AERO_KIND_SCENE: return scene_sdf(a, p, emission);
    case AERO_KIND_W17: {
        const w17::Ctx c = w17_ctx(a);
        const w17::R2 r = w17::sdW17(c, p);
        *emission = w17::w17Emission(c, p);
        return Res{r.d, r.m};
    }
    default: return Res{sd_custom(a, p), 6.0f};
    }
}

__global__ void __launch_bounds__(256) k_voxelize(VoxArgs a) {
    const uint32_t cx = blockIdx.x * 64 + (threadIdx.x & 63);
    const uint32_t cy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const uint32_t cz = blockIdx.z;
    if (cx >= a.nx || cy >= a.ny) return;
    const uint64_t idx = (uint64_t)cx + (uint64_t)a.nx * ((uint64_t)cy + (uint64_t)a.ny * cz);
    const v3 p = mk3(a.ox + ((float)cx + 0.5f) * a.h, a.oy + ((float)cy + 0.5f) * a.h, a.oz + ((float)cz + 0.5f) * a.h);
    float em = 0.0f;
    Res r = scene(a, p, &em);
    // W17: the sidepod cooling ducts are carved out of the body (material 12) and
    // the car is voxelized with a wider band (wgsl.ts:313-314,319-320).
    const bool is_w17 = a.shape.kind == AERO_KIND_W17;
    float duct = 1e3f;
    if (is_w17) {
        duct = w17::w17Duct(w17_ctx(a), p);
        if (fabs_(r.m - 12.0f) < 0.5f) r.d = fmax_(r.d, -duct);
    }
    const bool inside = r.d < (is_w17 ? 0.45f : 0.3f) * a.h && cx > 1u && cx < a.nx - 2u;
    uint32_t group = 1u;
    if (fabs_(r.m - 7.0f) < 0.5f) group = 2u;
    if (fabs_(r.m - 8.0f) < 0.5f) group = 3u;
    a.solid[idx] = inside ? group : 0u;
    a.emit[idx] = inside ? em : (duct < 0.0f ? -1.0f : 0.0f);
}

hipError_t launch_voxelize(aero_ctx* c, const aero_shape* s, const float* d_profile, const aero_prim* d_prims,
                           uint32_t nprims, const float* host_extra) {
    VoxArgs a{};
    a.shape = *s;
    a.profile = d_profile;
    a.prims = d_prims;
    a.nprims = nprims;
    a.nx = c->grid.nx; a.ny = c->grid.ny; a.nz = c->grid.nz;
    a.ox = c->grid.ox; a.oy = c->grid.oy; a.oz = c->grid.oz; a.h = c->grid.h;
    a.rake_c = host_extra[0]; a.rake_s = host_extra[1];
    a.wing_c = host_extra[2]; a.wing_s = host_extra[3];
    a.rw_cos = host_extra[4]; a.rw_sin = host_extra[5];
    a.car_body = c->car_body;
    a.car_parts = c->car_parts;
    a.atlas[0] = c->parts_dims[0]; a.atlas[1] = c->parts_dims[1]; a.atlas[2] = c->parts_dims[2];
    a.rig = c->d_rig;
    a.solid = c->solid;
    a.emit = c->emit;
    const dim3 grid((c->grid.nx + 63) / 64, (c->grid.ny + 3) / 4, c->grid.nz);
    hipLaunchKernelGGL(k_voxelize, grid, dim3(256), 0, c->stream, a);
    return hipGetLastError();
}

}  // namespace aero
