/*
 * CPU ORACLE — test infrastructure only (see lbm_oracle.c for the rules).
 *
 * FP32 restatement of the reference's VOXELIZE kernel and the signed-distance
 * primitives it evaluates.  Reference text followed:
 *   src/engine/wgsl.ts:9-18     Shape uniform (24 scalars)
 *   src/engine/wgsl.ts:20-76    sdBox, sdRoundBox, sdCapsule, sdRoundRect, opExtrude,
 *                               sdWheel, sdTaper, sdUnevenCapsule, sdWing, opU
 *   src/engine/wgsl.ts:78-116   profileAt, sdProfile2D, sdCustom, sceneSdf, sceneEmission
 *   src/engine/wgsl.ts:294-322  VOXELIZE.cs_main
 *   src/engine/f1.ts:37-46      rakeAngle/toBodyFrame (rake rotation about the front axle)
 *
 * Kind 16 ("primitive scene") is this build's own extension: a union of the
 * reference's primitives with per-primitive material ids, evaluated in the world
 * frame or in the raked body frame, plus point heat emitters.  It describes the
 * generic F1 body of the 320x96x96 workload.
 *
 * Pinning (tests/test_oracle.py): closed-form classification of cube and sphere
 * cells away from the threshold band, the 22-cell cube extent, material grouping,
 * the x-margin rule, sdWing/sdWheel/sdRoundBox against brute-force geometric
 * distances, and the custom-profile path against an analytic circle profile.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

typedef struct { uint32_t nx, ny, nz, n; float ox, oy, oz, h; } og_grid;
typedef struct {
    uint32_t kind, profileMode, pw, ph;
    float frontFlap, rearMain, rearFlap, rideH;
    float aoa, width, axisY, pScale;
    float pOrigin[2], heatBrakes, heatEngine;
    float heatRad, frontWing, rearWing, sidepod;
    float cover, nose, rearRide, duct;
} og_shape; /* wgsl.ts:10-17, 96 bytes */

typedef struct {
    uint32_t type, material, frame, pad;
    float a[12];
} og_prim; /* 64 bytes; layout shared with include/aero_lbm.h aero_prim */

enum { PR_ROUNDBOX = 1, PR_WHEEL = 2, PR_WING = 3, PR_SPHERE = 4, PR_CAPSULE = 5, PR_CYLZ = 6, PR_EMITTER = 7 };

typedef struct { float x, y; } f2;
typedef struct { float x, y, z; } f3;

static f3 v3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static f2 v2(float x, float y) { f2 r = {x, y}; return r; }
static f3 sub3(f3 a, f3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static f2 sub2(f2 a, f2 b) { return v2(a.x - b.x, a.y - b.y); }
static float mx(float a, float b) { return a > b ? a : b; } /* WGSL max on non-NaN input */
static float mn(float a, float b) { return a < b ? a : b; }
static float clampf_(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
static float len3(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static float len2(f2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
static float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

/* wgsl.ts:21-24 */
static float sdBox(f3 p, f3 b) {
    f3 q = v3(fabsf(p.x) - b.x, fabsf(p.y) - b.y, fabsf(p.z) - b.z);
    return len3(v3(mx(q.x, 0.0f), mx(q.y, 0.0f), mx(q.z, 0.0f))) + mn(mx(q.x, mx(q.y, q.z)), 0.0f);
}
/* wgsl.ts:25 */
static float sdRoundBox(f3 p, f3 b, float r) { return sdBox(p, v3(b.x - r, b.y - r, b.z - r)) - r; }
/* wgsl.ts:26-30 */
static float sdCapsule(f3 p, f3 a, f3 b, float r) {
    f3 pa = sub3(p, a), ba = sub3(b, a);
    float hh = clampf_(dot3(pa, ba) / dot3(ba, ba), 0.0f, 1.0f);
    return len3(v3(pa.x - ba.x * hh, pa.y - ba.y * hh, pa.z - ba.z * hh)) - r;
}
/* wgsl.ts:31-34 */
static float sdRoundRect(f2 p, f2 b, float r) {
    f2 q = v2(fabsf(p.x) - b.x + r, fabsf(p.y) - b.y + r);
    return len2(v2(mx(q.x, 0.0f), mx(q.y, 0.0f))) + mn(mx(q.x, q.y), 0.0f) - r;
}
/* wgsl.ts:35-38 */
static float opExtrude(float d2, float dz) {
    f2 w = v2(d2, dz);
    return mn(mx(w.x, w.y), 0.0f) + len2(v2(mx(w.x, 0.0f), mx(w.y, 0.0f)));
}
/* wgsl.ts:39-44 */
static float sdWheel(f3 p, f3 c, float r, float hw) {
    f3 q = sub3(p, c);
    const float rr = 0.045f;
    f2 d = v2(len2(v2(q.x, q.y)) - (r - rr), fabsf(q.z) - (hw - rr));
    return mn(mx(d.x, d.y), 0.0f) + len2(v2(mx(d.x, 0.0f), mx(d.y, 0.0f))) - rr;
}
/* wgsl.ts:56-64 */
static float sdUnevenCapsule(f2 p0, float r1, float r2, float hh) {
    f2 p = v2(fabsf(p0.x), p0.y);
    float b = (r1 - r2) / hh;
    float a = sqrtf(mx(1.0f - b * b, 0.0f));
    float k = dot2(p, v2(-b, a));
    if (k < 0.0f) return len2(p) - r1;
    if (k > a * hh) return len2(sub2(p, v2(0.0f, hh))) - r2;
    return dot2(p, v2(a, b)) - r1;
}
/* wgsl.ts:66-75, with cos(ang)/sin(ang) passed in (they depend on ang only) */
static float sdWingCS(f3 p, f2 le, float chord, float c, float s, float thick, float z0, float z1) {
    f2 q = sub2(v2(p.x, p.y), le);
    float lx = q.x * c + q.y * s;
    float ly = -q.x * s + q.y * c;
    float r1 = thick * 0.5f;
    float r2 = mx(thick * 0.12f, 0.006f);
    float d2 = sdUnevenCapsule(v2(ly, lx - r1), r1, r2, mx(chord - r1 - r2, 0.01f));
    return opExtrude(d2, fabsf(p.z - 0.5f * (z0 + z1)) - 0.5f * (z1 - z0));
}
static float sdWing(f3 p, f2 le, float chord, float ang, float thick, float z0, float z1) {
    return sdWingCS(p, le, chord, cosf(ang), sinf(ang), thick, z0, z1);
}
static float sdCylZ(f3 p, float r, float hz) { /* f1.ts:60-63 */
    f2 d = v2(len2(v2(p.x, p.y)) - r, fabsf(p.z) - hz);
    return mn(mx(d.x, d.y), 0.0f) + len2(v2(mx(d.x, 0.0f), mx(d.y, 0.0f)));
}
typedef struct { float d, m; } res2;
static res2 opU(res2 a, res2 b) { return (b.d < a.d) ? b : a; } /* wgsl.ts:76 */

/* ---- custom profile, wgsl.ts:79-100 ---- */
static float profileAt(const og_shape* sh, const float* prof, int ix, int iy) {
    int x = ix < 0 ? 0 : (ix > (int)sh->pw - 1 ? (int)sh->pw - 1 : ix);
    int y = iy < 0 ? 0 : (iy > (int)sh->ph - 1 ? (int)sh->ph - 1 : iy);
    return prof[(uint32_t)y * sh->pw + (uint32_t)x];
}
static float sdProfile2D(const og_shape* sh, const float* prof, f2 q) {
    f2 g = v2((q.x - sh->pOrigin[0]) / sh->pScale - 0.5f, (q.y - sh->pOrigin[1]) / sh->pScale - 0.5f);
    f2 size = v2((float)sh->pw, (float)sh->ph);
    f2 gc = v2(clampf_(g.x, 0.0f, size.x - 1.0f), clampf_(g.y, 0.0f, size.y - 1.0f));
    int i0x = (int)floorf(gc.x), i0y = (int)floorf(gc.y);
    f2 f = v2(gc.x - (float)i0x, gc.y - (float)i0y);
    float a = mixf(profileAt(sh, prof, i0x, i0y), profileAt(sh, prof, i0x + 1, i0y), f.x);
    float c = mixf(profileAt(sh, prof, i0x, i0y + 1), profileAt(sh, prof, i0x + 1, i0y + 1), f.x);
    return mixf(a, c, f.y) + len2(sub2(g, gc)) * sh->pScale;
}
static float sdCustom(const og_shape* sh, const float* prof, f3 p) {
    if (sh->profileMode == 1u) {
        float r = len2(v2(p.y - sh->axisY, p.z));
        return sdProfile2D(sh, prof, v2(p.x, sh->axisY + r));
    }
    return opExtrude(sdProfile2D(sh, prof, v2(p.x, p.y)), fabsf(p.z) - sh->width);
}

/* ---- raked body frame, f1.ts:37-46 (rotation part) ---- */
typedef struct { float c, s, rideH; } rake_t;
static rake_t rake_of(const og_shape* sh) {
    const float FA = 1.0f, RA26 = 4.4f;
    float th = atanf((sh->rearRide - sh->rideH) / (RA26 - FA));
    rake_t r = {cosf(th), sinf(th), sh->rideH};
    return r;
}
static f3 rake_apply(rake_t rk, f3 pw) {
    const float FA = 1.0f;
    float dx = pw.x - FA, dy = pw.y - rk.rideH;
    return v3(FA + dx * rk.c + dy * rk.s, rk.rideH - dx * rk.s + dy * rk.c, pw.z);
}

/* ---- primitive scene (kind 16) ---- */
static float prim_sdf(const og_prim* q, f3 p) {
    const float* a = q->a;
    switch (q->type) {
    case PR_ROUNDBOX: return sdRoundBox(sub3(p, v3(a[0], a[1], a[2])), v3(a[3], a[4], a[5]), a[6]);
    case PR_WHEEL: return sdWheel(p, v3(a[0], a[1], a[2]), a[3], a[4]);
    case PR_WING: return sdWing(p, v2(a[0], a[1]), a[2], a[3], a[4], a[5], a[6]);
    case PR_SPHERE: return len3(sub3(p, v3(a[0], a[1], a[2]))) - a[3];
    case PR_CAPSULE: return sdCapsule(p, v3(a[0], a[1], a[2]), v3(a[3], a[4], a[5]), a[6]);
    case PR_CYLZ: return sdCylZ(sub3(p, v3(a[0], a[1], a[2])), a[3], a[4]);
    default: return 1e3f;
    }
}
static res2 scene_sdf(const og_shape* sh, const og_prim* prims, uint32_t np, f3 pw) {
    rake_t rk = rake_of(sh);
    f3 pa = v3(pw.x, pw.y, fabsf(pw.z));
    f3 pb = rake_apply(rk, pa);
    res2 r = {1e3f, 0.0f};
    for (uint32_t k = 0; k < np; k++) {
        if (prims[k].type == PR_EMITTER) continue;
        f3 p = prims[k].frame == 1u ? pb : pa;
        res2 c = {prim_sdf(&prims[k], p), (float)prims[k].material};
        r = opU(r, c);
    }
    return r;
}
static float scene_emission(const og_shape* sh, const og_prim* prims, uint32_t np, f3 pw) {
    rake_t rk = rake_of(sh);
    f3 pa = v3(pw.x, pw.y, fabsf(pw.z));
    f3 pb = rake_apply(rk, pa);
    float e = 0.0f;
    for (uint32_t k = 0; k < np; k++) {
        if (prims[k].type != PR_EMITTER) continue;
        f3 p = prims[k].frame == 1u ? pb : pa;
        const float* a = prims[k].a;
        if (len3(sub3(p, v3(a[0], a[1], a[2]))) < a[3]) e = mx(e, a[4]);
    }
    return e;
}

/* sceneSdf, wgsl.ts:101-111 (kinds 0 and 5 are not restated here) */
static res2 sceneSdf(const og_shape* sh, const float* prof, const og_prim* prims, uint32_t np, f3 p) {
    res2 r;
    switch (sh->kind) {
    case 1u: r.d = len3(sub3(p, v3(2.6f, 1.0f, 0.0f))) - 0.6f; r.m = 6.0f; return r;
    case 2u: r.d = sdRoundBox(sub3(p, v3(2.6f, 1.0f, 0.0f)), v3(0.5f, 0.5f, 0.5f), 0.03f); r.m = 6.0f; return r;
    case 3u: r.d = sdWing(p, v2(1.6f, 1.2f), 1.8f, -sh->aoa, 0.24f, -1.25f, 1.25f); r.m = 6.0f; return r;
    case 6u: r.d = 1e3f; r.m = 0.0f; return r;
    case 16u: return scene_sdf(sh, prims, np, p);
    default: r.d = sdCustom(sh, prof, p); r.m = 6.0f; return r;
    }
}
static float sceneEmission(const og_shape* sh, const og_prim* prims, uint32_t np, f3 p) {
    if (sh->kind == 16u) return scene_emission(sh, prims, np, p);
    return 0.0f;
}

/* VOXELIZE.cs_main, wgsl.ts:305-321.  Also returns the raw SDF value per cell so
 * tests can enumerate the threshold band. */
EXPORT void orc_voxelize(const og_grid* g, const og_shape* sh, const float* profile,
                         const og_prim* prims, uint32_t nprims, uint32_t* solid, float* emit,
                         float* sdf_out) {
    const uint32_t NX = g->nx, NY = g->ny;
    for (uint64_t idx = 0; idx < g->n; idx++) {
        uint32_t cx = (uint32_t)(idx % NX), cy = (uint32_t)((idx / NX) % NY), cz = (uint32_t)(idx / ((uint64_t)NX * NY));
        f3 p = v3(g->ox + ((float)cx + 0.5f) * g->h, g->oy + ((float)cy + 0.5f) * g->h,
                  g->oz + ((float)cz + 0.5f) * g->h);
        res2 r = sceneSdf(sh, profile, prims, nprims, p);
        int inside = r.d < 0.3f * g->h && cx > 1u && cx < NX - 2u;
        uint32_t group = 1u;
        if (fabsf(r.m - 7.0f) < 0.5f) group = 2u;
        if (fabsf(r.m - 8.0f) < 0.5f) group = 3u;
        solid[idx] = inside ? group : 0u;
        emit[idx] = inside ? sceneEmission(sh, prims, nprims, p) : 0.0f;
        if (sdf_out) sdf_out[idx] = r.d;
    }
}

/* Point evaluation for tests. */
EXPORT float orc_sdf_point(const og_shape* sh, const float* profile, const og_prim* prims, uint32_t np,
                           float x, float y, float z, float* material) {
    res2 r = sceneSdf(sh, profile, prims, np, v3(x, y, z));
    if (material) *material = r.m;
    return r.d;
}
EXPORT float orc_sd_wing(float px, float py, float pz, float lex, float ley, float chord, float ang,
                         float thick, float z0, float z1) {
    return sdWing(v3(px, py, pz), v2(lex, ley), chord, ang, thick, z0, z1);
}
EXPORT float orc_sd_wheel(float px, float py, float pz, float cx, float cy, float cz, float r, float hw) {
    return sdWheel(v3(px, py, pz), v3(cx, cy, cz), r, hw);
}
EXPORT float orc_sd_roundbox(float px, float py, float pz, float bx, float by, float bz, float r) {
    return sdRoundBox(v3(px, py, pz), v3(bx, by, bz), r);
}
