/*
 * CPU ORACLE — test infrastructure only (tests/, smoke() and bench.py's CPU
 * baseline load it; the product never does).
 *
 * FP32 restatement of the reference's field consumers, the readers of the
 * `fields` (ux, uy, uz, rho) and `temps` buffers that run after every frame's LBM
 * and heat steps:
 *   sampleFlow        src/engine/wgsl.ts:331-355 (cellIdx :325-329)
 *   vorticityAt       src/engine/wgsl.ts:356-363
 *   View, fieldValue  src/engine/wgsl.ts:430-457
 *   pcg, rand, spawn  src/engine/wgsl.ts:537-567
 *   PARTICLES_SIM     src/engine/wgsl.ts:517-595 (cs_main :569-594)
 *   STREAM_SIM        src/engine/render-wgsl.ts:520-576
 * Each WGSL vector expression is evaluated component by component, left to right,
 * in the order the source writes it; mix(a, b, t) is the WGSL definition
 * a*(1-t) + b*t; normalize(v) is v / length(v), length(v) = sqrt((x*x + y*y) + z*z).
 * Built with -ffp-contract=off (oracle/Makefile).
 *
 * What pins it (tests/test_flow_oracle.py): trilinear sampling reproduces a linear
 * field exactly at interior points and a uniform field everywhere; the solid-corner
 * renormalisation and the all-solid default; the z < 0 mirror (u_z changes sign);
 * the rotation field u = w x r has vorticityAt = 2 w h; the inverse of the PCG
 * permutation, derived from its three invertible steps; particles and streamlines
 * in a uniform flow advance by the closed-form step and die / respawn where the
 * domain rules say.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "sincos_r64.h"

typedef struct { uint32_t nx, ny, nz, n; float ox, oy, oz, h; } fo_grid;
typedef struct { float uIn, vReal, cpMin, cpMax, tMin, tMax, speedMax, vortMax; } fo_view;
typedef struct {
    uint32_t count, spawnMode, colorMode, frame;
    float dt, life;
    uint32_t scatter, rakeRows;
    float spawnMin[3];
    uint32_t rakeCols;
    float spawnMax[3];
    float pad;
} fo_psim;
typedef struct {
    uint32_t lines, steps, seedMode, colorMode;
    float seedMin[3];
    uint32_t rows;
    float seedMax[3];
    uint32_t cols;
    float stepLen, time;
    uint32_t perEmitter, emitters;
} fo_stream;
typedef struct { float u[3], rho, temp, fluid; } fo_sample;

typedef struct {
    const fo_grid* g;
    const float* fields; /* n x 4 */
    const float* temps;  /* n */
} fo_ctx;

static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static float len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
static float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* i32(floor value): saturating, NaN -> 0 */
static int32_t to_i32(float v) {
    if (!(v == v)) return 0;
    if (v <= -2147483648.0f) return INT32_MIN;
    if (v >= 2147483520.0f) return 2147483520;
    return (int32_t)v;
}

/* cellIdx (wgsl.ts:325-329): clamp each coordinate into the grid */
static uint32_t cell_idx(const fo_grid* g, int32_t x, int32_t y, int32_t z) {
    const int32_t dx = (int32_t)g->nx, dy = (int32_t)g->ny, dz = (int32_t)g->nz;
    x = x < 0 ? 0 : (x > dx - 1 ? dx - 1 : x);
    y = y < 0 ? 0 : (y > dy - 1 ? dy - 1 : y);
    z = z < 0 ? 0 : (z > dz - 1 ? dz - 1 : z);
    return (uint32_t)x + g->nx * ((uint32_t)y + g->ny * (uint32_t)z);
}

/* sampleFlow (wgsl.ts:331-355) */
static fo_sample sample_flow(const fo_ctx* c, float px, float py, float pz) {
    const fo_grid* g = c->g;
    const float zs = pz < 0.0f ? -1.0f : 1.0f;
    const float p[3] = {px, py, fabsf(pz)};
    const float o[3] = {g->ox, g->oy, g->oz};
    float gg[3], f[3];
    int32_t i0[3];
    for (int k = 0; k < 3; k++) {
        gg[k] = (p[k] - o[k]) / g->h - 0.5f;
        i0[k] = to_i32(floorf(gg[k]));
        f[k] = gg[k] - (float)i0[k];
    }
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f}, tacc = 0.0f, wsum = 0.0f;
    for (int k = 0; k < 8; k++) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        int32_t cx = i0[0] + ox, cy = i0[1] + oy, cz = i0[2] + oz;
        if (cz < 0) cz = -cz - 1;
        const uint32_t id = cell_idx(g, cx, cy, cz);
        const float* v = c->fields + 4 * (size_t)id;
        const float w = ((ox ? f[0] : 1.0f - f[0]) * (oy ? f[1] : 1.0f - f[1])) * (oz ? f[2] : 1.0f - f[2]);
        if (v[3] > 0.0f) {
            for (int j = 0; j < 4; j++) acc[j] = acc[j] + v[j] * w;
            tacc = tacc + c->temps[id] * w;
            wsum = wsum + w;
        }
    }
    fo_sample s;
    if (wsum < 1e-5f) {
        s.u[0] = s.u[1] = s.u[2] = 0.0f;
        s.rho = 1.0f;
        s.temp = 0.0f;
        s.fluid = 0.0f;
        return s;
    }
    s.u[0] = acc[0] / wsum;
    s.u[1] = acc[1] / wsum;
    s.u[2] = (acc[2] / wsum) * zs;
    s.rho = acc[3] / wsum;
    s.temp = tacc / wsum;
    s.fluid = wsum;
    return s;
}

/* vorticityAt (wgsl.ts:356-363) */
static float vorticity_at(const fo_ctx* c, float px, float py, float pz) {
    const float e = c->g->h;
    const fo_sample dx0 = sample_flow(c, px - e, py - 0.0f, pz - 0.0f), dx1 = sample_flow(c, px + e, py + 0.0f, pz + 0.0f);
    const fo_sample dy0 = sample_flow(c, px - 0.0f, py - e, pz - 0.0f), dy1 = sample_flow(c, px + 0.0f, py + e, pz + 0.0f);
    const fo_sample dz0 = sample_flow(c, px - 0.0f, py - 0.0f, pz - e), dz1 = sample_flow(c, px + 0.0f, py + 0.0f, pz + e);
    const float wx = ((dy1.u[2] - dy0.u[2]) - (dz1.u[1] - dz0.u[1])) * 0.5f;
    const float wy = ((dz1.u[0] - dz0.u[0]) - (dx1.u[2] - dx0.u[2])) * 0.5f;
    const float wz = ((dx1.u[1] - dx0.u[1]) - (dy1.u[0] - dy0.u[0])) * 0.5f;
    return len3(wx, wy, wz);
}

/* rhoRef, pressureCoef, adiabaticDelta, fieldValue (wgsl.ts:435-457) */
static float rho_ref(const fo_ctx* c) {
    const fo_grid* g = c->g;
    const float r = c->fields[4 * (size_t)cell_idx(g, 6, (int32_t)g->ny - 6, (int32_t)g->nz - 6) + 3];
    return r > 0.5f ? r : 1.0f;
}
static float pressure_coef(const fo_ctx* c, const fo_view* v, const fo_sample* s) {
    return ((s->rho - rho_ref(c)) / 3.0f) / ((0.5f * v->uIn) * v->uIn);
}
static float field_value(const fo_ctx* c, const fo_view* v, uint32_t id, const fo_sample* s, float px, float py,
                         float pz) {
    switch (id) {
    case 1u: return (pressure_coef(c, v, s) - v->cpMin) / (v->cpMax - v->cpMin);
    case 2u: {
        const float ad = (((pressure_coef(c, v, s) * 0.5f) * v->vReal) * v->vReal) / 1005.0f;
        const float t = s->temp + ad;
        return (t - v->tMin) / (v->tMax - v->tMin);
    }
    case 3u: return vorticity_at(c, px, py, pz) / (v->uIn * v->vortMax);
    case 4u: return clampf((1.0f - s->u[0] / v->uIn) / 0.9f, 0.0f, 1.0f);
    default: return len3(s->u[0], s->u[1], s->u[2]) / (v->uIn * v->speedMax);
    }
}

/* pcg, rand (wgsl.ts:537-544) */
uint32_t fo_pcg(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
static float rand01(uint32_t* seed) {
    *seed = fo_pcg(*seed);
    return (float)(*seed) / 4294967295.0f;
}

/* spawn (wgsl.ts:547-567); particle = pos.xyz, age, vel.xyz, colour value */
static void spawn(const fo_ctx* c, const fo_psim* ps, uint32_t i, uint32_t* seed, float* out) {
    const float* lo = ps->spawnMin;
    const float* hi = ps->spawnMax;
    float pos[3];
    if (ps->spawnMode == 1u) {
        const uint32_t lanes = ps->rakeRows * ps->rakeCols;
        const uint32_t lane = i % lanes;
        const uint32_t row = lane / ps->rakeCols, col = lane % ps->rakeCols;
        const float fy = ((float)row + 0.5f) / (float)ps->rakeRows;
        const float fz = ((float)col + 0.5f) / (float)ps->rakeCols;
        pos[0] = lo[0];
        pos[1] = mixf(lo[1], hi[1], fy);
        pos[2] = mixf(lo[2], hi[2], fz);
        const float r0 = rand01(seed);
        const float r1 = rand01(seed);
        const float r2 = rand01(seed);
        pos[0] = pos[0] + r0 * 0.12f;
        pos[1] = pos[1] + (r1 - 0.5f) * 0.012f;
        pos[2] = pos[2] + (r2 - 0.5f) * 0.012f;
    } else {
        const float r0 = rand01(seed);
        const float r1 = rand01(seed);
        const float r2 = rand01(seed);
        pos[0] = lo[0] + r0 * 0.15f;
        pos[1] = mixf(lo[1], hi[1], r1);
        pos[2] = mixf(lo[2], hi[2], r2);
    }
    if (ps->scatter != 0u) {
        const float dmx = c->g->ox + (float)c->g->nx * c->g->h;
        pos[0] = mixf(lo[0], dmx - c->g->h * 2.0f, rand01(seed));
    }
    out[0] = pos[0]; out[1] = pos[1]; out[2] = pos[2]; out[3] = 0.0f;
    out[4] = 0.0f; out[5] = 0.0f; out[6] = 0.0f; out[7] = -1.0f;
}

/* PARTICLES_SIM cs_main (wgsl.ts:569-594) for every particle i < count */
void fo_particles_step(const fo_grid* g, const float* fields, const float* temps, const fo_psim* ps,
                       const fo_view* view, float* parts) {
    const fo_ctx c = {g, fields, temps};
    const float hx = g->ox + (float)g->nx * g->h, hy = g->oy + (float)g->ny * g->h, hz = g->oz + (float)g->nz * g->h;
    for (uint32_t i = 0; i < ps->count; i++) {
        float* P = parts + 8 * (size_t)i;
        uint32_t seed = fo_pcg(i * 9781u + ps->frame * 6271u + 17u);
        float p[8];
        memcpy(p, P, sizeof p);
        const float maxLife = ps->life * (0.6f + (0.4f * (float)(fo_pcg(i * 31u) % 1000u)) / 1000.0f);
        if (ps->scatter != 0u || p[3] > maxLife) spawn(&c, ps, i, &seed, p);
        const fo_sample s1 = sample_flow(&c, p[0], p[1], p[2]);
        float mid[3];
        for (int k = 0; k < 3; k++) mid[k] = p[k] + ((s1.u[k] * g->h) * ps->dt) * 0.5f;
        const fo_sample s2 = sample_flow(&c, mid[0], mid[1], mid[2]);
        float d[3];
        for (int k = 0; k < 3; k++) d[k] = (s2.u[k] * g->h) * ps->dt;
        const float dl = len3(d[0], d[1], d[2]);
        if (dl > g->h * 2.0f) {
            const float sc = (g->h * 2.0f) / dl;
            for (int k = 0; k < 3; k++) d[k] = d[k] * sc;
        }
        float np[3];
        for (int k = 0; k < 3; k++) np[k] = p[k] + d[k];
        const int dead = np[0] > hx - g->h * 1.5f || np[1] < 0.0f || np[1] > hy || np[2] > hz || s2.fluid < 0.05f;
        if (dead) {
            spawn(&c, ps, i, &seed, p);
            p[3] = 0.0f;
            memcpy(P, p, sizeof p);
            continue;
        }
        if (np[2] < 0.0f) np[2] = -np[2];
        const fo_sample sv = sample_flow(&c, np[0], np[1], np[2]);
        P[0] = np[0]; P[1] = np[1]; P[2] = np[2]; P[3] = p[3] + 1.0f;
        P[4] = d[0]; P[5] = d[1]; P[6] = d[2];
        P[7] = clampf(field_value(&c, view, ps->colorMode, &sv, np[0], np[1], np[2]), 0.0f, 1.0f);
    }
}

/* seedOf (render-wgsl.ts:536-550) */
static void seed_of(const fo_stream* st, const float* em, uint32_t line, float* p) {
    if (st->seedMode == 1u) {
        const uint32_t last = st->emitters - 1u;
        uint32_t ei = line / st->perEmitter;
        ei = ei < last ? ei : last;
        const uint32_t k = line % st->perEmitter;
        const float* e = em + 4 * ei;
        const float a = (((float)k + 0.5f) / (float)st->perEmitter) * 6.2831853f + st->time * 0.6f;
        const float ring = 0.35f + (0.65f * (float)(k % 3u)) / 2.0f;
        float sa, ca; /* sincos_r64, as the device evaluates the ring angle */
        sincos_r64(a, &sa, &ca);
        p[0] = e[0] + (0.0f * e[3]) * ring;
        p[1] = e[1] + (ca * e[3]) * ring;
        p[2] = e[2] + (sa * e[3]) * ring;
        return;
    }
    const uint32_t row = line / st->cols, col = line % st->cols;
    const float fy = ((float)row + 0.5f) / (float)st->rows;
    const float fz = ((float)col + 0.5f) / (float)st->cols;
    p[0] = st->seedMin[0];
    p[1] = mixf(st->seedMin[1], st->seedMax[1], fy);
    p[2] = mixf(st->seedMin[2], st->seedMax[2], fz);
}

/* STREAM_SIM cs_main (render-wgsl.ts:551-575): pts[line * steps + k] = (p, value or -1) */
void fo_streamlines(const fo_grid* g, const float* fields, const float* temps, const fo_stream* st,
                    const fo_view* view, const float* emitters, float* pts) {
    const fo_ctx c = {g, fields, temps};
    const float hx = g->ox + (float)g->nx * g->h, hy = g->oy + (float)g->ny * g->h, hz = g->oz + (float)g->nz * g->h;
    for (uint32_t line = 0; line < st->lines; line++) {
        float* out = pts + 4 * (size_t)line * st->steps;
        float p[3];
        seed_of(st, emitters, line, p);
        int alive = 1;
        for (uint32_t k = 0; k < st->steps; k++) {
            float* q = out + 4 * (size_t)k;
            if (!alive) { q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; q[3] = -1.0f; continue; }
            const fo_sample s1 = sample_flow(&c, p[0], p[1], p[2]);
            if (s1.fluid < 0.3f || len3(s1.u[0], s1.u[1], s1.u[2]) < view->uIn * 0.03f) {
                alive = 0;
                q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; q[3] = -1.0f;
                continue;
            }
            q[0] = p[0]; q[1] = p[1]; q[2] = p[2];
            q[3] = clampf(field_value(&c, view, st->colorMode, &s1, p[0], p[1], p[2]), 0.0f, 1.0f);
            const float l1 = len3(s1.u[0], s1.u[1], s1.u[2]);
            float mid[3];
            for (int j = 0; j < 3; j++) mid[j] = p[j] + ((s1.u[j] / l1) * st->stepLen) * 0.5f;
            const fo_sample s2 = sample_flow(&c, mid[0], mid[1], mid[2]);
            const float* dir = s2.fluid > 0.3f ? s2.u : s1.u;
            const float l2 = len3(dir[0], dir[1], dir[2]);
            for (int j = 0; j < 3; j++) p[j] = p[j] + (dir[j] / l2) * st->stepLen;
            if (p[0] > hx - g->h * 3.0f || p[1] < 0.01f || p[1] > hy || fabsf(p[2]) > hz) alive = 0;
        }
    }
}

/* Point probes: out[8 i ...] = ux, uy, uz, rho, temp, fluid, vorticityAt, fieldValue(mode) */
void fo_sample_points(const fo_grid* g, const float* fields, const float* temps, const fo_view* view,
                      uint32_t mode, const float* pts, uint32_t n, float* out) {
    const fo_ctx c = {g, fields, temps};
    for (uint32_t i = 0; i < n; i++) {
        const float* p = pts + 3 * (size_t)i;
        const fo_sample s = sample_flow(&c, p[0], p[1], p[2]);
        float* o = out + 8 * (size_t)i;
        o[0] = s.u[0]; o[1] = s.u[1]; o[2] = s.u[2]; o[3] = s.rho; o[4] = s.temp; o[5] = s.fluid;
        o[6] = vorticity_at(&c, p[0], p[1], p[2]);
        o[7] = field_value(&c, view, mode, &s, p[0], p[1], p[2]);
    }
}
