/*
 * CPU ORACLE — test infrastructure only.
 *
 * A literal C restatement of the reference's WGSL simulation kernels, used as the
 * checker for the HIP library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; no product code path calls it.
 *
 * Compiled twice:
 *   -DREAL=float  -ffp-contract=off  -> FP32 mirror of the WGSL arithmetic
 *                                        (same constants, same evaluation order)
 *   -DREAL=double                     -> the FP64 oracle (exact rational weights)
 *
 * Reference text followed (paths relative to the reference repository):
 *   src/engine/wgsl.ts:5-7      Grid uniform
 *   src/engine/wgsl.ts:123-139  D3Q19 tables CX/CY/CZ/OPP/ZMIR/WT and feq()
 *   src/engine/wgsl.ts:141-155  LBM_INIT.cs_main
 *   src/engine/wgsl.ts:157-292  LBM_STEP.stepCell + cs_main force reduction
 *   src/engine/wgsl.ts:460-511  HEAT.cs_main
 *
 * Evaluation order follows WGSL's left-to-right associativity; WGSL abstract-float
 * literals are materialised to REAL at their use (R()).  mix() and smoothstep()
 * follow the WGSL specification formulas (e1*(1-e3)+e2*e3 and t*t*(3-2t)).
 * WGSL round() is round-half-to-even (rint).
 *
 * Pinning (see tests/test_oracle.py): lattice moment identities, feq moments,
 * the exact uniform-flow fixed point of the whole step (inlet, outlet, moving
 * floor, far field and symmetry plane together), mass/momentum conservation of
 * the collision, the half-tunnel == mirrored full-tunnel identity, hand-summed
 * single-cell momentum exchange, the stability guard reset, and closed-form
 * heat cases (uniform cooling, integer-shift advection, solid copy).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif
#define R(x) ((REAL)(x))

#define EXPORT __attribute__((visibility("default")))

typedef struct { uint32_t nx, ny, nz, n; float ox, oy, oz, h; } og_grid; /* wgsl.ts:6 */
typedef struct { float uIn, tau0, smag, pad; } og_sim;                   /* wgsl.ts:142 */
typedef struct { float dt, cool, diffuse, pad; } og_heat;                /* wgsl.ts:461 */

typedef struct { REAL x, y, z; } v3;

/* wgsl.ts:124-133 */
static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const uint32_t OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
static const uint32_t ZMIR[19] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};
#define W0 (1.0 / 3.0)
#define W1 (1.0 / 18.0)
#define W2 (1.0 / 36.0)
static const REAL WT[19] = {R(W0), R(W1), R(W1), R(W1), R(W1), R(W1), R(W1),
                            R(W2), R(W2), R(W2), R(W2), R(W2), R(W2),
                            R(W2), R(W2), R(W2), R(W2), R(W2), R(W2)};

static REAL dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* Square root in REAL: libm's sqrt returns double, which would silently lift the
 * surrounding f32 expression to double.  sqrt of a float computed in double and
 * rounded once to float is the correctly rounded f32 square root. */
static REAL sqrt_r(REAL x) { return (REAL)sqrt((double)x); }

/* wgsl.ts:135-138 */
static REAL feq(uint32_t i, REAL rho, v3 u) {
    REAL cu = (REAL)CX[i] * u.x + (REAL)CY[i] * u.y + (REAL)CZ[i] * u.z;
    return WT[i] * rho * (R(1.0) + R(3.0) * cu + R(4.5) * cu * cu - R(1.5) * dot3(u, u));
}

static REAL clampr(REAL x, REAL lo, REAL hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* WGSL smoothstep: t = clamp((x - e0) / (e1 - e0), 0, 1); t * t * (3 - 2t) */
static REAL smoothstep_r(REAL e0, REAL e1, REAL x) {
    REAL t = clampr((x - e0) / (e1 - e0), R(0.0), R(1.0));
    return t * t * (R(3.0) - R(2.0) * t);
}

/* wgsl.ts:176: exponent-all-ones test on the f32 bit pattern (inf or NaN) */
static int is_bad(REAL x) { return !isfinite(x); }

EXPORT int orc_real_size(void) { return (int)sizeof(REAL); }

EXPORT REAL orc_feq(uint32_t i, REAL rho, REAL ux, REAL uy, REAL uz) {
    v3 u = {ux, uy, uz};
    return feq(i, rho, u);
}

EXPORT void orc_tables(int32_t* cx, int32_t* cy, int32_t* cz, uint32_t* opp, uint32_t* zmir, REAL* wt) {
    for (int i = 0; i < 19; i++) {
        cx[i] = CX[i]; cy[i] = CY[i]; cz[i] = CZ[i];
        opp[i] = OPP[i]; zmir[i] = ZMIR[i]; wt[i] = WT[i];
    }
}

/* LBM_INIT.cs_main, wgsl.ts:147-154 */
EXPORT void orc_init(const og_grid* g, const og_sim* s, REAL* fa, REAL* fields) {
    const uint64_t n = g->n;
    v3 u = {(REAL)s->uIn, R(0.0), R(0.0)};
    for (uint64_t idx = 0; idx < n; idx++) {
        for (uint32_t i = 0; i < 19; i++) fa[i * n + idx] = feq(i, R(1.0), u);
        if (fields) {
            fields[idx * 4 + 0] = u.x; fields[idx * 4 + 1] = u.y;
            fields[idx * 4 + 2] = u.z; fields[idx * 4 + 3] = R(1.0);
        }
    }
}

/* Per-step diagnostics returned to the caller. */
typedef struct {
    int64_t forces_q[12];  /* reference i32 fixed-point sums (scale 1e4), slot g*3+k */
    double forces_d[9];    /* FP64 sums of the per-cell force, slot g*3+k (k = x,y,z) */
    uint64_t guard_hits;   /* cells where the stability guard fired */
} og_stats;

/* One fluid/solid cell of LBM_STEP.stepCell, wgsl.ts:179-273.  gforce is the
 * invocation's private accumulator (wgsl.ts:178). */
static void step_cell(const og_grid* g, const og_sim* sim, const REAL* fin, REAL* fout,
                      const uint32_t* solid, REAL* fields, const float* emit, uint64_t idx,
                      v3 gforce[3], uint64_t* guard_hits) {
    const uint64_t n = g->n;
    const int NX = (int)g->nx, NY = (int)g->ny, NZ = (int)g->nz;
    if (solid[idx] != 0u) {
        for (uint32_t i = 0; i < 19; i++) fout[i * n + idx] = WT[i];
        for (int k = 0; k < 4; k++) fields[idx * 4 + k] = R(0.0);
        return;
    }
    const int x = (int)(idx % (uint64_t)NX);
    const int y = (int)((idx / (uint64_t)NX) % (uint64_t)NY);
    const int z = (int)(idx / ((uint64_t)NX * (uint64_t)NY));
    const v3 uin = {(REAL)sim->uIn, R(0.0), R(0.0)};
    REAL floorU = (REAL)sim->uIn;
    if (y == 0) {
        const uint64_t row = (uint64_t)z * NX * NY;
        const uint64_t xl = (uint64_t)(x - 2 > 0 ? x - 2 : 0);
        const uint64_t xr = (uint64_t)(x + 2 < NX - 1 ? x + 2 : NX - 1);
        const uint64_t zl = (uint64_t)(z - 1 > 0 ? z - 1 : 0);
        const uint64_t zr = (uint64_t)(z + 1 < NZ - 1 ? z + 1 : NZ - 1);
        if (solid[row + xl] != 0u || solid[row + xr] != 0u || solid[idx + NX] != 0u ||
            solid[zl * NX * NY + (uint64_t)x] != 0u || solid[zr * NX * NY + (uint64_t)x] != 0u)
            floorU = R(0.0);
    }
    const v3 prevU = {fields[idx * 4 + 0], fields[idx * 4 + 1], fields[idx * 4 + 2]};
    const int ductCell = emit[idx] < 0.0f;
    REAL f[19];
    for (uint32_t i = 0; i < 19; i++) {
        const int cx = CX[i], cy = CY[i], cz = CZ[i];
        int sx = x - cx, sy = y - cy, sz = z - cz;
        REAL v;
        if (sx < 0) {
            v = feq(i, R(1.0), uin);
        } else if (sx >= NX) {
            v = feq(i, R(1.0), prevU);
        } else if (sy < 0) {
            v = fin[OPP[i] * n + idx] + R(6.0) * WT[i] * (REAL)cx * floorU;
        } else if (sy >= NY || sz >= NZ) {
            v = feq(i, R(1.0), uin);
        } else {
            uint32_t src = i;
            int mirrored = 0;
            if (sz < 0) { sz = 0; src = ZMIR[i]; mirrored = 1; }
            const uint64_t sidx = (uint64_t)sx + (uint64_t)NX * ((uint64_t)sy + (uint64_t)NY * (uint64_t)sz);
            const uint32_t sg = solid[sidx];
            if (sg != 0u) {
                v = fin[OPP[i] * n + idx];
                if (!mirrored && !ductCell) {
                    /* gforce[sg-1] -= 2 (v - w) c, componentwise */
                    REAL m = R(2.0) * (v - WT[i]);
                    v3* gf = &gforce[sg - 1u];
                    gf->x = gf->x - m * (REAL)cx;
                    gf->y = gf->y - m * (REAL)cy;
                    gf->z = gf->z - m * (REAL)cz;
                }
            } else {
                v = fin[src * n + sidx];
            }
        }
        f[i] = v;
    }

    REAL rho = R(0.0);
    v3 mom = {R(0.0), R(0.0), R(0.0)};
    for (uint32_t i = 0; i < 19; i++) {
        rho += f[i];
        mom.x += f[i] * (REAL)CX[i];
        mom.y += f[i] * (REAL)CY[i];
        mom.z += f[i] * (REAL)CZ[i];
    }
    REAL den = rho > R(1e-4) ? rho : R(1e-4); /* max(rho, 1e-4) */
    v3 u = {mom.x / den, mom.y / den, mom.z / den};
    REAL ulen = sqrt_r(dot3(u, u));
    if (is_bad(rho) || is_bad(u.x) || is_bad(u.y) || is_bad(u.z) || rho < R(0.3) || rho > R(3.0) ||
        ulen > R(0.35)) {
        rho = R(1.0);
        u.x = uin.x * R(0.5); u.y = uin.y * R(0.5); u.z = uin.z * R(0.5);
        for (uint32_t i = 0; i < 19; i++) f[i] = feq(i, rho, u);
        for (int k = 0; k < 3; k++) { gforce[k].x = R(0.0); gforce[k].y = R(0.0); gforce[k].z = R(0.0); }
        (*guard_hits)++;
    }

    if (ductCell) {
        /* mix(u, (0.55 uIn, u.y, u.z), 0.04) = a*(1-t) + b*t per component */
        const REAL t = R(0.04);
        const REAL omt = R(1.0) - t;
        v3 b = {R(0.55) * (REAL)sim->uIn, u.y, u.z};
        u.x = u.x * omt + b.x * t;
        u.y = u.y * omt + b.y * t;
        u.z = u.z * omt + b.z * t;
    }

    REAL pxx = R(0.0), pyy = R(0.0), pzz = R(0.0), pxy = R(0.0), pxz = R(0.0), pyz = R(0.0);
    REAL eq[19];
    for (uint32_t i = 0; i < 19; i++) {
        eq[i] = feq(i, rho, u);
        const REAL ne = f[i] - eq[i];
        const REAL cx = (REAL)CX[i], cy = (REAL)CY[i], cz = (REAL)CZ[i];
        pxx += ne * cx * cx; pyy += ne * cy * cy; pzz += ne * cz * cz;
        pxy += ne * cx * cy; pxz += ne * cx * cz; pyz += ne * cy * cz;
    }
    const REAL q = sqrt_r(pxx * pxx + pyy * pyy + pzz * pzz + R(2.0) * (pxy * pxy + pxz * pxz + pyz * pyz));
    const REAL fx = (REAL)x, fnx = (REAL)NX;
    const REAL tau0 = (REAL)sim->tau0 + R(0.4) * smoothstep_r(fnx - R(18.0), fnx - R(1.0), fx) +
                      R(0.25) * (R(1.0) - smoothstep_r(R(0.0), R(10.0), fx));
    const REAL smag = (REAL)sim->smag;
    const REAL tau = R(0.5) * (tau0 + sqrt_r(tau0 * tau0 + R(18.0 * 1.41421356) * smag * smag * q / rho));
    const REAL omega = R(1.0) / tau;
    const REAL third = R(1.0 / 3.0);
    for (uint32_t i = 0; i < 19; i++) {
        const REAL cx = (REAL)CX[i], cy = (REAL)CY[i], cz = (REAL)CZ[i];
        const REAL qp = (cx * cx - third) * pxx + (cy * cy - third) * pyy + (cz * cz - third) * pzz +
                        R(2.0) * (cx * cy * pxy + cx * cz * pxz + cy * cz * pyz);
        fout[i * n + idx] = eq[i] + (R(1.0) - omega) * R(4.5) * WT[i] * qp;
    }
    fields[idx * 4 + 0] = u.x; fields[idx * 4 + 1] = u.y;
    fields[idx * 4 + 2] = u.z; fields[idx * 4 + 3] = rho;
}

/* LBM_STEP.cs_main, wgsl.ts:275-291: per invocation the private force is rounded to
 * i32 at scale 1e4 (x and y only) and summed; the integer sum is order-free, so the
 * workgroup/global atomics reduce to a plain sum.  forces_d additionally sums the
 * unrounded per-cell force in FP64 (all three components). */
static void step_range(const og_grid* g, const og_sim* s, const REAL* fin, REAL* fout,
                       const uint32_t* solid, REAL* fields, const float* emit, og_stats* st) {
    const uint64_t n = g->n;
#ifdef _OPENMP
#pragma omp parallel
    {
        og_stats loc;
        memset(&loc, 0, sizeof(loc));
#pragma omp for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            uint64_t idx = (uint64_t)ii;
            v3 gf[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            step_cell(g, s, fin, fout, solid, fields, emit, idx, gf, &loc.guard_hits);
            for (int k = 0; k < 3; k++) {
                loc.forces_d[k * 3 + 0] += (double)gf[k].x;
                loc.forces_d[k * 3 + 1] += (double)gf[k].y;
                loc.forces_d[k * 3 + 2] += (double)gf[k].z;
                if (gf[k].x != R(0.0)) loc.forces_q[k * 3 + 0] += (int32_t)rint((float)(gf[k].x * R(10000.0)));
                if (gf[k].y != R(0.0)) loc.forces_q[k * 3 + 1] += (int32_t)rint((float)(gf[k].y * R(10000.0)));
            }
        }
#pragma omp critical
        {
            for (int k = 0; k < 12; k++) st->forces_q[k] += loc.forces_q[k];
            for (int k = 0; k < 9; k++) st->forces_d[k] += loc.forces_d[k];
            st->guard_hits += loc.guard_hits;
        }
    }
#else
    for (uint64_t idx = 0; idx < n; idx++) {
        v3 gf[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        step_cell(g, s, fin, fout, solid, fields, emit, idx, gf, &st->guard_hits);
        for (int k = 0; k < 3; k++) {
            st->forces_d[k * 3 + 0] += (double)gf[k].x;
            st->forces_d[k * 3 + 1] += (double)gf[k].y;
            st->forces_d[k * 3 + 2] += (double)gf[k].z;
            if (gf[k].x != R(0.0)) st->forces_q[k * 3 + 0] += (int32_t)rint((float)(gf[k].x * R(10000.0)));
            if (gf[k].y != R(0.0)) st->forces_q[k * 3 + 1] += (int32_t)rint((float)(gf[k].y * R(10000.0)));
        }
    }
#endif
}

/* Runs nsteps steps with the reference's ping-pong (step s reads buffer s%2, writes
 * the other); after an odd count the newest populations are in fb.  Statistics are
 * accumulated into *st (caller zeroes it). */
EXPORT void orc_run(const og_grid* g, const og_sim* s, REAL* fa, REAL* fb, const uint32_t* solid,
                    REAL* fields, const float* emit, uint32_t nsteps, og_stats* st) {
    for (uint32_t k = 0; k < nsteps; k++) {
        if ((k & 1u) == 0u) step_range(g, s, fa, fb, solid, fields, emit, st);
        else step_range(g, s, fb, fa, solid, fields, emit, st);
    }
}

/* ---- HEAT.cs_main, wgsl.ts:460-511 ---- */
static uint64_t heat_idx(const og_grid* g, int cx, int cy, int cz) {
    /* idxOf: z < 0 mirrored to -z-1, then clamped to the grid (wgsl.ts:469-475) */
    if (cz < 0) cz = -cz - 1;
    const int dx = (int)g->nx, dy = (int)g->ny, dz = (int)g->nz;
    cx = cx < 0 ? 0 : (cx > dx - 1 ? dx - 1 : cx);
    cy = cy < 0 ? 0 : (cy > dy - 1 ? dy - 1 : cy);
    cz = cz < 0 ? 0 : (cz > dz - 1 ? dz - 1 : cz);
    return (uint64_t)cx + (uint64_t)g->nx * ((uint64_t)cy + (uint64_t)g->ny * (uint64_t)cz);
}
static REAL heat_at(const og_grid* g, const REAL* tin, int cx, int cy, int cz) {
    if (cx < 0) return R(0.0);
    return tin[heat_idx(g, cx, cy, cz)];
}

EXPORT void orc_heat(const og_grid* g, const og_heat* hp, const REAL* fields, const float* emit,
                     const REAL* tin, REAL* tout) {
    const uint64_t n = g->n;
    const int NX = (int)g->nx, NY = (int)g->ny, NZ = (int)g->nz;
    static const int OFF[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint64_t idx = (uint64_t)ii;
        const int x = (int)(idx % (uint64_t)NX);
        const int y = (int)((idx / (uint64_t)NX) % (uint64_t)NY);
        const int z = (int)(idx / ((uint64_t)NX * (uint64_t)NY));
        const REAL* fv = fields + idx * 4;
        if (fv[3] <= R(0.0)) { tout[idx] = (REAL)emit[idx]; continue; }
        const REAL dt = (REAL)hp->dt;
        const REAL gx = (REAL)x - fv[0] * dt, gy = (REAL)y - fv[1] * dt, gz = (REAL)z - fv[2] * dt;
        const int ix = (int)floor(gx), iy = (int)floor(gy), iz = (int)floor(gz);
        const REAL ffx = gx - (REAL)ix, ffy = gy - (REAL)iy, ffz = gz - (REAL)iz;
        REAL t = R(0.0);
        for (int k = 0; k < 8; k++) {
            const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
            const REAL w = (ox == 1 ? ffx : R(1.0) - ffx) * (oy == 1 ? ffy : R(1.0) - ffy) *
                           (oz == 1 ? ffz : R(1.0) - ffz);
            t += heat_at(g, tin, ix + ox, iy + oy, iz + oz) * w;
        }
        REAL nb = R(0.0);
        float src = 0.0f;
        for (int k = 0; k < 6; k++) {
            const int ccx = x + OFF[k][0], ccy = y + OFF[k][1], ccz = z + OFF[k][2];
            nb += heat_at(g, tin, ccx, ccy, ccz);
            if (ccx >= 0 && ccy >= 0 && ccx < NX && ccy < NY && ccz < NZ) {
                const float e = emit[heat_idx(g, ccx, ccy, ccz)];
                src = src > e ? src : e; /* max(src, e) */
            }
        }
        const REAL dif = (REAL)hp->diffuse;
        t = t * (R(1.0) - dif) + (nb / R(6.0)) * dif;
        if (src > 0.0f) {
            const REAL a = R(0.35);
            t = t * (R(1.0) - a) + (REAL)src * a;
        }
        const REAL o = t * (R(1.0) - (REAL)hp->cool);
        tout[idx] = o > R(0.0) ? o : R(0.0);
    }
}
