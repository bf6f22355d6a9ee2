/*
 * CPU ORACLE — test infrastructure only (tests/ load it; the product never does).
 *
 * Restatement of the reference's distance-grid baker, bakeGrid of
 * scripts/bake-car.ts:237-316 (closestOnTriangle :201-224, toHalf :226-235):
 *   1. every selected triangle (FP64 corners) lowers the unsigned distance of the grid
 *      points in its bounding box grown by `band` (points at origin + i*h); distances
 *      are kept in an f32 array initialised to band, and a triangle whose corners all
 *      lie below z = -band is skipped;
 *   2. morphological closing: a 6-connected flood from every face but z = 0 through
 *      cells whose distance is >= closing, then a 26-connected regrowth of the flooded
 *      set by ceil(closing / h) cells;
 *   3. output f16 (truncating toHalf of the f32 value) of d - shell outside (regrown
 *      cells) and -max(d, 0.001) - shell inside.
 * FP64 arithmetic in the source's order; built with -ffp-contract=off.
 *
 * What pins it (tests/test_bake_oracle.py): the shipped w17.sdf.bin and every grid of
 * w17.parts.bin, which the reference's own baker wrote, are reproduced bit for bit from
 * public/models/w17.glb and w17.rig.bin (where the reference is present); the
 * point-triangle distance against brute force over the closed triangle; closed forms
 * for a sphere mesh.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* closestOnTriangle (bake-car.ts:201-224): distance from p to triangle t[9] */
double bo_closest(double px, double py, double pz, const double* t) {
    const double ax = t[0], ay = t[1], az = t[2], bx = t[3], by = t[4], bz = t[5], cx = t[6], cy = t[7], cz = t[8];
    const double abx = bx - ax, aby = by - ay, abz = bz - az, acx = cx - ax, acy = cy - ay, acz = cz - az;
    const double apx = px - ax, apy = py - ay, apz = pz - az;
    const double d1 = (abx * apx + aby * apy) + abz * apz, d2 = (acx * apx + acy * apy) + acz * apz;
#define SQ(x, y, z) sqrt(((x) * (x) + (y) * (y)) + (z) * (z))
    if (d1 <= 0 && d2 <= 0) return SQ(apx, apy, apz);
    const double bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const double d3 = (abx * bpx + aby * bpy) + abz * bpz, d4 = (acx * bpx + acy * bpy) + acz * bpz;
    if (d3 >= 0 && d4 <= d3) return SQ(bpx, bpy, bpz);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        return SQ(apx - v * abx, apy - v * aby, apz - v * abz);
    }
    const double cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    const double d5 = (abx * cpx + aby * cpy) + abz * cpz, d6 = (acx * cpx + acy * cpy) + acz * cpz;
    if (d6 >= 0 && d5 <= d6) return SQ(cpx, cpy, cpz);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const double w = d2 / (d2 - d6);
        return SQ(apx - w * acx, apy - w * acy, apz - w * acz);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return SQ(px - (bx + w * (cx - bx)), py - (by + w * (cy - by)), pz - (bz + w * (cz - bz)));
    }
    const double denom = 1 / ((va + vb) + vc);
    const double v = vb * denom, w = vc * denom;
    return SQ((apx - abx * v) - acx * w, (apy - aby * v) - acy * w, (apz - abz * v) - acz * w);
#undef SQ
}

/* toHalf (bake-car.ts:226-235): f32 value -> f16 bits, mantissa truncated, exponent
 * underflow to signed zero, overflow to infinity */
uint16_t bo_to_half(double v) {
    const float f = (float)v;
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int exp = (int)((x >> 23) & 0xffu) - 127 + 15;
    const uint32_t mant = x & 0x7fffffu;
    if (exp <= 0) return (uint16_t)sign;
    if (exp >= 31) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | ((uint32_t)exp << 10) | (mant >> 13));
}

static double maxd(double a, double b) { return a > b ? a : b; }
static double mind(double a, double b) { return a < b ? a : b; }

/* Step 1 only: the f32 distance array (for tests of the sweep alone). */
void bo_distance(const double* tris, uint32_t ntri, const double* origin, double h, const uint32_t* dims, double band,
                 float* dist) {
    const uint32_t NX = dims[0], NY = dims[1], NZ = dims[2];
    const size_t n = (size_t)NX * NY * NZ;
    for (size_t i = 0; i < n; i++) dist[i] = (float)band;
    for (uint32_t f = 0; f < ntri; f++) {
        const double* t = tris + 9 * (size_t)f;
        if (maxd(maxd(t[2], t[5]), t[8]) < -band) continue;
        long lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            const double mn = mind(mind(t[a], t[a + 3]), t[a + 6]);
            const double mx = maxd(maxd(t[a], t[a + 3]), t[a + 6]);
            const double l = floor((mn - band - origin[a]) / h), u = ceil((mx + band - origin[a]) / h);
            lo[a] = l > 0 ? (long)l : 0;
            hi[a] = u < (double)dims[a] - 1 ? (long)u : (long)dims[a] - 1;
        }
        for (long k = lo[2]; k <= hi[2]; k++)
            for (long j = lo[1]; j <= hi[1]; j++)
                for (long i = lo[0]; i <= hi[0]; i++) {
                    const double d = bo_closest(origin[0] + (double)i * h, origin[1] + (double)j * h,
                                                origin[2] + (double)k * h, t);
                    const size_t idx = ((size_t)k * NY + (size_t)j) * NX + (size_t)i;
                    if (d < (double)dist[idx]) dist[idx] = (float)d;
                }
    }
}

/* bakeGrid (bake-car.ts:237-316) */
int bo_bake_grid(const double* tris, uint32_t ntri, const double* origin, double h, const uint32_t* dims, double band,
                 double closing, double shell, uint16_t* out) {
    const long NX = dims[0], NY = dims[1], NZ = dims[2];
    const size_t n = (size_t)NX * NY * NZ;
    float* dist = (float*)malloc(n * sizeof(float));
    uint8_t* reached = (uint8_t*)calloc(n, 1);
    int32_t* queue = (int32_t*)malloc(n * sizeof(int32_t));
    int16_t* grow = (int16_t*)malloc(n * sizeof(int16_t));
    if (!dist || !reached || !queue || !grow) {
        free(dist); free(reached); free(queue); free(grow);
        return -1;
    }
    bo_distance(tris, ntri, origin, h, dims, band, dist);
#define AT(i, j, k) (((size_t)(k) * NY + (size_t)(j)) * NX + (size_t)(i))
    size_t qh = 0, qt = 0;
#define PUSH(i, j, k)                                                               \
    do {                                                                            \
        const size_t id_ = AT(i, j, k);                                             \
        if (!reached[id_] && !((double)dist[id_] < closing)) {                      \
            reached[id_] = 1;                                                       \
            queue[qt++] = (int32_t)id_;                                             \
        }                                                                           \
    } while (0)
    for (long k = 0; k < NZ; k++) for (long j = 0; j < NY; j++) { PUSH(0, j, k); PUSH(NX - 1, j, k); }
    for (long k = 0; k < NZ; k++) for (long i = 0; i < NX; i++) { PUSH(i, 0, k); PUSH(i, NY - 1, k); }
    for (long j = 0; j < NY; j++) for (long i = 0; i < NX; i++) PUSH(i, j, NZ - 1);
    while (qh < qt) {
        const size_t idx = (size_t)queue[qh++];
        const long i = (long)(idx % (size_t)NX), j = (long)((idx / (size_t)NX) % (size_t)NY),
                   k = (long)(idx / ((size_t)NX * NY));
        if (i > 0) PUSH(i - 1, j, k);
        if (i < NX - 1) PUSH(i + 1, j, k);
        if (j > 0) PUSH(i, j - 1, k);
        if (j < NY - 1) PUSH(i, j + 1, k);
        if (k > 0) PUSH(i, j, k - 1);
        if (k < NZ - 1) PUSH(i, j, k + 1);
    }
    for (size_t i = 0; i < n; i++) grow[i] = -1;
    qh = 0;
    qt = 0;
    for (size_t idx = 0; idx < n; idx++)
        if (reached[idx]) { grow[idx] = 0; queue[qt++] = (int32_t)idx; }
    const long max_steps = (long)ceil(closing / h);
    while (qh < qt) {
        const size_t idx = (size_t)queue[qh++];
        const int16_t g = grow[idx];
        if (g >= max_steps) continue;
        const long i = (long)(idx % (size_t)NX), j = (long)((idx / (size_t)NX) % (size_t)NY),
                   k = (long)(idx / ((size_t)NX * NY));
        for (long dk = -1; dk <= 1; dk++) for (long dj = -1; dj <= 1; dj++) for (long di = -1; di <= 1; di++) {
            const long a = i + di, b = j + dj, c = k + dk;
            if (a < 0 || b < 0 || c < 0 || a >= NX || b >= NY || c >= NZ) continue;
            const size_t m = AT(a, b, c);
            if (grow[m] >= 0) continue;
            grow[m] = (int16_t)(g + 1);
            queue[qt++] = (int32_t)m;
        }
    }
    for (size_t idx = 0; idx < n; idx++) {
        const double d = (double)dist[idx];
        out[idx] = bo_to_half(grow[idx] >= 0 ? d - shell : -maxd(d, 0.001) - shell);
    }
#undef PUSH
#undef AT
    free(dist); free(reached); free(queue); free(grow);
    return 0;
}
