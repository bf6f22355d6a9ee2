/*
 * CPU ORACLE — test infrastructure only (tests/ load it; the product never does).
 *
 * FP64 restatement of the reference's vortex-lattice wing loads, src/engine/vlm.ts:
 *   segment (:42-51)            regularised Biot-Savart velocity of a straight vortex
 *                               segment (core 0.01 m);
 *   horseshoeVelocity (:54-66)  a horseshoe whose trailing legs follow the surface to
 *                               the trailing edge and then run 60 m downstream, with its
 *                               ground image (y -> -y, circulation negated);
 *   solve (:115-138)            Gaussian elimination with partial pivoting, rows swapped,
 *                               zero multipliers skipped, zero or NaN pivots replaced by
 *                               1e-12 (the source's `d || 1e-12`);
 *   wingLoads (:140-182)        influence matrix of normal velocities at the collocation
 *                               points, circulations, Kutta-Joukowski force at each bound
 *                               vortex midpoint, drag + CD0 q area, centres of pressure.
 * A horseshoe is 20 doubles: a, b (bound vortex), ta, tb (trailing-edge points), cp
 * (collocation point), n (unit normal), area, pad.  Math.hypot is V8's (the engine of
 * Node and Chrome): NaN / infinity first, then the magnitudes scaled by the largest and
 * their squares Kahan-summed (vo_hypot3).  Built with -ffp-contract=off.
 *
 * What pins it (tests/test_vlm_oracle.py): vo_hypot3 against Math.hypot outputs that
 * Node wrote (tests/golden/js_hypot.npz); the segment far from its core against the
 * classical Biot-Savart formula and a long segment against the infinite line 1/(2 pi d);
 * the ground image cancelling the normal velocity on the ground plane; the solver
 * against numpy's LU and on matrices that need row swaps; a flat rectangular wing high
 * above the ground against Helmbold's lift slope, its centre of pressure at the quarter
 * chord and its induced drag against CL^2 / (pi AR); ground effect raising the lift.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define VO_RHO 1.225
#define VO_CD0 0.018
#define VO_CORE 0.01
#define VO_FAR 60.0
#define VO_REC 20
#define VO_PI 3.141592653589793 /* Math.PI */

/* Math.hypot(x, y, z) as V8 evaluates it */
double vo_hypot3(double x, double y, double z) {
    const double v[3] = {x, y, z};
    double a[3] = {0, 0, 0}, mx = 0;
    int nan = 0;
    for (int i = 0; i < 3; i++) {
        if (v[i] != v[i]) {
            nan = 1;
        } else {
            a[i] = fabs(v[i]);
            if (a[i] > mx) mx = a[i];
        }
    }
    if (mx == INFINITY) return INFINITY;
    if (nan) return NAN;
    if (mx == 0) return 0;
    double sum = 0, comp = 0;
    for (int i = 0; i < 3; i++) {
        const double n = a[i] / mx;
        const double s = n * n - comp;
        const double p = sum + s;
        comp = (p - sum) - s;
        sum = p;
    }
    return sqrt(sum) * mx;
}

static double vo_len(const double* v) { return vo_hypot3(v[0], v[1], v[2]); }
static double vo_dot(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* segment (vlm.ts:42-51): out += velocity at p of the segment a -> b of circulation gamma */
void vo_segment(const double* p, const double* a, const double* b, double* out, double gamma) {
    const double r1[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    const double r2[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    const double c[3] = {r1[1] * r2[2] - r1[2] * r2[1], r1[2] * r2[0] - r1[0] * r2[2], r1[0] * r2[1] - r1[1] * r2[0]};
    const double r0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double l1 = vo_len(r1), l2 = vo_len(r2);
    const double c2 = vo_dot(c, c) + (VO_CORE * VO_CORE) * vo_dot(r0, r0);
    if (l1 < 1e-6 || l2 < 1e-6) return;
    const double k = (gamma / ((4 * VO_PI) * c2)) * (vo_dot(r0, r1) / l1 - vo_dot(r0, r2) / l2);
    out[0] += c[0] * k;
    out[1] += c[1] * k;
    out[2] += c[2] * k;
}

/* horseshoeVelocity (vlm.ts:54-66) */
void vo_horseshoe(const double* p, const double* h, double gamma, double* out) {
    const double* a = h;
    const double* b = h + 3;
    const double* ta = h + 6;
    const double* tb = h + 9;
    const double afar[3] = {ta[0] + VO_FAR, ta[1], ta[2]};
    const double bfar[3] = {tb[0] + VO_FAR, tb[1], tb[2]};
    for (int s = 0; s < 2; s++) {
        const double sign = s == 0 ? 1.0 : -1.0;
#define M(dst, v) const double dst[3] = {(v)[0], (v)[1] * sign, (v)[2]}
        M(maf, afar); M(mta, ta); M(ma, a); M(mb, b); M(mtb, tb); M(mbf, bfar);
#undef M
        const double g = gamma * sign;
        vo_segment(p, maf, mta, out, g);
        vo_segment(p, mta, ma, out, g);
        vo_segment(p, ma, mb, out, g);
        vo_segment(p, mb, mtb, out, g);
        vo_segment(p, mtb, mbf, out, g);
    }
}

/* solve (vlm.ts:115-138), in place; x receives the solution */
void vo_solve(double* A, double* rhs, int n, double* x) {
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(A[(size_t)i * n + k]) > fabs(A[(size_t)piv * n + k])) piv = i;
        if (piv != k) {
            for (int j = 0; j < n; j++) {
                const double t = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)piv * n + j];
                A[(size_t)piv * n + j] = t;
            }
            const double t = rhs[k];
            rhs[k] = rhs[piv];
            rhs[piv] = t;
        }
        double d = A[(size_t)k * n + k];
        if (d == 0 || d != d) d = 1e-12;
        for (int i = k + 1; i < n; i++) {
            const double f = A[(size_t)i * n + k] / d;
            if (f == 0) continue;
            for (int j = k; j < n; j++) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
            rhs[i] -= f * rhs[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = rhs[i];
        for (int j = i + 1; j < n; j++) s -= A[(size_t)i * n + j] * x[j];
        double d = A[(size_t)i * n + i];
        if (d == 0 || d != d) d = 1e-12;
        x[i] = s / d;
    }
}

/* wingLoads (vlm.ts:140-182) after buildLattice: out = downforce, drag, copX, dragY */
int vo_loads(const double* hs, int n, double speed, double* out, double* gamma_out) {
    const size_t m = n > 0 ? (size_t)n : 1;
    double* A = (double*)malloc(m * m * sizeof(double));
    double* rhs = (double*)malloc(m * sizeof(double));
    double* gamma = (double*)malloc(m * sizeof(double));
    if (!A || !rhs || !gamma) { free(A); free(rhs); free(gamma); return -1; }
    const double vinf[3] = {speed, 0, 0};
    for (int i = 0; i < n; i++) {
        const double* hi = hs + (size_t)VO_REC * i;
        for (int j = 0; j < n; j++) {
            double v[3] = {0, 0, 0};
            vo_horseshoe(hi + 12, hs + (size_t)VO_REC * j, 1, v);
            A[(size_t)i * n + j] = vo_dot(v, hi + 15);
        }
        rhs[i] = -vo_dot(vinf, hi + 15);
    }
    vo_solve(A, rhs, n, gamma);
    double force[3] = {0, 0, 0}, area = 0, mx = 0, my = 0;
    for (int i = 0; i < n; i++) {
        const double* h = hs + (size_t)VO_REC * i;
        const double m[3] = {(h[0] + h[3]) / 2, (h[1] + h[4]) / 2, (h[2] + h[5]) / 2};
        double vel[3] = {vinf[0], vinf[1], vinf[2]};
        for (int j = 0; j < n; j++) vo_horseshoe(m, hs + (size_t)VO_REC * j, gamma[j], vel);
        const double ab[3] = {h[3] - h[0], h[4] - h[1], h[5] - h[2]};
        const double f[3] = {vel[1] * ab[2] - vel[2] * ab[1], vel[2] * ab[0] - vel[0] * ab[2],
                             vel[0] * ab[1] - vel[1] * ab[0]};
        force[0] += (VO_RHO * gamma[i]) * f[0];
        force[1] += (VO_RHO * gamma[i]) * f[1];
        force[2] += (VO_RHO * gamma[i]) * f[2];
        mx += (((-VO_RHO) * gamma[i]) * f[1]) * m[0];
        my += ((VO_RHO * gamma[i]) * f[0]) * m[1];
        area += h[18];
    }
    const double q = (0.5 * VO_RHO) * speed * speed;
    const double down = -force[1];
    double cx = 0, cy = 0;
    for (int i = 0; i < n; i++) cx = cx + hs[(size_t)VO_REC * i + 12];
    for (int i = 0; i < n; i++) cy = cy + hs[(size_t)VO_REC * i + 13];
    const double nn = n > 1 ? (double)n : 1.0;
    out[0] = down;
    out[1] = force[0] + ((VO_CD0 * q) * area);
    out[2] = fabs(down) > 1 ? mx / down : cx / nn;
    out[3] = fabs(force[0]) > 1 ? my / force[0] : cy / nn;
    if (gamma_out) memcpy(gamma_out, gamma, (size_t)n * sizeof(double));
    free(A); free(rhs); free(gamma);
    return 0;
}
