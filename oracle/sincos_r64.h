/*
 * CPU ORACLE — test infrastructure only (tests/ load it; the product never does).
 *
 * sin and cos of an f32 angle in FP64, rounded once to f32: Cody-Waite reduction by pi/2
 * (fdlibm's two-part constant: the first part has 33 significant bits, so k * P1 is exact
 * for |k| < 2^20), Taylor series of sin to r^19 and cos to r^20 on |r| <= pi/4 (truncation
 * below 2^-70), Horner order.  The device evaluates the same steps for the angles that
 * vary per point (the kind-0 wings' spanwise curl, the W17 parts' rotation, the streamline
 * emitter rings), so both give identical bits where the C library's cosf/sinf and the
 * device's may differ in the last one.  Pinned against the C library's double sin/cos
 * (tests/test_sdf_oracle.py).  Included by sdf_oracle.c and flow_oracle.c.
 */
#ifndef ORACLE_SINCOS_R64_H
#define ORACLE_SINCOS_R64_H
#include <math.h>

static void sincos_r64(float a, float* sn, float* cs) {
    const double x = (double)a;
    if (!(fabs(x) <= 1e9)) {
        *sn = *cs = NAN;
        return;
    }
    const double kd = rint(x * 0x1.45f306dc9c883p-1);
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
    *sn = (float)(q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c)));
    *cs = (float)(q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s)));
}

#endif
