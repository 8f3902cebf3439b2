// Compact single-precision sin/cos for the nonlinear phase factors.
//
// Method: Cody-Waite argument reduction by pi/2 with a three-term split of
// pi/2 applied through fused multiply-adds, then Taylor polynomials of sin
// (to r^9) and cos (to r^10) on |r| <= pi/4, whose truncation errors
// (< 3e-9 and < 2e-10) are far below one single-precision ulp.  Every
// operation is an IEEE single-precision fma, multiply, add or round-to-nearest,
// so the host build of this header computes bit-identical results to the
// device build; the unit test checks the host build against double precision.
// Valid for |x| < 2^15 (the reduction constant products stay exact enough).
#pragma once

#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define BEC_HD __host__ __device__
#else
#define BEC_HD
#endif

namespace bec {

// pi/2 = PIO2_A + PIO2_B + PIO2_C, each the float nearest to the remainder
constexpr float PIO2_A = 1.57079637050628662109375f;
constexpr float PIO2_B = -4.371138828673793e-8f;
constexpr float PIO2_C = -1.7151245100058819e-15f;
constexpr float TWO_OVER_PI = 0.636619746685028076171875f;

BEC_HD inline void fast_sincosf(float x, float *s, float *c)
{
    const float q = rintf(x * TWO_OVER_PI);
    float r = fmaf(-q, PIO2_A, x);
    r = fmaf(-q, PIO2_B, r);
    r = fmaf(-q, PIO2_C, r);
    const float r2 = r * r;
    // sin r = r + r^3 (-1/3! + r^2 (1/5! + r^2 (-1/7! + r^2 / 9!)))
    float ps = fmaf(r2, 2.75573192239858906526e-6f, -1.98412698412698412698e-4f);
    ps = fmaf(r2, ps, 8.33333333333333333333e-3f);
    ps = fmaf(r2, ps, -1.66666666666666666667e-1f);
    const float sn = fmaf(r * r2, ps, r);
    // cos r = 1 + r^2 (-1/2 + r^2 (1/4! + r^2 (-1/6! + r^2 (1/8! - r^2 / 10!))))
    float pc = fmaf(r2, -2.75573192239858906526e-7f, 2.48015873015873015873e-5f);
    pc = fmaf(r2, pc, -1.38888888888888888889e-3f);
    pc = fmaf(r2, pc, 4.16666666666666666667e-2f);
    pc = fmaf(r2, pc, -0.5f);
    const float cs = fmaf(r2, pc, 1.0f);
    const int k = (int)q;
    const bool swap = (k & 1) != 0;
    float ss = swap ? cs : sn;
    float cc = swap ? sn : cs;
    if (k & 2) ss = -ss;
    if ((k + 1) & 2) cc = -cc;
    *s = ss;
    *c = cc;
}

} // namespace bec
