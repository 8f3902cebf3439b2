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
#include <string.h>

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

// pi = PI_A + PI_B + PI_C, each the float nearest to the remainder
constexpr float PI_A = 3.1415927410125732421875f;
constexpr float PI_B = -8.742277657347586e-8f;
constexpr float PI_C = -3.4302490200117637e-15f;
constexpr float INV_PI = 0.3183098733425140380859375f;

BEC_HD inline float xor_bits(float v, unsigned m)
{
    unsigned u;
    memcpy(&u, &v, sizeof(u));
    u ^= m;
    memcpy(&v, &u, sizeof(v));
    return v;
}

// Adding 1.5 * 2^23 to |v| < 2^22 rounds v to an integer held in the sum's low
// mantissa bits (the sum lies in [2^23, 2^24), where one ulp is 1).
constexpr float ROUND_MAGIC = 12582912.0f;

// The same accuracy for |x| < 1e5 with a reduction by pi instead of pi/2: x = q pi + r,
// |r| <= pi/2, sin x = (-1)^q sin r and cos x = (-1)^q cos r, so the quadrant
// fix-up is one sign-bit xor per result instead of a swap and two selects.
// Taylor polynomials of sin to r^13 and cos to r^14 (truncation errors
// < 7e-10 and < 7e-11 at pi/2).
//
// Both results come out multiplied by S, which must be a power of two (1 for
// the plain functions): every coefficient and the reduced argument are scaled
// by S before they enter the fused multiply-adds, and scaling by a power of two
// commutes with rounding (no overflow or underflow occurs here), so the results
// are exactly S sin x and S cos x as the unscaled evaluation rounds them.  A
// caller that multiplies sin and cos by a constant power of two anyway (the
// 1/N^2 of an unnormalised transform pair) gets that product for one multiply
// instead of two.  The quadrant q comes from one fused multiply-add with
// ROUND_MAGIC; its parity is the low bit of that sum.
BEC_HD inline void fast_sincos_pi_scaled(float x, float S, float *s, float *c)
{
    const float t = fmaf(x, INV_PI, ROUND_MAGIC);
    const float q = t - ROUND_MAGIC;
    float r = fmaf(-q, PI_A, x);
    r = fmaf(-q, PI_B, r);      // the PI_C term, |q| 3.4e-15, is below 1e-10 for |x| < 1e5
    const float r2 = r * r;
    float ps = fmaf(r2, S * 1.6059043836821614599e-10f, S * -2.5052108385441718775e-8f);
    ps = fmaf(r2, ps, S * 2.7557319223985890653e-6f);
    ps = fmaf(r2, ps, S * -1.9841269841269841270e-4f);
    ps = fmaf(r2, ps, S * 8.3333333333333333333e-3f);
    ps = fmaf(r2, ps, S * -1.6666666666666666667e-1f);
    const float sn = fmaf(r * r2, ps, S * r);
    float pc = fmaf(r2, S * -1.1470745597729724714e-11f, S * 2.0876756987868098979e-9f);
    pc = fmaf(r2, pc, S * -2.7557319223985890653e-7f);
    pc = fmaf(r2, pc, S * 2.4801587301587301587e-5f);
    pc = fmaf(r2, pc, S * -1.3888888888888888889e-3f);
    pc = fmaf(r2, pc, S * 4.1666666666666666667e-2f);
    pc = fmaf(r2, pc, S * -0.5f);
    const float cs = fmaf(r2, pc, S);
    unsigned tb;
    memcpy(&tb, &t, sizeof(tb));
    const unsigned sign = tb << 31;
    *s = xor_bits(sn, sign);
    *c = xor_bits(cs, sign);
}

BEC_HD inline void fast_sincos_pi(float x, float *s, float *c)
{
    fast_sincos_pi_scaled(x, 1.0f, s, c);
}

} // namespace bec
