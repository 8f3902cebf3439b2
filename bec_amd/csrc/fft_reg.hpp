// In-register DFTs of power-of-two length, fully unrolled at compile time.
// Input and output are in natural order; SIGN = -1 computes
// X[k] = sum_j x[j] exp(-2 pi i jk/N), SIGN = +1 the conjugate kernel.
#pragma once

#include "common.hpp"

#include <type_traits>

namespace bec {

template <int B, int E, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

constexpr int bitrev(int i, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

// a * exp(SIGN * 2 pi i K / L), with the trivial angles special-cased
template <typename T, long K, long L, int SIGN> __device__ __forceinline__ cpx<T> twmul(cpx<T> a)
{
    constexpr long k = ((K % L) + L) % L;
    if constexpr (k == 0) {
        return a;
    } else if constexpr (2 * k == L) {
        return mk<T>(-a.x, -a.y);
    } else if constexpr (4 * k == L) {          // exp(SIGN i pi/2) = SIGN i
        return SIGN > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
    } else if constexpr (4 * k == 3 * L) {      // exp(SIGN 3 i pi/2) = -SIGN i
        return SIGN > 0 ? mk<T>(a.y, -a.x) : mk<T>(-a.y, a.x);
    } else if constexpr (8 * k == L || 8 * k == 3 * L || 8 * k == 5 * L || 8 * k == 7 * L) {
        constexpr T h = (T)0.70710678118654752440084436210484903928;
        constexpr double c = ccos2pi(k, L), s = SIGN * csin2pi(k, L);
        constexpr T cs = c > 0 ? h : -h, ss = s > 0 ? h : -h;
        if constexpr (cs == ss) {
            return mk<T>((a.x - a.y) * cs, (a.x + a.y) * cs);
        } else {
            return mk<T>((a.x + a.y) * cs, (a.y - a.x) * cs);
        }
    } else {
        constexpr T c = (T)ccos2pi(k, L);
        constexpr T s = (T)(SIGN * csin2pi(k, L));
        return mk<T>(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// FP32 twiddle constants with a balanced modulus.  Rounding cos and sin to
// the nearest floats separately leaves every non-trivial twiddle of the
// radix-2 transforms with |w|^2 - 1 < 0 (-5.7e-8 at pi/8, -3.4e-8 at pi/4), so
// each forward/inverse pair of 16-point transforms loses 5.3e-8 of the energy
// (9e-8 at 32 points) whatever the data: a systematic norm drift of the FP32
// state.  A butterfly whose twiddle has modulus error e changes the energy of
// an N-point transform of white data by about e/N, so the errors of all the
// twiddles a transform applies should sum to zero.  For each angle, the float
// pairs within two ulps of (cos, sin) offer a best pair below modulus one and
// a best pair above it (at an odd multiple of pi/4, -3.4e-8 and +5.0e-8: one
// ulp in either coordinate moves c^2 + s^2 by 8.4e-8, so no pair is closer).
// tw_side() picks one of the two for every butterfly of the transform: the
// butterflies with the largest errors first, each taking the side that keeps
// the running sum nearest zero.  In an FP32 emulation of the transforms
// (tests/test_twiddle_modulus.py) the round-trip energy bias is +2e-10 at
// 16 points and -1e-11 at 32 points (it was +9.5e-9 and -3e-9 with the best
// pair of each angle, alternating at the odd multiples of pi/4); an 8-point
// transform has only two such butterflies and keeps +4e-9.  The transforms
// stay accurate to 2e-7 relative.
struct TwPair {
    double c, s;
};

constexpr double f32_ulp(double v)   // spacing of floats around v (normal range)
{
    double a = v < 0 ? -v : v, p = 1.0;
    if (a < 1.1754943508222875e-38) return 1.401298464324817e-45;   // zero and subnormals
    while (p > a) p *= 0.5;
    while (p * 2.0 <= a) p *= 2.0;
    return p * 1.1920928955078125e-7;
}

// The float pair within two ulps of (cos, sin) of angle sign 2 pi k / L whose
// modulus error c^2 + s^2 - 1 is smallest, on the side of zero that `want`
// asks for (-1 below, +1 above, 0 either); ties go to the smaller angle error.
constexpr TwPair tw_f32(long k, long L, int sign, int want)
{
    const double c0 = ccos2pi(k, L), s0 = sign * csin2pi(k, L);
    const double cf = (double)(float)c0, sf = (double)(float)s0;
    const double uc = f32_ulp(cf), us = f32_ulp(sf);
    TwPair best{cf, sf};
    double be = 1e300, bp = 1e300;
    for (int i = -2; i <= 2; ++i)
        for (int j = -2; j <= 2; ++j) {
            const double c = cf + i * uc, s = sf + j * us;
            if ((double)(float)c != c || (double)(float)s != s) continue;   // crossed a binade
            const double e = c * c + s * s - 1.0;
            if ((want < 0 && !(e < 0)) || (want > 0 && !(e > 0))) continue;
            const double ae = e < 0 ? -e : e;
            const double pe = c0 * s - s0 * c < 0 ? s0 * c - c0 * s : c0 * s - s0 * c;
            if (ae < be || (ae == be && pe < bp)) {
                best = TwPair{c, s};
                be = ae;
                bp = pe;
            }
        }
    return best;
}

// exp(2 pi i k / len) is 1, -1 or +-i: the butterfly is additions only
constexpr bool trivial_twiddle(long k, long len)
{
    const long r = ((k % len) + len) % len;
    return r == 0 || 2 * r == len || 4 * r == len || 4 * r == 3 * len;
}

constexpr double tw_err(TwPair p) { return p.c * p.c + p.s * p.s - 1.0; }

constexpr int TW_MAX_USES = 64;   // non-trivial butterflies of one transform, N <= 32

// The largest-first balancing of an N-point transform of kernel sign `sign`
// towards a sum `target` of its butterflies' modulus errors, over its
// non-trivial butterflies in the loop order of dft() (stage, then k, then
// block).  Returns the side (-1 below modulus one, +1 above) of butterfly
// (L0, k0, B0), and the sum reached in *sum_out.
constexpr int tw_balance(int N, int sign, double target, long L0, long k0, int B0, double *sum_out)
{
    double en[TW_MAX_USES] = {}, ep[TW_MAX_USES] = {};
    int side[TW_MAX_USES] = {}, done[TW_MAX_USES] = {};
    int n = 0, me = -1;
    for (long L = 2; L <= N; L *= 2)
        for (long k = 0; k < L / 2; ++k)
            for (int B = 0; B < N / L; ++B) {
                if (trivial_twiddle(k, L)) continue;
                en[n] = tw_err(tw_f32(k, L, sign, -1));
                ep[n] = tw_err(tw_f32(k, L, sign, +1));
                if (L == L0 && k == k0 && B == B0) me = n;
                ++n;
            }
    double run = 0.0;
    for (int step = 0; step < n; ++step) {
        int bi = -1;
        double bm = -1.0;
        for (int i = 0; i < n; ++i) {
            const double m = -en[i] > ep[i] ? -en[i] : ep[i];
            if (!done[i] && m > bm) {
                bm = m;
                bi = i;
            }
        }
        done[bi] = 1;
        const double a = run + en[bi] - target, b = run + ep[bi] - target;
        const double aa = a < 0 ? -a : a, ab = b < 0 ? -b : b;
        side[bi] = (aa < ab || (aa == ab && -en[bi] <= ep[bi])) ? -1 : 1;
        run += side[bi] < 0 ? en[bi] : ep[bi];
    }
    *sum_out = run;
    return me >= 0 ? side[me] : 0;
}

// Side of the pair that butterfly (L, k, B) of an N-point transform uses.
// PARTNER = 0: the transform balances its own errors to zero.  PARTNER = P:
// the transform is the first stage of a two-stage transform whose second
// stage is a P-point transform balanced to zero; a stage of length n changes
// the energy by its error sum / n, so this stage aims at -(N/P) times the sum
// the P-point stage is left with.  That matters where P cannot balance itself:
// an 8-point transform has only its two pi/4 butterflies (sum +1.6e-8).
constexpr int tw_side(int N, int sign, long L, long k, int B, int PARTNER)
{
    double target = 0.0, s = 0.0;
    if (PARTNER > 0) {
        tw_balance(PARTNER, sign, 0.0, -1, -1, -1, &s);
        target = -(double)N / PARTNER * s;
    }
    return tw_balance(N, sign, target, L, k, B, &s);
}

// u + b exp(SIGN 2 pi i K / L) for a non-trivial angle, accumulated in FMAs;
// in FP32 with the pair that butterfly block B of an N-point transform takes
template <typename T, int N, int PARTNER, long K, long L, int B, int SIGN>
__device__ __forceinline__ cpx<T> twiddle_add(cpx<T> u, cpx<T> b)
{
    constexpr long k = ((K % L) + L) % L;
    if constexpr (sizeof(T) == 4) {
        constexpr TwPair P = tw_f32(k, L, SIGN, tw_side(N, SIGN, L, k, B, PARTNER));
        constexpr T c = (T)P.c, s = (T)P.s;
        return mk<T>(fma(b.x, c, fma(-b.y, s, u.x)), fma(b.x, s, fma(b.y, c, u.y)));
    } else if constexpr (8 * k == L || 8 * k == 3 * L || 8 * k == 5 * L || 8 * k == 7 * L) {
        constexpr T h = (T)0.70710678118654752440084436210484903928;
        constexpr double c = ccos2pi(k, L), s = SIGN * csin2pi(k, L);
        constexpr T cs = c > 0 ? h : -h, ss = s > 0 ? h : -h;
        if constexpr (cs == ss) return mk<T>(fma(b.x - b.y, cs, u.x), fma(b.x + b.y, cs, u.y));
        else return mk<T>(fma(b.x + b.y, cs, u.x), fma(b.y - b.x, cs, u.y));
    } else {
        constexpr T c = (T)ccos2pi(k, L);
        constexpr T s = (T)(SIGN * csin2pi(k, L));
        return mk<T>(fma(b.x, c, fma(-b.y, s, u.x)), fma(b.x, s, fma(b.y, c, u.y)));
    }
}

// Radix-2 decimation-in-time, in place on v[0..N).  PARTNER: see tw_side (the
// length of the second stage when this is the first of two; 0 otherwise).
template <typename T, int N, int SIGN, int PARTNER = 0> __device__ __forceinline__ void dft(cpx<T> (&v)[N])
{
    constexpr int LOG = ilog2(N);
    static_assert((1 << LOG) == N, "power-of-two length");
    if constexpr (N == 1) {
        return;
    } else {
        cpx<T> t[N];
        static_for<0, N>([&](auto I) { t[bitrev(decltype(I)::value, LOG)] = v[decltype(I)::value]; });
        static_for<1, LOG + 1>([&](auto S) {
            constexpr int len = 1 << decltype(S)::value;
            constexpr int half = len / 2;
            static_for<0, half>([&](auto K) {
                constexpr int k = decltype(K)::value;
                static_for<0, N / len>([&](auto B_) {
                    constexpr int blk = decltype(B_)::value;
                    constexpr int i = blk * len + k;
                    cpx<T> u = t[i];
                    if constexpr (trivial_twiddle(k, len)) {
                        cpx<T> w = twmul<T, k, len, SIGN>(t[i + half]);
                        t[i] = cadd<T>(u, w);
                        t[i + half] = csub<T>(u, w);
                    } else {
                        // u + w b as nested FMAs, and the second output as 2u - (u + w b):
                        // 6 instructions instead of 8
                        const cpx<T> s1 = twiddle_add<T, N, PARTNER, k, len, blk, SIGN>(u, t[i + half]);
                        t[i] = s1;
                        t[i + half] = mk<T>(fma((T)2, u.x, -s1.x), fma((T)2, u.y, -s1.y));
                    }
                });
            });
        });
        static_for<0, N>([&](auto I) { v[decltype(I)::value] = t[decltype(I)::value]; });
    }
}

} // namespace bec
