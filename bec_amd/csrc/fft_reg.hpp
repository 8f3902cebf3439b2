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

// Radix-2 decimation-in-time, in place on v[0..N).
template <typename T, int N, int SIGN> __device__ __forceinline__ void dft(cpx<T> (&v)[N])
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
                static_for<0, N / len>([&](auto B) {
                    constexpr int i = decltype(B)::value * len + k;
                    cpx<T> u = t[i];
                    cpx<T> w = twmul<T, k, len, SIGN>(t[i + half]);
                    t[i] = cadd<T>(u, w);
                    t[i + half] = csub<T>(u, w);
                });
            });
        });
        static_for<0, N>([&](auto I) { v[decltype(I)::value] = t[decltype(I)::value]; });
    }
}

} // namespace bec
