// Shared device-side definitions for the gfx950 split-step kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bec {

// ---------------------------------------------------------------- complex
template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename T> using cpx = typename vec2<T>::type;

template <typename T> __host__ __device__ __forceinline__ cpx<T> mk(T re, T im)
{
    cpx<T> r;
    r.x = re;
    r.y = im;
    return r;
}
template <typename T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * b with its fused multiply-adds written out, so that every kernel that
// applies it rounds alike (which product of cmul the compiler fuses can differ
// between two kernels that inline it)
template <typename T> __device__ __forceinline__ cpx<T> cmul_fixed(cpx<T> a, cpx<T> b)
{
    if constexpr (sizeof(T) == 4) return mk<T>(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
    else return mk<T>(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
// conj(a) * b
template <typename T> __device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b)
{
    return mk<T>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename T> __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T> __device__ __forceinline__ T cnorm(cpx<T> a) { return a.x * a.x + a.y * a.y; }

// ------------------------------------------------- compile-time sin / cos
// Twiddles of the in-register DFTs are constants.  They are evaluated at
// compile time from the Taylor series on [-pi/4, pi/4] after reduction by the
// exact octant symmetries, which is accurate to a few ulp of double.
namespace ce {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double sin_series(double x)
{
    double term = x, sum = x, x2 = x * x;
    for (int n = 1; n < 14; ++n) {
        term *= -x2 / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cos_series(double x)
{
    double term = 1.0, sum = 1.0, x2 = x * x;
    for (int n = 1; n < 14; ++n) {
        term *= -x2 / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
// cos and sin of 2*pi*k/n for integers (exact octant reduction)
constexpr double cos2pi(long k, long n)
{
    k %= n;
    if (k < 0) k += n;
    // angle = 2 pi k / n in [0, 2 pi); reduce to octant using 8k/n
    long k8 = 8 * k;
    long oct = k8 / n;            // 0..7
    long rem = k8 - oct * n;      // angle within octant scaled: (rem / n) * (pi/4)
    double r = kPi / 4.0 * (double)rem / (double)n;
    double c = cos_series(r), s = sin_series(r);
    double c2 = cos_series(kPi / 4.0 - r), s2 = sin_series(kPi / 4.0 - r);
    switch (oct) {
    case 0: return c;
    case 1: return s2;
    case 2: return -s;
    case 3: return -c2;
    case 4: return -c;
    case 5: return -s2;
    case 6: return s;
    default: return c2;
    }
}
} // namespace ce

// sin(2 pi k / n) = cos(2 pi k / n - pi / 2) = cos(2 pi (4k - n) / (4n))
constexpr double csin2pi(long k, long n) { return ce::cos2pi(4 * k - n, 4 * n); }
constexpr double ccos2pi(long k, long n) { return ce::cos2pi(k, n); }

// ------------------------------------------------------------ parameters
// Per-record physical parameters in natural units.  One record serves a
// batch of trajectories that share a grid and a Hamiltonian.
struct TrajParams {
    double xmax, ymax, zmax;
    double dx, dy, dz;
    double inv_lambda2;          // 1 / lambda^2
    double g11, g12, g22;
    double l111, l12, l22;
    double detuning;
    double dt;                   // real-time step (natural units)
    double dt_steady;            // imaginary-time step
    double N0;                   // target atom number (ground state)
    double vac;                  // V / (2 dV): vacuum density subtracted by density()
    int itmax;
    int pad_;
};

__device__ __forceinline__ double potential_at(const TrajParams &p, int x, int y, int z)
{
    double xx = -p.xmax + p.dx * x, yy = -p.ymax + p.dy * y, zz = -p.zmax + p.dz * z;
    return 0.5 * (xx * xx + yy * yy + zz * zz * p.inv_lambda2);
}

template <typename T> __device__ __forceinline__ T potential_t(const TrajParams &p, int x, int y, int z)
{
    T xx = (T)(-p.xmax) + (T)p.dx * (T)x;
    T yy = (T)(-p.ymax) + (T)p.dy * (T)y;
    T zz = (T)(-p.zmax) + (T)p.dz * (T)z;
    return (T)0.5 * (xx * xx + yy * yy + zz * zz * (T)p.inv_lambda2);
}

// Wave-wide sum over a group of G consecutive lanes (G a power of two <= 64).
template <int G> __device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

} // namespace bec
