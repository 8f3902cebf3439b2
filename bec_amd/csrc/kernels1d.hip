// The 1D two-component program (programs/bec_wigner_2d.cpp/wigner_1d.cpp) on
// the GPU: calculate_steady_state (wigner_1d.cpp:166-243) and main_loop
// (wigner_1d.cpp:245-445), FP64.
//
// A trajectory is one wavefront.  Its state lives in registers: lane l holds
// cells l, l + 64, ... (up to eight per lane, nv <= 512) of both components,
// in whichever space the step is in.  The program's lattice has an odd,
// generally prime, number of cells (41), so the transforms are direct DFTs:
// the state is staged into LDS, and each lane accumulates its outputs
// sum_j x_j w^(j k) over the staged inputs (broadcast reads) with the
// exponent index (j k) mod nv stepped incrementally into a twiddle table in
// LDS.  Both components share each table read.  HBM sees the ground state
// read once per trajectory and the samples written.
//
// Conventions of the program (and of the oracle's restatement): the k-space
// factor is -k^2 with k = dk j for 2j <= nv else dk (j - nv), dk = pi/zmax;
// x -> k is the +i transform over nv cells, k -> x the unnormalised -i
// transform; the trap term is mu - z^2/4 with z = -zmax + dz j.
#include <hip/hip_runtime.h>
#include <math.h>

#include "backend1d.hpp"
#include "philox.hpp"

namespace bec {

namespace {

constexpr int LANES = 64;
constexpr int MAXNV = W1D_MAX_NV;

__device__ __forceinline__ double2 cm(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cs(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double nrm(double2 a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ double kval(int j, int n, double dk) { return (2 * j > n) ? dk * (j - n) : dk * j; }

// tw[m] = exp(-2 pi i m / n)
__device__ void make_twiddles(double2 *tw, int n)
{
    for (int m = threadIdx.x; m < n; m += LANES) {
        double s, c;
        sincospi(-2.0 * m / n, &s, &c);
        tw[m] = make_double2(c, s);
    }
    __syncthreads();
}

// Stage the lane's cells of one or two vectors into LDS.
template <int R, bool TWO>
__device__ __forceinline__ void stage(const double2 (&a)[R], const double2 (&b)[R], double2 *sa, double2 *sb, int n)
{
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = threadIdx.x + LANES * r;
        if (k < n) {
            sa[k] = a[r];
            if (TWO) sb[k] = b[r];
        }
    }
    __syncthreads();
}

// out[k] = sum_j s[j] exp(sign 2 pi i j k / n) for the lane's k, from the staged
// inputs; ends with a barrier so the staging buffers may be overwritten.
template <int R, bool TWO>
__device__ __forceinline__ void dft(const double2 *sa, const double2 *sb, const double2 *tw, int n, int sign,
                                    double2 (&oa)[R], double2 (&ob)[R])
{
    const double sg = sign < 0 ? 1.0 : -1.0;
    int step[R], idx[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = threadIdx.x + LANES * r;
        step[r] = k < n ? k : 0;
        idx[r] = 0;
        oa[r] = make_double2(0.0, 0.0);
        ob[r] = make_double2(0.0, 0.0);
    }
    // unrolled so that several table reads are in flight: one wavefront per
    // trajectory leaves LDS latency exposed otherwise
#pragma unroll 4
    for (int j = 0; j < n; ++j) {
        const double2 xa = sa[j];
        const double2 xb = TWO ? sb[j] : xa;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double2 w = tw[idx[r]];
            w.y *= sg;
            oa[r].x = fma(xa.x, w.x, fma(-xa.y, w.y, oa[r].x));
            oa[r].y = fma(xa.x, w.y, fma(xa.y, w.x, oa[r].y));
            if (TWO) {
                ob[r].x = fma(xb.x, w.x, fma(-xb.y, w.y, ob[r].x));
                ob[r].y = fma(xb.x, w.y, fma(xb.y, w.x, ob[r].y));
            }
            idx[r] += step[r];
            idx[r] -= idx[r] >= n ? n : 0;
        }
    }
    __syncthreads();
}

// x -> k of the register state: the +i transform and 1/n
template <int R, bool TWO>
__device__ __forceinline__ void to_k(double2 (&a)[R], double2 (&b)[R], double2 *sa, double2 *sb, const double2 *tw,
                                     int n)
{
    stage<R, TWO>(a, b, sa, sb, n);
    dft<R, TWO>(sa, sb, tw, n, +1, a, b);
    const double inv = 1.0 / n;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        a[r] = cs(a[r], inv);
        b[r] = cs(b[r], inv);
    }
}

template <int R, bool TWO>
__device__ __forceinline__ void to_x(double2 (&a)[R], double2 (&b)[R], double2 *sa, double2 *sb, const double2 *tw,
                                     int n)
{
    stage<R, TWO>(a, b, sa, sb, n);
    dft<R, TWO>(sa, sb, tw, n, -1, a, b);
}

template <int K> __device__ __forceinline__ void wave_sum(double (&v)[K])
{
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, LANES);
}

// calculate_steady_state (wigner_1d.cpp:166-243): uniform n0 start, ntG - 1
// imaginary-time steps with the itmax midpoint iteration, the N series after
// each pass; then the residual of the x-space result.  One wavefront.
template <int R>
__global__ __launch_bounds__(LANES) void w1d_ground_state_kernel(W1dParams p, int n, double2 *gs_out, double *ntot,
                                                                 double *rep_out)
{
    extern __shared__ double2 lds[];   // twiddles, then two staging vectors: 3 nv
    double2 *tw = lds, *sa = lds + n, *sb = lds + 2 * n;
    make_twiddles(tw, n);
    const double dk = M_PI / p.zmax;
    double2 a[R], u[R];
    double propG[R], pG[R], k2[R], z[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = threadIdx.x + LANES * r;
        const double kv = kval(k, n, dk);
        z[r] = -p.zmax + p.dz * k;
        k2[r] = kv * kv;
        propG[r] = exp(-k2[r] * p.dt / 2.0);
        pG[r] = p.mu - z[r] * z[r] / 4.0;
        a[r] = make_double2(k < n ? p.n0 : 0.0, 0.0);
    }
    to_k<R, false>(a, u, sa, sb, tw, n);
    double N0 = 0.0;
    for (int it = 0; it < p.ntG; ++it) {
        if (it > 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) a[r] = cs(a[r], propG[r]);
            to_x<R, false>(a, u, sa, sb, tw, n);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double2 a1 = a[r];
                double2 v = a1;
                double da = 0.0;
                for (int i = 0; i < p.itmax; ++i) {
                    da = exp(p.dt / 2.0 * (pG[r] - p.g11 * nrm(v)));
                    v = cs(a1, da);
                }
                a[r] = cs(v, da);
            }
            to_k<R, false>(a, u, sa, sb, tw, n);
#pragma unroll
            for (int r = 0; r < R; ++r) a[r] = cs(a[r], propG[r]);
        }
        double s[1] = {0.0};
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (threadIdx.x + LANES * r < n) s[0] += nrm(a[r]);
        wave_sum<1>(s);
        N0 = s[0] * p.dz * n;
        if (ntot && threadIdx.x == 0) ntot[it] = N0;
    }
    to_x<R, false>(a, u, sa, sb, tw, n);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = threadIdx.x + LANES * r;
        if (k < n) gs_out[k] = a[r];
    }
    // residual of (-d^2/dz^2 + z^2/4 + g11 |psi|^2) psi = mu psi
    double2 h[R];
#pragma unroll
    for (int r = 0; r < R; ++r) h[r] = a[r];
    to_k<R, false>(h, u, sa, sb, tw, n);
#pragma unroll
    for (int r = 0; r < R; ++r) h[r] = cs(h[r], k2[r]);
    to_x<R, false>(h, u, sa, sb, tw, n);
    double s2[2] = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (threadIdx.x + LANES * r >= n) continue;
        const double v = (p.mu - pG[r]) + p.g11 * nrm(a[r]);
        h[r] = make_double2(h[r].x + v * a[r].x, h[r].y + v * a[r].y);
        s2[0] += a[r].x * h[r].x + a[r].y * h[r].y;
        s2[1] += nrm(a[r]);
    }
    wave_sum<2>(s2);
    const double mu = s2[0] / s2[1];
    double s1[1] = {0.0};
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (threadIdx.x + LANES * r < n) s1[0] += nrm(make_double2(h[r].x - mu * a[r].x, h[r].y - mu * a[r].y));
    wave_sum<1>(s1);
    if (threadIdx.x == 0) {
        rep_out[0] = N0;
        rep_out[1] = sqrt(s1[0] / s2[1]) / fabs(mu);
        rep_out[2] = mu;
    }
}

// main_loop (wigner_1d.cpp:245-445) for one trajectory per wavefront.  meas:
// [traj][ntW][3] = {N1, N2, visibility}; avn: [traj][nv] x-space density
// minus the vacuum term after the last step.
template <int R>
__global__ __launch_bounds__(LANES) void w1d_ensemble_kernel(W1dParams p, int n, const double2 *gs, double *meas,
                                                             double *avn, uint32_t k0, uint32_t k1)
{
    extern __shared__ double2 lds[];   // twiddles, then two staging vectors: 3 nv
    double2 *tw = lds, *sa = lds + n, *sb = lds + 2 * n;
    const int traj = blockIdx.x;
    make_twiddles(tw, n);
    const double dk = M_PI / p.zmax, nd = p.noise_dev / 0.5;
    double2 a[R], b[R], kw[R];
    double base[R];
    // initial state: ground state + noise in x-space (:281-290), then to k-space
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = threadIdx.x + LANES * r;
        const double kv = kval(k, n, dk), z = -p.zmax + p.dz * k;
        double sn, c;
        sincos(kv * kv * p.dt / 2.0, &sn, &c);
        kw[r] = make_double2(c, sn);   // exp(+i k^2 dt/2)
        base[r] = p.mu - z * z / 4.0;
        a[r] = b[r] = make_double2(0.0, 0.0);
        if (k < n) {
            double e[4];
            noise4((uint32_t)k, (uint32_t)traj, 0u, k0, k1, e);
            const double2 g = gs[k];
            a[r] = make_double2(g.x + p.Va * nd * e[0], g.y + p.Va * nd * e[1]);
            b[r] = make_double2(p.Vb * nd * e[2], p.Vb * nd * e[3]);
        }
    }
    to_k<R, true>(a, b, sa, sb, tw, n);
    const double r2 = sqrt(2.0), V = (p.Va + p.Vb) / 2.0;
    for (int nn = -p.ntW; nn < p.ntW; ++nn) {
        if (nn > -p.ntW) {
            if (nn == 0) {   // pi/2 pulse (:308-318)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double2 x = a[r], y = b[r];
                    b[r] = make_double2((y.x + x.y) / r2, (y.y - x.x) / r2);   // (b - i a)/sqrt 2
                    a[r] = make_double2((x.x + y.y) / r2, (x.y - y.x) / r2);   // (a - i b)/sqrt 2
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                a[r] = cm(a[r], kw[r]);
                b[r] = cm(b[r], kw[r]);
            }
            to_x<R, true>(a, b, sa, sb, tw, n);
#pragma unroll
            for (int r = 0; r < R; ++r) {   // single-shot nonlinear step (:331-337)
                const double na = nrm(a[r]), nb = nrm(b[r]);
                const double dA = -p.g11 * na - p.g12 * nb + base[r] + 0.5 * p.E;
                const double dB = -p.g22 * nb - p.g12 * na + base[r] - 0.5 * p.E;
                double sa_, ca, sb_, cb;
                sincos(-p.dt * dA, &sa_, &ca);
                sincos(-p.dt * dB, &sb_, &cb);
                a[r] = cm(a[r], make_double2(ca, sa_));
                b[r] = cm(b[r], make_double2(cb, sb_));
            }
            to_k<R, true>(a, b, sa, sb, tw, n);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                a[r] = cm(a[r], kw[r]);
                b[r] = cm(b[r], kw[r]);
            }
        }
        if (nn >= 0) {   // second pulse U^+ and measurement in k-space (:358-391)
            double s[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (threadIdx.x + LANES * r >= n) continue;
                const double2 x = a[r], y = b[r];
                const double2 b1 = make_double2((y.x - x.y) / r2, (y.y + x.x) / r2);   // (b + i a)/sqrt 2
                const double2 a1 = make_double2((x.x - y.y) / r2, (x.y + y.x) / r2);   // (a + i b)/sqrt 2
                s[0] += nrm(a1);
                s[1] += nrm(b1);
                s[2] += x.x * y.x + x.y * y.y;
                s[3] += x.x * y.y - x.y * y.x;
                s[4] += nrm(x);
                s[5] += nrm(y);
            }
            wave_sum<6>(s);
            if (threadIdx.x == 0) {
                double *m = meas + ((long)traj * p.ntW + nn) * 3;
                m[0] = n * (s[0] * p.dz - 0.5 * V);
                m[1] = n * (s[1] * p.dz - 0.5 * V);
                m[2] = 2.0 * sqrt(s[2] * s[2] + s[3] * s[3]) / (s[4] + s[5]);
            }
        }
    }
    // mean density after the last step (:396-400)
    to_x<R, false>(a, b, sa, sb, tw, n);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = threadIdx.x + LANES * r;
        if (k < n) avn[(long)traj * n + k] = nrm(a[r]) - 0.5 * V / p.dz;
    }
}

// Sums over trajectories in trajectory order (the order the program
// accumulates them in): thread t < ntW sums sample t, the next nv threads the
// cells of avn.
__global__ __launch_bounds__(256) void w1d_reduce_kernel(int ne, int ntW, int n, const double *meas, const double *avn,
                                                         double *sums, double *avn_sum)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < ntW) {
        double s[5] = {0, 0, 0, 0, 0};
        for (int j = 0; j < ne; ++j) {
            const double *q = meas + ((long)j * ntW + t) * 3;
            s[0] += q[0];
            s[1] += q[1];
            s[2] += q[0] * q[0];
            s[3] += q[1] * q[1];
            s[4] += q[2];
        }
        for (int i = 0; i < 5; ++i) sums[(long)i * ntW + t] = s[i];
    } else if (t < ntW + n) {
        const int c = t - ntW;
        double s = 0.0;
        for (int j = 0; j < ne; ++j) s += avn[(long)j * n + c];
        avn_sum[c] = s;
    }
}

size_t lds_bytes(int n) { return sizeof(double2) * 3 * (size_t)n; }

template <int R> hipError_t launch_gs(const W1dParams &p, int n, double2 *gs, double *ntot, double *rep, hipStream_t s)
{
    hipLaunchKernelGGL(w1d_ground_state_kernel<R>, dim3(1), dim3(LANES), lds_bytes(n), s, p, n, gs, ntot, rep);
    return hipGetLastError();
}

template <int R>
hipError_t launch_ens(const W1dParams &p, int n, const double2 *gs, double *meas, double *avn, hipStream_t s)
{
    const uint32_t k0 = (uint32_t)p.seed, k1 = (uint32_t)(p.seed >> 32);
    hipLaunchKernelGGL(w1d_ensemble_kernel<R>, dim3((unsigned)p.ne), dim3(LANES), lds_bytes(n), s, p, n, gs, meas, avn,
                       k0, k1);
    return hipGetLastError();
}

} // namespace

hipError_t run_w1d_ground_state(const W1dParams &p, int n, double2 *gs_out, double *ntot, double *rep_out,
                                hipStream_t s)
{
    if (n <= 64) return launch_gs<1>(p, n, gs_out, ntot, rep_out, s);
    if (n <= 128) return launch_gs<2>(p, n, gs_out, ntot, rep_out, s);
    if (n <= 256) return launch_gs<4>(p, n, gs_out, ntot, rep_out, s);
    if (n <= MAXNV) return launch_gs<8>(p, n, gs_out, ntot, rep_out, s);
    return hipErrorInvalidValue;
}

hipError_t run_w1d_ensemble(const W1dParams &p, int n, const double2 *gs, double *meas, double *avn, hipStream_t s)
{
    if (n <= 64) return launch_ens<1>(p, n, gs, meas, avn, s);
    if (n <= 128) return launch_ens<2>(p, n, gs, meas, avn, s);
    if (n <= 256) return launch_ens<4>(p, n, gs, meas, avn, s);
    if (n <= MAXNV) return launch_ens<8>(p, n, gs, meas, avn, s);
    return hipErrorInvalidValue;
}

hipError_t run_w1d_reduce(const W1dParams &p, int n, const double *meas, const double *avn, double *sums,
                          double *avn_sum, hipStream_t s)
{
    const int threads = p.ntW + n;
    hipLaunchKernelGGL(w1d_reduce_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, p.ne, p.ntW, n,
                       meas, avn, sums, avn_sum);
    return hipGetLastError();
}

} // namespace bec
