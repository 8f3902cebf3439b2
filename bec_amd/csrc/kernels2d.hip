// The 2D two-component program (programs/bec_wigner_2d.cpp/wigner.cpp) on the
// GPU: calculate_steady_state (wigner.cpp:196-275) and main_loop
// (wigner.cpp:278-477), FP64.
//
// At these lattices (up to 64 x 64) a whole trajectory fits one workgroup's
// LDS: both components, 64 x 65 complex doubles each (rows padded by one
// element against bank conflicts), 133 KB of the 160 KB.  Each workgroup (1024
// threads, four wavefronts per SIMD) runs
// its trajectory's entire evolution -- noise, pulses, every split step and the
// measurements -- with the state resident in LDS; HBM sees only the ground
// state read at the start and the measurements written.  The 2D FFTs are
// in-place radix-2 transforms along rows, then columns, with twiddle tables
// in LDS.  The ground state runs the same way in one workgroup.
//
// Conventions of the program (and of the oracle's restatement): the k-space
// kinetic factor is -k^2 with k = dk i for 2i <= n else dk (i - n),
// dk = pi/xmax, dx = 2 xmax/(n - 1); x -> k is the +i transform over n cells,
// k -> x the unnormalised -i transform; the trap term is mu - r^2/4.
#include <hip/hip_runtime.h>
#include <math.h>

#include "backend2d.hpp"
#include "philox.hpp"

namespace bec {

namespace {

constexpr int THREADS = 1024;   // 16 wavefronts: four per SIMD hide the LDS and FP64 latencies
constexpr int SCRATCH = THREADS / 64 * 8;   // block_sum: up to 8 sums per wavefront
constexpr int MAXN = 64;
constexpr int LDMAX = MAXN + 1;
constexpr int CPT = MAXN * MAXN / THREADS;   // cells per thread at the largest lattice

struct Geo {
    int nx, ny, ld, n, lgx, lgy;
    double dx, dy, dkx, dky, xmax, ymax;
};

__device__ __forceinline__ double2 cm(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cs(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double nrm(double2 a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ int brev(int i, int bits) { return (int)(__brev((unsigned)i) >> (32 - bits)); }

__device__ __forceinline__ double kval(int i, int n, double dk) { return (2 * i > n) ? dk * (i - n) : dk * i; }

// cell c = y nx + x of thread-owned element j
__device__ __forceinline__ int cell_of(int j) { return threadIdx.x + THREADS * j; }
__device__ __forceinline__ int lds_of(const Geo &g, int c) { return (c >> g.lgx) * g.ld + (c & (g.nx - 1)); }

// In-place radix-2 DIT transforms of `lines` lines of length n (stride es
// between elements, ls between lines) in each of the buffers; twiddles
// tw[k] = exp(-2 pi i k / n), conjugated for sign = +1.
__device__ void fft_lines(double2 *b0, double2 *b1, int nb, int n, int lg, int lines, int es, int ls, int sign,
                          const double2 *tw)
{
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nb * lines * n; idx += THREADS) {
        const int b = idx / (lines * n), r = idx - b * lines * n, line = r / n, i = r - line * n;
        const int j = brev(i, lg);
        if (i < j) {
            double2 *p = (b ? b1 : b0) + line * ls;
            const double2 t = p[i * es];
            p[i * es] = p[j * es];
            p[j * es] = t;
        }
    }
    __syncthreads();
    const int hn = n / 2;
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int idx = tid; idx < nb * lines * hn; idx += THREADS) {
            const int b = idx / (lines * hn), r = idx - b * lines * hn, line = r / hn, q = r - line * hn;
            const int grp = q / half, k = q - grp * half;
            const int i = grp * len + k, j = i + half;
            double2 w = tw[k * step];
            if (sign > 0) w.y = -w.y;
            double2 *p = (b ? b1 : b0) + line * ls;
            const double2 u = p[i * es], v = cm(p[j * es], w);
            p[i * es] = make_double2(u.x + v.x, u.y + v.y);
            p[j * es] = make_double2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
    }
}

// 2D transform of one (b1 = null) or two buffers: rows (x), then columns (y)
__device__ void fft2d(const Geo &g, double2 *b0, double2 *b1, int sign, const double2 *twx, const double2 *twy)
{
    const int nb = b1 ? 2 : 1;
    fft_lines(b0, b1, nb, g.nx, g.lgx, g.ny, 1, g.ld, sign, twx);
    fft_lines(b0, b1, nb, g.ny, g.lgy, g.nx, g.ld, 1, sign, twy);
}

// x -> k: the +i transform and 1/n
__device__ void to_k(const Geo &g, double2 *b0, double2 *b1, const double2 *twx, const double2 *twy)
{
    fft2d(g, b0, b1, +1, twx, twy);
    const double inv = 1.0 / g.n;
    for (int c = threadIdx.x; c < g.n; c += THREADS) {
        const int o = lds_of(g, c);
        b0[o] = cs(b0[o], inv);
        if (b1) b1[o] = cs(b1[o], inv);
    }
    __syncthreads();
}

// Sum of K doubles over the workgroup (wave shuffles, then LDS across waves);
// every thread receives the totals.
template <int K> __device__ void block_sum(double (&v)[K], double *scratch)
{
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) scratch[wave * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int w = 0; w < THREADS / 64; ++w) s += scratch[w * K + k];
        v[k] = s;
    }
    __syncthreads();
}

__device__ void make_twiddles(const Geo &g, double2 *twx, double2 *twy)
{
    for (int k = threadIdx.x; k < g.nx / 2; k += THREADS) {
        double s, c;
        sincospi(-2.0 * k / g.nx, &s, &c);
        twx[k] = make_double2(c, s);
    }
    for (int k = threadIdx.x; k < g.ny / 2; k += THREADS) {
        double s, c;
        sincospi(-2.0 * k / g.ny, &s, &c);
        twy[k] = make_double2(c, s);
    }
    __syncthreads();
}

__device__ __forceinline__ double r2_of(const Geo &g, int c)
{
    const int x = c & (g.nx - 1), y = c >> g.lgx;
    const double xx = -g.xmax + g.dx * x, yy = -g.ymax + g.dy * y;
    return xx * xx + yy * yy;
}

__device__ __forceinline__ double k2_of(const Geo &g, int c)
{
    const int x = c & (g.nx - 1), y = c >> g.lgx;
    const double kx = kval(x, g.nx, g.dkx), ky = kval(y, g.ny, g.dky);
    return kx * kx + ky * ky;
}

// Residual ||H psi - mu psi|| / (|mu| ||psi||) of the x-space state in P,
// H = -lap + r^2/4 + g11 |psi|^2; Hs is scratch.  Returns mu in *mu_out.
__device__ double residual(const Geo &g, const W2dParams &p, double2 *P, double2 *Hs, const double2 *twx,
                           const double2 *twy, double *scratch, double *mu_out)
{
    for (int c = threadIdx.x; c < g.n; c += THREADS) Hs[lds_of(g, c)] = P[lds_of(g, c)];
    __syncthreads();
    to_k(g, Hs, nullptr, twx, twy);
    for (int c = threadIdx.x; c < g.n; c += THREADS) Hs[lds_of(g, c)] = cs(Hs[lds_of(g, c)], k2_of(g, c));
    __syncthreads();
    fft2d(g, Hs, nullptr, -1, twx, twy);
    double s2[2] = {0.0, 0.0};
    for (int c = threadIdx.x; c < g.n; c += THREADS) {
        const int o = lds_of(g, c);
        const double2 ps = P[o];
        const double v = r2_of(g, c) / 4.0 + p.g11 * nrm(ps);
        const double2 h = make_double2(Hs[o].x + v * ps.x, Hs[o].y + v * ps.y);
        Hs[o] = h;
        s2[0] += ps.x * h.x + ps.y * h.y;
        s2[1] += nrm(ps);
    }
    block_sum<2>(s2, scratch);
    const double mu = s2[0] / s2[1];
    double s1[1] = {0.0};
    for (int c = threadIdx.x; c < g.n; c += THREADS) {
        const int o = lds_of(g, c);
        s1[0] += nrm(make_double2(Hs[o].x - mu * P[o].x, Hs[o].y - mu * P[o].y));
    }
    block_sum<1>(s1, scratch);
    if (mu_out) *mu_out = mu;
    return sqrt(s1[0] / s2[1]) / fabs(mu);
}

__device__ Geo make_geo(const W2dParams &p)
{
    Geo g;
    g.nx = p.nvx;
    g.ny = p.nvy;
    g.ld = p.nvx + 1;
    g.n = p.nvx * p.nvy;
    g.lgx = 31 - __clz(p.nvx);
    g.lgy = 31 - __clz(p.nvy);
    g.xmax = p.xmax;
    g.ymax = p.ymax;
    g.dx = 2.0 * p.xmax / (p.nvx - 1);
    g.dy = 2.0 * p.ymax / (p.nvy - 1);
    g.dkx = M_PI / p.xmax;
    g.dky = M_PI / p.ymax;
    return g;
}

// calculate_steady_state (wigner.cpp:196-275): uniform n0 start, imaginary
// time with the itmax midpoint iteration; ntG - 1 steps (reference rule) or
// until the residual is below gs_tol.  One workgroup.
__global__ __launch_bounds__(THREADS) void w2d_ground_state_kernel(W2dParams p, double2 *gs_out, double *rep_out)
{
    __shared__ double2 A[MAXN * LDMAX], S[MAXN * LDMAX];
    __shared__ double2 twx[MAXN / 2], twy[MAXN / 2];
    __shared__ double scratch[SCRATCH];
    const Geo g = make_geo(p);
    make_twiddles(g, twx, twy);
    for (int c = threadIdx.x; c < g.n; c += THREADS) A[lds_of(g, c)] = make_double2(p.n0, 0.0);
    __syncthreads();
    to_k(g, A, nullptr, twx, twy);
    const int total = p.gs_rule == 0 ? p.ntG - 1 : p.gs_max_iter;
    int iters = 0;
    double res = -1.0;
    for (int it = 0; it < total; ++it) {
        for (int c = threadIdx.x; c < g.n; c += THREADS)
            A[lds_of(g, c)] = cs(A[lds_of(g, c)], exp(-k2_of(g, c) * p.dtG / 2.0));
        __syncthreads();
        fft2d(g, A, nullptr, -1, twx, twy);
        for (int j = 0; j < CPT; ++j) {
            const int c = cell_of(j);
            if (c >= g.n) break;
            const int o = lds_of(g, c);
            const double pG = p.mu - r2_of(g, c) / 4.0;
            const double2 a1 = A[o];
            double2 a = a1;
            double da = 0.0;
            for (int k = 0; k < p.itmax; ++k) {
                da = exp(p.dtG / 2.0 * (pG - p.g11 * nrm(a)));
                a = cs(a1, da);
            }
            A[o] = cs(a, da);
        }
        __syncthreads();
        to_k(g, A, nullptr, twx, twy);
        for (int c = threadIdx.x; c < g.n; c += THREADS)
            A[lds_of(g, c)] = cs(A[lds_of(g, c)], exp(-k2_of(g, c) * p.dtG / 2.0));
        __syncthreads();
        ++iters;
        if (p.gs_rule == 1) {
            // residual of the x-space iterate: the k-space iterate is parked in
            // gs_out while both LDS buffers serve the residual, then restored
            for (int c = threadIdx.x; c < g.n; c += THREADS) {
                gs_out[c] = A[lds_of(g, c)];
                S[lds_of(g, c)] = A[lds_of(g, c)];
            }
            __syncthreads();
            fft2d(g, S, nullptr, -1, twx, twy);
            res = residual(g, p, S, A, twx, twy, scratch, nullptr);
            for (int c = threadIdx.x; c < g.n; c += THREADS) A[lds_of(g, c)] = gs_out[c];
            __syncthreads();
            if (res < p.gs_tol) break;
        }
    }
    double s1[1] = {0.0};
    for (int c = threadIdx.x; c < g.n; c += THREADS) s1[0] += nrm(A[lds_of(g, c)]);
    block_sum<1>(s1, scratch);
    fft2d(g, A, nullptr, -1, twx, twy);
    for (int c = threadIdx.x; c < g.n; c += THREADS) gs_out[c] = A[lds_of(g, c)];
    __syncthreads();
    double mu = 0.0;
    const double r = residual(g, p, A, S, twx, twy, scratch, &mu);
    if (threadIdx.x == 0) {
        rep_out[0] = s1[0] * g.dx * g.dy * g.n;
        rep_out[1] = iters;
        rep_out[2] = r;
        rep_out[3] = mu;
    }
}

// Ground state of n_atoms atoms (gs_rule 2; the program's own loop fixes mu):
// the preconditioned imaginary-time gradient flow
//   psi <- psi - tau P (H - mu) psi + beta (psi - psi_prev),  P = 1/(alpha + k^2),
// renormalised to dx dy sum |psi|^2 = n_atoms after every update, from the
// Thomas-Fermi profile sqrt(max(mu - r^2/4, 0) / g11), alpha = mu, until the
// residual ||H psi - mu psi|| / (|mu| ||psi||) is below gs_tol.  One
// workgroup, the x-space iterate and H psi in LDS, the previous iterate in
// global memory (pv, one cell per thread, read and written by that thread).
// The tail (atoms, x-space state, final residual) is that of the program's
// loop below.
__global__ __launch_bounds__(THREADS) void w2d_flow_gs_kernel(W2dParams p, double2 *gs_out, double2 *pv,
                                                             double *rep_out)
{
    __shared__ double2 A[MAXN * LDMAX], S[MAXN * LDMAX];
    __shared__ double2 twx[MAXN / 2], twy[MAXN / 2];
    __shared__ double scratch[SCRATCH];
    const Geo g = make_geo(p);
    make_twiddles(g, twx, twy);
    const double alpha = p.mu > 0 ? p.mu : 1.0;
    double s0[2] = {0.0, 0.0};
    for (int c = threadIdx.x; c < g.n; c += THREADS) {
        double v = p.mu - r2_of(g, c) / 4.0;
        v = v > 0 ? sqrt(v / p.g11) : 0.0;
        A[lds_of(g, c)] = make_double2(v, 0.0);
        s0[0] += v * v;
        s0[1] += v > 0 ? 1.0 : 0.0;
    }
    block_sum<2>(s0, scratch);
    if (s0[1] == 0.0) {   // no Thomas-Fermi cloud on the lattice: uniform start
        for (int c = threadIdx.x; c < g.n; c += THREADS) A[lds_of(g, c)] = make_double2(1.0, 0.0);
        s0[0] = g.n;
    }
    double sc = sqrt(p.n_atoms / (s0[0] * g.dx * g.dy));
    for (int c = threadIdx.x; c < g.n; c += THREADS) {
        const int o = lds_of(g, c);
        A[o] = cs(A[o], sc);
        pv[c] = A[o];
    }
    __syncthreads();
    int iters = 0;
    for (;;) {
        // S = H A
        for (int c = threadIdx.x; c < g.n; c += THREADS) S[lds_of(g, c)] = A[lds_of(g, c)];
        __syncthreads();
        to_k(g, S, nullptr, twx, twy);
        for (int c = threadIdx.x; c < g.n; c += THREADS) S[lds_of(g, c)] = cs(S[lds_of(g, c)], k2_of(g, c));
        __syncthreads();
        fft2d(g, S, nullptr, -1, twx, twy);
        double s2[2] = {0.0, 0.0};
        for (int c = threadIdx.x; c < g.n; c += THREADS) {
            const int o = lds_of(g, c);
            const double2 ps = A[o];
            const double v = r2_of(g, c) / 4.0 + p.g11 * nrm(ps);
            const double2 h = make_double2(S[o].x + v * ps.x, S[o].y + v * ps.y);
            S[o] = h;
            s2[0] += ps.x * h.x + ps.y * h.y;
            s2[1] += nrm(ps);
        }
        block_sum<2>(s2, scratch);
        const double mu = s2[0] / s2[1];
        double s3[1] = {0.0};
        for (int c = threadIdx.x; c < g.n; c += THREADS) {
            const int o = lds_of(g, c);
            const double2 r = make_double2(S[o].x - mu * A[o].x, S[o].y - mu * A[o].y);
            S[o] = r;
            s3[0] += nrm(r);
        }
        block_sum<1>(s3, scratch);
        const double res = sqrt(s3[0] / s2[1]) / fabs(mu);
        if (!(res == res) || res < p.gs_tol || iters >= p.gs_max_iter) break;   // uniform across the workgroup
        // S = P (H - mu) A
        to_k(g, S, nullptr, twx, twy);
        for (int c = threadIdx.x; c < g.n; c += THREADS)
            S[lds_of(g, c)] = cs(S[lds_of(g, c)], 1.0 / (alpha + k2_of(g, c)));
        __syncthreads();
        fft2d(g, S, nullptr, -1, twx, twy);
        double s4[1] = {0.0};
        for (int c = threadIdx.x; c < g.n; c += THREADS) {
            const int o = lds_of(g, c);
            const double2 cur = A[o], prev = pv[c], d = S[o];
            const double2 nw = make_double2(cur.x - p.gs_tau * d.x + p.gs_beta * (cur.x - prev.x),
                                            cur.y - p.gs_tau * d.y + p.gs_beta * (cur.y - prev.y));
            pv[c] = cur;
            A[o] = nw;
            s4[0] += nrm(nw);
        }
        block_sum<1>(s4, scratch);
        sc = sqrt(p.n_atoms / (s4[0] * g.dx * g.dy));
        for (int c = threadIdx.x; c < g.n; c += THREADS) A[lds_of(g, c)] = cs(A[lds_of(g, c)], sc);
        __syncthreads();
        ++iters;
    }
    // the program loop's tail, from k-space: atoms, x-space state, final residual
    to_k(g, A, nullptr, twx, twy);
    double s1[1] = {0.0};
    for (int c = threadIdx.x; c < g.n; c += THREADS) s1[0] += nrm(A[lds_of(g, c)]);
    block_sum<1>(s1, scratch);
    fft2d(g, A, nullptr, -1, twx, twy);
    for (int c = threadIdx.x; c < g.n; c += THREADS) gs_out[c] = A[lds_of(g, c)];
    __syncthreads();
    double mu = 0.0;
    const double r = residual(g, p, A, S, twx, twy, scratch, &mu);
    if (threadIdx.x == 0) {
        rep_out[0] = s1[0] * g.dx * g.dy * g.n;
        rep_out[1] = iters;
        rep_out[2] = r;
        rep_out[3] = mu;
    }
}

// main_loop (wigner.cpp:278-477) for one trajectory per workgroup.  meas:
// [traj][n_meas][3] = {N1, N2, visibility}; avn: [traj][cells] x-space
// density minus the vacuum term after the last step (may be null).
__global__ __launch_bounds__(THREADS) void w2d_ensemble_kernel(W2dParams p, const double2 *gs, double *meas,
                                                               double *avn, uint32_t k0, uint32_t k1)
{
    __shared__ double2 A[MAXN * LDMAX], Bf[MAXN * LDMAX];
    __shared__ double2 twx[MAXN / 2], twy[MAXN / 2];
    __shared__ double scratch[SCRATCH];
    const Geo g = make_geo(p);
    const int traj = blockIdx.x;
    make_twiddles(g, twx, twy);
    // initial state: ground state + noise (wigner.cpp:315-324); the program's
    // deviation 1/sqrt(dx dy) on top of the 0.5 of the normal deviates
    const double nd = p.noise_dev / 0.5;
    for (int c = threadIdx.x; c < g.n; c += THREADS) {
        double z[4];
        noise4((uint32_t)c, (uint32_t)traj, 0u, k0, k1, z);
        const int o = lds_of(g, c);
        A[o] = make_double2(gs[c].x + p.Va * nd * z[0], gs[c].y + p.Va * nd * z[1]);
        Bf[o] = make_double2(p.Vb * nd * z[2], p.Vb * nd * z[3]);
    }
    __syncthreads();
    to_k(g, A, Bf, twx, twy);
    // kinetic factors exp(+i k^2 dtW/2) of the thread's cells, in registers
    double2 kw[CPT];
    for (int j = 0; j < CPT; ++j) {
        const int c = cell_of(j);
        double s = 0.0, co = 1.0;
        if (c < g.n) sincos(k2_of(g, c) * p.dtW / 2.0, &s, &co);
        kw[j] = make_double2(co, s);
    }
    const double r2 = sqrt(2.0), V = (p.Va + p.Vb) / 2.0;
    auto kinetic = [&]() {
        for (int j = 0; j < CPT; ++j) {
            const int c = cell_of(j);
            if (c >= g.n) break;
            const int o = lds_of(g, c);
            A[o] = cm(A[o], kw[j]);
            Bf[o] = cm(Bf[o], kw[j]);
        }
        __syncthreads();
    };
    for (int nn = -p.ntW; nn < p.n_meas; ++nn) {
        if (nn > -p.ntW) {
            if (nn == 0) {   // first pi/2 pulse (wigner.cpp:341-353)
                for (int c = threadIdx.x; c < g.n; c += THREADS) {
                    const int o = lds_of(g, c);
                    const double2 a = A[o], b = Bf[o];
                    Bf[o] = make_double2((b.x + a.y) / r2, (b.y - a.x) / r2);   // (b - i a)/sqrt 2
                    A[o] = make_double2((a.x + b.y) / r2, (a.y - b.x) / r2);    // (a - i b)/sqrt 2
                }
                __syncthreads();
            }
            kinetic();
            fft2d(g, A, Bf, -1, twx, twy);
            for (int j = 0; j < CPT; ++j) {
                const int c = cell_of(j);
                if (c >= g.n) break;
                const int o = lds_of(g, c);
                const double2 a = A[o], b = Bf[o];
                const double na = nrm(a), nb = nrm(b);
                const double base = p.mu - r2_of(g, c) / 4.0;
                const double dA = -p.g11 * na - p.g12 * nb + base + 0.5 * p.E;
                const double dB = -p.g22 * nb - p.g12 * na + base - 0.5 * p.E;
                double sa, ca, sb, cb;
                sincos(-p.dtW * dA, &sa, &ca);
                sincos(-p.dtW * dB, &sb, &cb);
                A[o] = cm(a, make_double2(ca, sa));
                Bf[o] = cm(b, make_double2(cb, sb));
            }
            __syncthreads();
            to_k(g, A, Bf, twx, twy);
            kinetic();
        }
        if (nn >= 0) {   // second pulse U^+ and measurement (wigner.cpp:393-425)
            double s[6] = {0, 0, 0, 0, 0, 0};
            for (int c = threadIdx.x; c < g.n; c += THREADS) {
                const int o = lds_of(g, c);
                const double2 a = A[o], b = Bf[o];
                const double2 b1 = make_double2((b.x - a.y) / r2, (b.y + a.x) / r2);   // (b + i a)/sqrt 2
                const double2 aa = make_double2((a.x - b.y) / r2, (a.y + b.x) / r2);   // (a + i b)/sqrt 2
                s[0] += nrm(aa);
                s[1] += nrm(b1);
                s[2] += a.x * b.x + a.y * b.y;
                s[3] += a.x * b.y - a.y * b.x;
                s[4] += nrm(a);
                s[5] += nrm(b);
            }
            block_sum<6>(s, scratch);
            if (threadIdx.x == 0) {
                double *m = meas + ((long)traj * p.n_meas + nn) * 3;
                m[0] = g.n * (s[0] * g.dx * g.dy - 0.5 * V);
                m[1] = g.n * (s[1] * g.dx * g.dy - 0.5 * V);
                m[2] = 2.0 * sqrt(s[2] * s[2] + s[3] * s[3]) / (s[4] + s[5]);
            }
        }
    }
    if (avn) {
        fft2d(g, A, nullptr, -1, twx, twy);
        for (int c = threadIdx.x; c < g.n; c += THREADS)
            avn[(long)traj * g.n + c] = nrm(A[lds_of(g, c)]) - 0.5 * V / g.dx / g.dy;
    }
}

} // namespace

hipError_t run_w2d_ground_state(const W2dParams &p, double2 *gs_out, double2 *pv, double *rep_out, hipStream_t s)
{
    if (p.gs_rule == 2) {
        if (!pv) return hipErrorInvalidValue;
        hipLaunchKernelGGL(w2d_flow_gs_kernel, dim3(1), dim3(THREADS), 0, s, p, gs_out, pv, rep_out);
    } else {
        hipLaunchKernelGGL(w2d_ground_state_kernel, dim3(1), dim3(THREADS), 0, s, p, gs_out, rep_out);
    }
    return hipGetLastError();
}

hipError_t run_w2d_ensemble(const W2dParams &p, const double2 *gs, double *meas, double *avn, hipStream_t s)
{
    const uint32_t k0 = (uint32_t)p.seed, k1 = (uint32_t)(p.seed >> 32);
    hipLaunchKernelGGL(w2d_ensemble_kernel, dim3((unsigned)p.ne), dim3(THREADS), 0, s, p, gs, meas, avn, k0, k1);
    return hipGetLastError();
}

} // namespace bec
