// Ground-state kernels (imaginary time) built on the pass templates.
//
// Split step: calculateSteadyState of programs/bec_wigner_3d.cuda/wigner.cu:84-175
// with the kernels of wigner_kernel.cu:128-134 (k-space half step),
// :154-181 (x-space midpoint step) and :232-258 (energy integral, quirk Q3).
//
// Gradient flow: psi <- psi - tau P (H - mu) psi + beta (psi - psi_prev), then
// renormalisation to N, with P = (alpha + k^2/2)^-1 applied in k-space.  Its
// fixed points are exactly the solutions of H psi = mu psi on the lattice.
#include "gs_backend.hpp"
#include "launch.hpp"

namespace bec {

namespace {

__device__ __forceinline__ double dv_of(const TrajParams &p) { return p.dx * p.dy * p.dz; }

// Sum of one trajectory's per-workgroup partials, by one wave: lane l adds
// slots l, l + 64, ... and the lanes combine through shuffles (a fixed order,
// so results are reproducible and every wave that computes it gets the same
// bits).  Every lane receives the total.
__device__ __forceinline__ double psum(const double *partial, int per, int nsum, int traj, int k)
{
    const double *p = partial + (long)traj * per * nsum + k;
    double s = 0.0;
    for (int j = (int)(threadIdx.x & 63); j < per; j += 64) s += p[(long)j * nsum];
    return group_sum<64>(s);
}

// --------------------------------------------------------------- z-passes

template <typename T, int NZ, int R1> struct OpGsZImag : SkipDone {
    static constexpr int NC = 1;
    static constexpr int NW = NZ / R1, KPT = R1 / NW;
    cpx<T> *psi;
    const cpx<T> *kxy, *kz;   // KT_IM_HALF tables [records][...]
    const double *scale;      // per trajectory
    cpx<T> *kcopy;
    int nxy, rec_shift;
    __device__ const cpx<T> *src(int) const { return psi; }
    __device__ cpx<T> *dst(int) const { return psi; }
    __device__ __forceinline__ void kspace(int, cpx<T> (&X)[KPT][NW], int kk0, int q, int traj) const
    {
        const int rec = traj >> rec_shift;
        cpx<T> fxy = kxy[(long)rec * nxy + q];
        const T s = (T)scale[traj];
        fxy = mk<T>(fxy.x * s, fxy.y * s);
        const cpx<T> *kzr = kz + (long)rec * NZ;
#pragma unroll
        for (int i = 0; i < KPT; ++i)
#pragma unroll
            for (int m = 0; m < NW; ++m) {
                const int kzi = kk0 + i + R1 * m;
                X[i][m] = cmul<T>(X[i][m], cmul<T>(fxy, kzr[kzi]));
                if (kcopy) kcopy[((long)traj * NZ + kzi) * nxy + q] = X[i][m];
            }
    }
};

// Fused residual test (ph != nullptr): every workgroup evaluates its
// trajectory's residual from the partials of the Hamiltonian and residual
// passes (the same fixed-order sums as scalar_residual_kernel), stops on
// convergence or a non-finite value, and the first workgroup records it.
struct FusedResidual {
    const double *ph = nullptr, *pr = nullptr;   // partials {.., <psi,psi>, ..}, {|r|^2}
    int per = 0, check = 0, pair = 0;
    double tol = 0.0;
    GsScalars sc{};
    bool stop = false;
    __device__ __forceinline__ void run(int traj, bool first, const int *done)
    {
        if (!ph || (done && done[traj])) return;
        const double mu = sc.mu[traj];
        double r = sqrt(psum(pr, per, 1, traj, 0) / psum(ph, per, 5, traj, 1)) / fabs(mu);
        if (pair) {
            const int q = traj ^ 1;
            const double rq = sqrt(psum(pr, per, 1, q, 0) / psum(ph, per, 5, q, 1)) / fabs(sc.mu[q]);
            r = (r == r && rq == rq) ? fmax(r, rq) : rq + r;
        }
        const int flag = !(r == r) ? 2 : (check && r < tol) ? 1 : 0;
        stop = flag != 0;
        if (first && threadIdx.x == 0) {
            sc.res[traj] = r;
            if (flag) sc.done[traj] = flag;
        }
    }
};

template <typename T, int NZ, int R1> struct OpGsZDiag : SkipDone {
    static constexpr int NC = 1;
    static constexpr int NW = NZ / R1, KPT = R1 / NW;
    cpx<T> *w;
    const double *exy, *ez;
    const double *alpha_rec;   // per-record preconditioner shift (nullptr: alpha)
    int nxy, rec_shift, variant;
    double alpha;
    FusedResidual fr;
    __device__ __forceinline__ void begin(int traj, bool first) { fr.run(traj, first, done); }
    __device__ __forceinline__ bool skip(int traj) const { return (done && done[traj]) || fr.stop; }
    __device__ const cpx<T> *src(int) const { return w; }
    __device__ cpx<T> *dst(int) const { return w; }
    __device__ __forceinline__ void kspace(int, cpx<T> (&X)[KPT][NW], int kk0, int q, int traj) const
    {
        const int rec = traj >> rec_shift;
        const double e0 = exy[(long)rec * nxy + q];
        const double *ezr = ez + (long)rec * NZ;
        const double al = alpha_rec ? alpha_rec[rec] : alpha;
#pragma unroll
        for (int i = 0; i < KPT; ++i)
#pragma unroll
            for (int m = 0; m < NW; ++m) {
                const double e = e0 + ezr[kk0 + i + R1 * m];
                const T f = (T)(variant == 0 ? e / NZ : 1.0 / ((al + e) * NZ));
                X[i][m] = cscale<T>(X[i][m], f);
            }
    }
};

// -------------------------------------------------------------- xy-passes

// propagateXSpaceOneComponent (wigner_kernel.cu:154-181) + 1/(NX NY)
template <typename T, int N> struct OpGsNonlinIm : SkipDone {
    static constexpr int NC = 1, SIGN_IN = -1, SIGN_OUT = +1, NSUM = 0;
    static constexpr bool WRITE = true;
    cpx<T> *psi;
    const TrajParams *P;
    int rec_shift;
    __device__ const cpx<T> *src(int) const { return psi; }
    __device__ cpx<T> *dst(int) const { return psi; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[1][N], int y, int z, int traj, long, double *) const
    {
        const TrajParams &p = P[traj >> rec_shift];
        const T hdt = (T)(0.5 * p.dt_steady), g11 = (T)p.g11, scale = (T)(1.0 / (N * N));
        const int itmax = p.itmax;
        static_for<0, N>([&](auto X_) {
            constexpr int x = decltype(X_)::value;
            const T V = potential_t<T>(p, x, y, z);
            const cpx<T> a0 = r[0][x];
            cpx<T> a = a0;
            T da = (T)0;
            for (int it = 0; it < itmax; ++it) {
                if constexpr (sizeof(T) == 4) da = expf(hdt * (-V - g11 * cnorm<T>(a)));
                else da = exp(hdt * (-V - g11 * cnorm<T>(a)));
                a = cscale<T>(a0, da);
            }
            r[0][x] = cscale<T>(a, da * scale);
        });
    }
};

// Reference energy integral pieces on the x-space state (before the pending
// renormalisation): {sum n, sum n V, sum n^2, sum Re(psi_x[i] psi_k[i] k_i^2/2)}
template <typename T, int N> struct OpGsMeasureRef : SkipDone {
    static constexpr int NC = 1, SIGN_IN = -1, SIGN_OUT = 0, NSUM = 4;
    static constexpr bool WRITE = false;
    const cpx<T> *psi, *kcopy;
    const TrajParams *P;
    const double *exy, *ez;
    int rec_shift, nz;
    __device__ const cpx<T> *src(int) const { return psi; }
    __device__ cpx<T> *dst(int) const { return nullptr; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[1][N], int y, int z, int traj, long row, double *s) const
    {
        const int rec = traj >> rec_shift;
        const TrajParams &p = P[rec];
        const double ezv = ez[(long)rec * nz + z];
        static_for<0, N>([&](auto X_) {
            constexpr int x = decltype(X_)::value;
            const double ar = r[0][x].x, ai = r[0][x].y;
            const double n = ar * ar + ai * ai;
            const cpx<T> k = kcopy[row + x];
            const double e = exy[(long)rec * N * N + y * N + x] + ezv;
            s[0] += n;
            s[1] += n * potential_at(p, x, y, z);
            s[2] += n * n;
            s[3] += (ar * (double)k.x - ai * (double)k.y) * e;
        });
    }
};

// x-space psi * s / (NX NY) -> forward xy transform (x -> hybrid)
// Fused renormalisation (pn != nullptr): the scale of the previous update,
// s = sqrt(N_target / (dV sum |psi|^2)), from its partials (the sums of
// scalar_norm_kernel); the first workgroup records s and counts the iteration.
template <typename T, int N> struct OpGsForward : SkipDone {
    static constexpr int NC = 1, SIGN_IN = 0, SIGN_OUT = +1, NSUM = 0;
    static constexpr bool WRITE = true;
    const cpx<T> *psi;
    cpx<T> *w;
    const double *scale;
    const double *pn = nullptr;
    int per = 0, rec_shift = 0;
    const TrajParams *P = nullptr;
    GsScalars sc{};
    double s_v = 1.0;
    bool bad = false;
    __device__ __forceinline__ void begin(int traj, bool first)
    {
        if (done && done[traj]) return;
        if (!pn) {
            s_v = scale[traj];
            return;
        }
        const TrajParams &p = P[traj >> rec_shift];
        const double n = psum(pn, per, 1, traj, 0) * dv_of(p);
        s_v = sqrt((sc.target ? sc.target[traj] : p.N0) / n);
        bad = !(n == n);
        if (first && threadIdx.x == 0) {
            sc.s[traj] = s_v;
            sc.iters[traj] += 1;
            if (bad) sc.done[traj] = 2;
        }
    }
    __device__ __forceinline__ bool skip(int traj) const { return (done && done[traj]) || bad; }
    __device__ const cpx<T> *src(int) const { return psi; }
    __device__ cpx<T> *dst(int) const { return w; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[1][N], int, int, int traj, long, double *) const
    {
        const T s = (T)(s_v / (N * N));
#pragma unroll
        for (int x = 0; x < N; ++x) r[0][x] = cscale<T>(r[0][x], s);
    }
};

// h = K psi + (V + g |psi_s|^2) psi_s with psi_s = s psi and g = g11.
// Pair mode (two-component ground state, report.tex:200-211): trajectory
// 2 j + c holds component c of set j; g = g11 (c = 0) or g22 (c = 1) and the
// partner's density adds g12 |psi_partner|^2.
// partials {Re<psi_s,h>, <psi_s,psi_s>, Re<psi_s,K psi_s>, sum V n, sum n^2}
template <typename T, int N> struct OpGsHam : SkipDone {
    static constexpr int NC = 1, SIGN_IN = -1, SIGN_OUT = 0, NSUM = 5;
    static constexpr bool WRITE = true;
    const cpx<T> *psi;
    cpx<T> *w;
    const double *scale;
    const TrajParams *P;
    int rec_shift, pair;
    long cells;
    __device__ const cpx<T> *src(int) const { return w; }
    __device__ cpx<T> *dst(int) const { return w; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[1][N], int y, int z, int traj, long row, double *s) const
    {
        const TrajParams &p = P[traj >> rec_shift];
        const T sc = (T)scale[traj];
        const T g = (T)((pair && (traj & 1)) ? p.g22 : p.g11), g12 = (T)p.g12;
        const T scp = pair ? (T)scale[traj ^ 1] : (T)0;
        const long poff = (traj & 1) ? -cells : cells;
        static_for<0, N>([&](auto X_) {
            constexpr int x = decltype(X_)::value;
            const cpx<T> ps = cscale<T>(psi[row + x], sc);
            const T n = cnorm<T>(ps);
            const T V = potential_t<T>(p, x, y, z);
            T coup = (T)0;
            if (pair) coup = g12 * cnorm<T>(cscale<T>(psi[row + poff + x], scp));
            const cpx<T> kp = r[0][x];
            const cpx<T> h = cadd<T>(kp, cscale<T>(ps, V + g * n + coup));
            r[0][x] = h;
            s[0] += (double)ps.x * h.x + (double)ps.y * h.y;
            s[1] += (double)n;
            s[2] += (double)ps.x * kp.x + (double)ps.y * kp.y;
            s[3] += (double)V * n;
            s[4] += (double)n * n;
        });
    }
};

// r = h - mu psi_s ; partial |r|^2 ; r / (NX NY) -> forward xy transform
// Fused Rayleigh quotient (ph != nullptr): mu = Re<psi,h>/<psi,psi> from the
// Hamiltonian pass's partials (the sums of scalar_mu_kernel); the first
// workgroup records it.
template <typename T, int N> struct OpGsResid : SkipDone {
    static constexpr int NC = 1, SIGN_IN = 0, SIGN_OUT = +1, NSUM = 1;
    static constexpr bool WRITE = true;
    const cpx<T> *psi;
    cpx<T> *w;
    const double *scale, *mu;
    const double *ph = nullptr;
    int per = 0;
    double *mu_out = nullptr;
    double mu_v = 0.0;
    __device__ __forceinline__ void begin(int traj, bool first)
    {
        if (done && done[traj]) return;
        if (!ph) {
            mu_v = mu[traj];
            return;
        }
        mu_v = psum(ph, per, 5, traj, 0) / psum(ph, per, 5, traj, 1);
        if (first && threadIdx.x == 0) mu_out[traj] = mu_v;
    }
    __device__ const cpx<T> *src(int) const { return w; }
    __device__ cpx<T> *dst(int) const { return w; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[1][N], int, int, int traj, long row, double *s) const
    {
        const double sc = scale[traj], m = mu_v;
        const T inv = (T)(1.0 / (N * N));
        static_for<0, N>([&](auto X_) {
            constexpr int x = decltype(X_)::value;
            const cpx<T> p = psi[row + x];
            const double rr = (double)r[0][x].x - m * sc * (double)p.x;
            const double ri = (double)r[0][x].y - m * sc * (double)p.y;
            s[0] += rr * rr + ri * ri;
            r[0][x] = mk<T>((T)rr * inv, (T)ri * inv);
        });
    }
};

// psi_new = s psi - tau d + beta (s psi - prev); prev <- s psi ; partial |psi_new|^2
template <typename T, int N> struct OpGsUpdate : SkipDone {
    static constexpr int NC = 1, SIGN_IN = -1, SIGN_OUT = 0, NSUM = 1;
    static constexpr bool WRITE = false;
    cpx<T> *psi, *prev;
    const cpx<T> *w;
    const double *scale;
    T tau, beta;
    __device__ const cpx<T> *src(int) const { return w; }
    __device__ cpx<T> *dst(int) const { return nullptr; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[1][N], int, int, int traj, long row, double *s) const
    {
        const T sc = (T)scale[traj];
        static_for<0, N>([&](auto X_) {
            constexpr int x = decltype(X_)::value;
            const cpx<T> ps = cscale<T>(psi[row + x], sc);
            const cpx<T> pv = prev[row + x];
            const cpx<T> d = r[0][x];
            const cpx<T> nw = mk<T>(ps.x - tau * d.x + beta * (ps.x - pv.x), ps.y - tau * d.y + beta * (ps.y - pv.y));
            prev[row + x] = ps;
            psi[row + x] = nw;
            s[0] += (double)nw.x * nw.x + (double)nw.y * nw.y;
        });
    }
};

// partial |psi|^2 of an x-space state (no transform)
template <typename T, int N> struct OpGsNorm : SkipDone {
    static constexpr int NC = 1, SIGN_IN = 0, SIGN_OUT = 0, NSUM = 1;
    static constexpr bool WRITE = false;
    const cpx<T> *psi;
    __device__ const cpx<T> *src(int) const { return psi; }
    __device__ cpx<T> *dst(int) const { return nullptr; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[1][N], int, int, int, long, double *s) const
    {
#pragma unroll
        for (int x = 0; x < N; ++x) s[0] += (double)r[0][x].x * r[0][x].x + (double)r[0][x].y * r[0][x].y;
    }
};

// ---------------------------------------------------------- scalar kernels

// The scalar kernels below run one wave per trajectory (psum): their loads
// are issued in parallel instead of one dependent chain per thread.

__global__ void scalar_ref_kernel(const double *partial, int per, int ntraj, const TrajParams *P, int rec_shift,
                                  GsScalars sc, double tol, int use_rule)
{
    const int traj = blockIdx.x;   // one wave per trajectory
    if (traj >= ntraj || sc.done[traj]) return;
    const bool lead = threadIdx.x == 0;
    const TrajParams &p = P[traj >> rec_shift];
    const double dV = dv_of(p);
    const double D = psum(partial, per, 4, traj, 0), A = psum(partial, per, 4, traj, 1);
    const double B = psum(partial, per, 4, traj, 2), C = psum(partial, per, 4, traj, 3);
    const double s2 = p.N0 / (D * dV);
    const double E = dV * (s2 * (A + C) + p.g11 * s2 * s2 * B / 2.0);
    if (!lead) return;
    sc.iters[traj] += 1;
    if (use_rule && fabs((E - sc.e_prev[traj]) / E) < tol) sc.done[traj] = 1;
    if (!(E == E) || !(s2 == s2)) sc.done[traj] = 2;   // non-finite
    sc.e_prev[traj] = E;
    sc.e_ref[traj] = E;
    sc.s[traj] = sqrt(s2);
}

__global__ void scalar_mu_kernel(const double *partial, int per, int ntraj, GsScalars sc)
{
    const int traj = blockIdx.x;   // one wave per trajectory
    if (traj >= ntraj || sc.done[traj]) return;
    const bool lead = threadIdx.x == 0;
    const double mu = psum(partial, per, 5, traj, 0) / psum(partial, per, 5, traj, 1);
    if (lead) sc.mu[traj] = mu;
}

// Pair mode: both components of a set take max(r_a, r_b), so they stop together.
__global__ void scalar_residual_kernel(const double *ph, const double *pr, int per, int ntraj, GsScalars sc,
                                       double tol, int check, int pair)
{
    const int traj = blockIdx.x;   // one wave per trajectory
    if (traj >= ntraj || sc.done[traj]) return;
    const bool lead = threadIdx.x == 0;
    const double nn = psum(ph, per, 5, traj, 1);
    const double rr = psum(pr, per, 1, traj, 0);
    double r = sqrt(rr / nn) / fabs(sc.mu[traj]);
    if (pair) {
        const int q = traj ^ 1;
        const double rq = sqrt(psum(pr, per, 1, q, 0) / psum(ph, per, 5, q, 1)) / fabs(sc.mu[q]);
        r = (r == r && rq == rq) ? fmax(r, rq) : rq + r;   // NaN propagates
    }
    if (!lead) return;
    sc.res[traj] = r;
    if (!(r == r)) sc.done[traj] = 2;
    else if (check && r < tol) sc.done[traj] = 1;
}

__global__ void scalar_norm_kernel(const double *partial, int per, int ntraj, const TrajParams *P, int rec_shift,
                                   GsScalars sc, int count_iter)
{
    const int traj = blockIdx.x;   // one wave per trajectory
    if (traj >= ntraj || sc.done[traj]) return;
    const bool lead = threadIdx.x == 0;
    const TrajParams &p = P[traj >> rec_shift];
    const double n = psum(partial, per, 1, traj, 0) * dv_of(p);
    if (!lead) return;
    sc.s[traj] = sqrt((sc.target ? sc.target[traj] : p.N0) / n);
    if (count_iter) sc.iters[traj] += 1;
    if (!(n == n)) sc.done[traj] = 2;
}

template <typename T>
__global__ void apply_scale_kernel(cpx<T> *psi, long cells, long n, double *s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T f = (T)s[i / cells];
    psi[i] = cscale<T>(psi[i], f);
}

__global__ void reset_scale_kernel(double *s, int ntraj)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntraj) s[i] = 1.0;
}

template <typename T>
__global__ void tf_kernel(cpx<T> *psi, int nx, int ny, int nz, long n, const TrajParams *P, int rec_shift,
                          const double *mu)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long cells = (long)nx * ny * nz;
    const int traj = (int)(i / cells);
    const long c = i % cells;
    const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((long)nx * ny));
    const int rec = traj >> rec_shift;
    const TrajParams &p = P[rec];
    const double e = mu[rec] - potential_at(p, x, y, z);
    psi[i] = mk<T>(e > 0 ? (T)sqrt(e / p.g11) : (T)0, (T)0);
}

} // namespace

// ------------------------------------------------------------- launchers

hipError_t gs_z_imag(const DevGroup &g, void *psi, const GsScalars &sc, void *kcopy, hipStream_t s)
{
    int pw = g.pw ? g.pw : default_pw(g.prec, g.nx, g.nz);
    return dispatch(g, [&](auto Tt, auto, auto NZt) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int NZ = decltype(NZt)::value;
        constexpr int R1 = z_r1<NZ>();
        const int nxy = g.nx * g.ny;
        OpGsZImag<T, NZ, R1> op;
        op.done = sc.done;
        op.psi = (cpx<T> *)psi;
        op.kxy = (const cpx<T> *)g.kxy + (long)KT_IM_HALF * g.n_records * nxy;
        op.kz = (const cpx<T> *)g.kz + (long)KT_IM_HALF * g.n_records * NZ;
        op.scale = sc.s;
        op.kcopy = (cpx<T> *)kcopy;
        op.nxy = nxy;
        op.rec_shift = g.rec_shift;
        return with_pw(pw, [&](auto PWt) { return launch_z<T, NZ, decltype(PWt)::value, +1, true>(g, op, g.ntraj, s); });
    });
}

hipError_t gs_z_diag(const DevGroup &g, void *w, const GsScalars &sc, const GsTables &t, int variant, double alpha,
                     const double *alpha_rec, hipStream_t s, const GsResidualTest *test)
{
    int pw = g.pw ? g.pw : default_pw(g.prec, g.nx, g.nz);
    return dispatch(g, [&](auto Tt, auto, auto NZt) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int NZ = decltype(NZt)::value;
        constexpr int R1 = z_r1<NZ>();
        OpGsZDiag<T, NZ, R1> op;
        op.done = sc.done;
        op.w = (cpx<T> *)w;
        op.exy = t.exy;
        op.ez = t.ez;
        op.nxy = g.nx * g.ny;
        op.rec_shift = g.rec_shift;
        op.variant = variant;
        op.alpha = alpha;
        op.alpha_rec = alpha_rec;
        if (test) {
            op.fr.ph = test->ph;
            op.fr.pr = test->pr;
            op.fr.per = xy_wg_per_traj(g);
            op.fr.check = test->check;
            op.fr.pair = sc.pair;
            op.fr.tol = test->tol;
            op.fr.sc = sc;
        }
        return with_pw(pw, [&](auto PWt) { return launch_z<T, NZ, decltype(PWt)::value, +1, true>(g, op, g.ntraj, s); });
    });
}

hipError_t gs_xy_nonlinear_imag(const DevGroup &g, void *psi, const GsScalars &sc, hipStream_t s)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        OpGsNonlinIm<T, N> op;
        op.done = sc.done;
        op.psi = (cpx<T> *)psi;
        op.P = g.params;
        op.rec_shift = g.rec_shift;
        return launch_xy<T, N>(g, op, g.ntraj, nullptr, s);
    });
}

hipError_t gs_xy_measure_ref(const DevGroup &g, const void *psi, const void *kcopy, const GsScalars &sc,
                             const GsTables &t, double *partial, hipStream_t s)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        OpGsMeasureRef<T, N> op;
        op.done = sc.done;
        op.psi = (const cpx<T> *)psi;
        op.kcopy = (const cpx<T> *)kcopy;
        op.P = g.params;
        op.exy = t.exy;
        op.ez = t.ez;
        op.rec_shift = g.rec_shift;
        op.nz = g.nz;
        return launch_xy<T, N>(g, op, g.ntraj, partial, s);
    });
}

hipError_t gs_scalar_ref(const DevGroup &g, const double *partial, const GsScalars &sc, double tol, int use_rule,
                         hipStream_t s)
{
    hipLaunchKernelGGL(scalar_ref_kernel, dim3(g.ntraj), dim3(64), 0, s, partial, xy_wg_per_traj(g),
                       g.ntraj, g.params, g.rec_shift, sc, tol, use_rule);
    return hipGetLastError();
}

hipError_t gs_xy_forward(const DevGroup &g, const void *psi, void *w, const GsScalars &sc, hipStream_t s,
                         const double *pn)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        OpGsForward<T, N> op;
        op.done = sc.done;
        op.psi = (const cpx<T> *)psi;
        op.w = (cpx<T> *)w;
        op.scale = sc.s;
        op.pn = pn;
        op.per = xy_wg_per_traj(g);
        op.P = g.params;
        op.rec_shift = g.rec_shift;
        op.sc = sc;
        return launch_xy<T, N>(g, op, g.ntraj, nullptr, s);
    });
}

hipError_t gs_xy_hamiltonian(const DevGroup &g, const void *psi, void *w, const GsScalars &sc, double *partial,
                             hipStream_t s)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        OpGsHam<T, N> op;
        op.done = sc.done;
        op.psi = (const cpx<T> *)psi;
        op.w = (cpx<T> *)w;
        op.scale = sc.s;
        op.P = g.params;
        op.rec_shift = g.rec_shift;
        op.pair = sc.pair;
        op.cells = (long)g.nx * g.ny * g.nz;
        return launch_xy<T, N>(g, op, g.ntraj, partial, s);
    });
}

hipError_t gs_scalar_mu(const DevGroup &g, const double *partial, const GsScalars &sc, hipStream_t s)
{
    hipLaunchKernelGGL(scalar_mu_kernel, dim3(g.ntraj), dim3(64), 0, s, partial, xy_wg_per_traj(g),
                       g.ntraj, sc);
    return hipGetLastError();
}

hipError_t gs_xy_residual(const DevGroup &g, const void *psi, void *w, const GsScalars &sc, double *partial,
                          hipStream_t s, const double *ph)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        OpGsResid<T, N> op;
        op.done = sc.done;
        op.psi = (const cpx<T> *)psi;
        op.w = (cpx<T> *)w;
        op.scale = sc.s;
        op.mu = sc.mu;
        op.ph = ph;
        op.per = xy_wg_per_traj(g);
        op.mu_out = sc.mu;
        return launch_xy<T, N>(g, op, g.ntraj, partial, s);
    });
}

hipError_t gs_scalar_residual(const DevGroup &g, const double *partial_h, const double *partial_r,
                              const GsScalars &sc, double tol, int check, hipStream_t s)
{
    hipLaunchKernelGGL(scalar_residual_kernel, dim3(g.ntraj), dim3(64), 0, s, partial_h, partial_r,
                       xy_wg_per_traj(g), g.ntraj, sc, tol, check, sc.pair);
    return hipGetLastError();
}

hipError_t gs_xy_update(const DevGroup &g, void *psi, void *prev, const void *w, const GsScalars &sc, double tau,
                        double beta, double *partial, hipStream_t s)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        OpGsUpdate<T, N> op;
        op.done = sc.done;
        op.psi = (cpx<T> *)psi;
        op.prev = (cpx<T> *)prev;
        op.w = (const cpx<T> *)w;
        op.scale = sc.s;
        op.tau = (T)tau;
        op.beta = (T)beta;
        return launch_xy<T, N>(g, op, g.ntraj, partial, s);
    });
}

hipError_t gs_scalar_norm(const DevGroup &g, const double *partial, const GsScalars &sc, hipStream_t s)
{
    hipLaunchKernelGGL(scalar_norm_kernel, dim3(g.ntraj), dim3(64), 0, s, partial, xy_wg_per_traj(g),
                       g.ntraj, g.params, g.rec_shift, sc, 1);
    return hipGetLastError();
}

hipError_t gs_norm_only(const DevGroup &g, const void *psi, const GsScalars &sc, double *partial, hipStream_t s)
{
    hipError_t e = dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        OpGsNorm<T, N> op;
        op.done = sc.done;
        op.psi = (const cpx<T> *)psi;
        return launch_xy<T, N>(g, op, g.ntraj, partial, s);
    });
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scalar_norm_kernel, dim3(g.ntraj), dim3(64), 0, s, partial, xy_wg_per_traj(g),
                       g.ntraj, g.params, g.rec_shift, sc, 0);
    return hipGetLastError();
}

hipError_t gs_apply_scale(const DevGroup &g, void *psi, const GsScalars &sc, hipStream_t s)
{
    const long cells = (long)g.nx * g.ny * g.nz, n = cells * g.ntraj;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (g.prec == 0)
        hipLaunchKernelGGL(apply_scale_kernel<float>, grid, block, 0, s, (float2 *)psi, cells, n, sc.s);
    else
        hipLaunchKernelGGL(apply_scale_kernel<double>, grid, block, 0, s, (double2 *)psi, cells, n, sc.s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(reset_scale_kernel, dim3((g.ntraj + 63) / 64), dim3(64), 0, s, sc.s, g.ntraj);
    return hipGetLastError();
}

template <typename T>
__global__ void net_start_kernel(const float *__restrict__ net, cpx<T> *__restrict__ psi, long cells, long n,
                                 const double *__restrict__ unit, double w, double r)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;   // i = set * cells + cell
    if (i >= n) return;
    const long j = i / cells, cell = i - j * cells;
    const double sv = (double)net[2 * i], dv = (double)net[2 * i + 1] / w;
    const double u = unit[j];
    const double a = ((sv + dv) * r) * u, b = ((sv - dv) * r) * u;
    psi[(2 * j) * cells + cell] = mk<T>((T)a, (T)0);
    psi[(2 * j + 1) * cells + cell] = mk<T>((T)b, (T)0);
}

hipError_t gs_net_start(const DevGroup &g, void *psi, const float *net, const double *unit, double spin_weight,
                        double r, hipStream_t s)
{
    const long cells = (long)g.nx * g.ny * g.nz;
    const long n = cells * (g.ntraj / 2);
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (g.prec == 0)
        hipLaunchKernelGGL(net_start_kernel<float>, grid, block, 0, s, net, (cpx<float> *)psi, cells, n, unit,
                           spin_weight, r);
    else
        hipLaunchKernelGGL(net_start_kernel<double>, grid, block, 0, s, net, (cpx<double> *)psi, cells, n, unit,
                           spin_weight, r);
    return hipGetLastError();
}

hipError_t gs_tf_start(const DevGroup &g, void *psi, const double *mu_tf, hipStream_t s)
{
    const long n = (long)g.nx * g.ny * g.nz * g.ntraj;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (g.prec == 0)
        hipLaunchKernelGGL(tf_kernel<float>, grid, block, 0, s, (float2 *)psi, g.nx, g.ny, g.nz, n, g.params,
                           g.rec_shift, mu_tf);
    else
        hipLaunchKernelGGL(tf_kernel<double>, grid, block, 0, s, (double2 *)psi, g.nx, g.ny, g.nz, n, g.params,
                           g.rec_shift, mu_tf);
    return hipGetLastError();
}

} // namespace bec
