// The two streaming passes of one split-step, as kernel templates.
//
// State layout in HBM: two arrays a, b of interleaved complex, index
// traj*cells + z*(NX*NY) + y*NX + x.  Between steps the state is held in a
// "hybrid" representation H(z, ky, kx): transformed to k-space along x and y
// only.  One real-time step is then exactly two passes over HBM:
//
//   z-pass : H --F_z(+)--> psi_k --* exp(i c dt k^2/2)/NZ--> --F_z(-)--> H
//   xy-pass: H --F_xy(-)--> psi_x --nonlinear/loss--> --F_xy(+)/(NX NY)--> H
//
// which is propagate() of programs/bec_wigner_3d.cuda/wigner.cu:178-199 with
// the trailing kinetic half-step of one step merged with the leading one of
// the next (c = 1) and the 1/cells normalisation of wigner.cu:193 folded in.
#pragma once

#include "common.hpp"
#include "philox.hpp"
#include "fastmath.hpp"
#include "fft_reg.hpp"

namespace bec {

// ------------------------------------------------------------ memory access
// NT = true issues non-temporal (streaming) loads and stores: the state is
// touched once per pass, so it should not displace reusable lines.
typedef float v2f_t __attribute__((ext_vector_type(2)));
typedef double v2d_t __attribute__((ext_vector_type(2)));

template <bool NT, typename T> __device__ __forceinline__ cpx<T> ld(const cpx<T> *p)
{
    if constexpr (NT) {
        if constexpr (sizeof(T) == 4) {
            const v2f_t r = __builtin_nontemporal_load((const v2f_t *)p);
            return mk<T>(r.x, r.y);
        } else {
            const v2d_t r = __builtin_nontemporal_load((const v2d_t *)p);
            return mk<T>(r.x, r.y);
        }
    } else {
        return *p;
    }
}

template <bool NT, typename T> __device__ __forceinline__ void st(cpx<T> *p, cpx<T> v)
{
    if constexpr (NT) {
        if constexpr (sizeof(T) == 4) {
            v2f_t r = {v.x, v.y};
            __builtin_nontemporal_store(r, (v2f_t *)p);
        } else {
            v2d_t r = {v.x, v.y};
            __builtin_nontemporal_store(r, (v2d_t *)p);
        }
    } else {
        *p = v;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8).  With
// the remap every tile of trajectory `traj` runs on XCD traj % 8 in both
// passes, so a trajectory's lines stay in one XCD's L2 between passes.
// Requires ntraj % 8 == 0; tiles_per_traj tiles per trajectory.
__device__ __forceinline__ long xcd_remap(long b, int tiles_per_traj)
{
    const long xcd = b & 7, slot = b >> 3;
    const long tl = slot / tiles_per_traj, w = slot - tl * tiles_per_traj;
    return (tl * 8 + xcd) * tiles_per_traj + w;
}

// ============================================================ xy-plane pass
// One wavefront (64 lanes) holds 64/N planes of N x N cells.  Lane t of a
// plane first owns column x = t (coalesced loads), transforms along y in
// registers, transposes through LDS to own row y = t, transforms along x,
// applies the pointwise operator, and undoes the path.

template <typename T, int N>
__device__ __forceinline__ void col2row(cpx<T> *tile, int t, cpx<T> (&v)[N])
{
#pragma unroll
    for (int y = 0; y < N; ++y) tile[y * (N + 1) + t] = v[y];
    __syncthreads();
#pragma unroll
    for (int x = 0; x < N; ++x) v[x] = tile[t * (N + 1) + x];
    __syncthreads();
}

template <typename T, int N>
__device__ __forceinline__ void row2col(cpx<T> *tile, int t, cpx<T> (&v)[N])
{
#pragma unroll
    for (int x = 0; x < N; ++x) tile[t * (N + 1) + x] = v[x];
    __syncthreads();
#pragma unroll
    for (int y = 0; y < N; ++y) v[y] = tile[y * (N + 1) + t];
    __syncthreads();
}

// A copy of v that the compiler cannot see through: values derived from it
// are recomputed where they are used rather than hoisted out of a loop.
__device__ __forceinline__ int opaque_index(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

struct XyGeom {
    int nz;          // planes per trajectory
    long n_planes;   // planes in the launch
    int xcd;         // remap workgroups so a trajectory stays on one XCD
};

// Op interface:
//   static constexpr int NC;         components carried through the pass
//   static constexpr int SIGN_IN;    transform on the way in (0: none)
//   static constexpr int SIGN_OUT;   transform on the way out (0: none)
//   static constexpr bool WRITE;     write the components back
//   static constexpr int NSUM;       per-WG FP64 partial sums produced
//   const cpx<T>* src(int c); cpx<T>* dst(int c);
//   void apply(cpx<T> (&rows)[NC][N], int y, int z, int traj, long row_index, double* sums)
// MINW: minimum waves per SIMD the register allocation must allow.
// One tile (the 64/N planes blk*PPW ..) of the xy pass, by one 64-lane
// workgroup, exchanging through `lds` (PPW * N * (N + 1) complex values).
// lane is threadIdx.x; a caller that loops over tiles passes it through an
// opaque copy (opaque_index), so that the per-lane LDS addresses stay inside
// its loop instead of being hoisted out and held in registers.
template <typename T, int N, class Op, bool NT>
__device__ __forceinline__ void xy_tile(const Op &op, const XyGeom &g, double *partial, long blk, cpx<T> *lds,
                                        int lane);

template <typename T, int N, class Op, bool NT = false, int MINW = 1>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MINW, 8)))
xy_kernel(Op op, XyGeom g, double *partial)
{
    constexpr int PPW = 64 / N;
    __shared__ cpx<T> lds[PPW * N * (N + 1)];
    const long blk = g.xcd ? xcd_remap(blockIdx.x, g.nz / PPW) : (long)blockIdx.x;
    xy_tile<T, N, Op, NT>(op, g, partial, blk, lds, threadIdx.x);
}

template <typename T, int N, class Op, bool NT>
__device__ __forceinline__ void xy_tile(const Op &op_, const XyGeom &g, double *partial, long blk, cpx<T> *lds,
                                        int lane)
{
    constexpr int PPW = 64 / N;
    constexpr int NC = Op::NC;
    Op op = op_;
    const int pl = lane / N, t = lane % N;
    const long plane = blk * PPW + pl;
    const int z = (int)(plane % g.nz);
    // PPW divides nz, so all planes of the tile belong to one trajectory; taken
    // from blk alone it is uniform, and its parameters load into scalar registers
    const int traj = (int)((blk * PPW) / g.nz);
    op.begin(traj, blk % (g.nz / PPW) == 0);
    if (op.skip(traj)) return;   // uniform: a workgroup never spans two trajectories
    cpx<T> *tile = lds + pl * N * (N + 1);
    const long base = plane * (long)(N * N);

    cpx<T> v[NC][N];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const cpx<T> *s = op.src(c);
#pragma unroll
        for (int y = 0; y < N; ++y) v[c][y] = ld<NT, T>(s + base + y * N + t);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if constexpr (Op::SIGN_IN != 0) dft<T, N, Op::SIGN_IN>(v[c]);
        col2row<T, N>(tile, t, v[c]);
        if constexpr (Op::SIGN_IN != 0) dft<T, N, Op::SIGN_IN>(v[c]);
    }

    double sums[Op::NSUM > 0 ? Op::NSUM : 1];
#pragma unroll
    for (int i = 0; i < (Op::NSUM > 0 ? Op::NSUM : 1); ++i) sums[i] = 0.0;
    op.apply(v, t, z, traj, base + (long)t * N, sums);

    if constexpr (Op::NSUM > 0) {
#pragma unroll
        for (int i = 0; i < Op::NSUM; ++i) {
            double s = group_sum<64>(sums[i]);
            if (lane == 0) partial[blk * Op::NSUM + i] = s;
        }
    }

    if constexpr (Op::WRITE) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if constexpr (Op::SIGN_OUT != 0) dft<T, N, Op::SIGN_OUT>(v[c]);
            row2col<T, N>(tile, t, v[c]);
            if constexpr (Op::SIGN_OUT != 0) dft<T, N, Op::SIGN_OUT>(v[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            cpx<T> *d = op.dst(c);
#pragma unroll
            for (int y = 0; y < N; ++y) st<NT, T>(d + base + y * N + t, v[c][y]);
        }
    }
}

// ------------------------------------------------------------ xy operators

// Operators that never skip a trajectory.  begin(traj, first) runs once per
// workgroup, by every thread, before skip(): operators may compute per-
// trajectory scalars there (first: the workgroup is its trajectory's first).
struct NoSkip {
    __device__ __forceinline__ void begin(int, bool) {}
    __device__ __forceinline__ bool skip(int) const { return false; }
};
// Operators that skip trajectories whose done flag is set.
struct SkipDone {
    const int *done;
    __device__ __forceinline__ void begin(int, bool) {}
    __device__ __forceinline__ bool skip(int traj) const { return done && done[traj]; }
};

// Fused two-component nonlinear + loss + detuning step in x-space
// (propagateXSpaceTwoComponent, wigner_kernel.cu:184-229), followed by the
// 1/(NX NY) share of the x->k normalisation.  IT1 = true is the straight-line
// form for itmax = 1 (the reference default, wigner.cpp:90): densities from
// the input, d = exp(p dt/2), output psi0 d d.  IT1 = false runs the general
// midpoint iteration with a runtime trip count.
//
// NOISE adds the loss-induced Wigner noise that the reference derives
// (docs/wigner_simulation.report/report.tex:624-656) but does not integrate:
// the Ito increment B dW of the leading-order loss diffusion, evaluated at the
// input state, with the Philox stream {cell, traj, step, 1 + group}.
template <typename T, int N, bool IT1 = true, bool NOISE = false> struct OpNonlinRT : NoSkip {
    static constexpr int NC = 2, SIGN_IN = -1, SIGN_OUT = +1, NSUM = 0;
    static constexpr bool WRITE = true;
    cpx<T> *a, *b;
    const TrajParams *P;
    int rec_shift;
    const unsigned *step;          // NOISE: the step counter
    uint32_t k0, k1, group;        // NOISE: Philox key, group index
    __device__ const cpx<T> *src(int c) const { return c == 0 ? a : b; }
    __device__ cpx<T> *dst(int c) const { return c == 0 ? a : b; }
    // B dW for one cell: D11 = 3/2 l111 n1^2 + 1/2 l12 n2, D22 = 1/2 l12 n1 + l22 n2,
    // D12 = 1/2 l12 a0 conj(b0), <d psi_i d psi_j*> = D_ij dt/dV; B is the lower
    // Cholesky factor and dW_k = (z_2k + i z_2k+1) sqrt(dt / (2 dV)), times `scale`.
    static __device__ __forceinline__ void loss_increment(T l111, T l12, T l22, cpx<T> a0, cpx<T> b0, const T z[4],
                                                          T ws, cpx<T> &da, cpx<T> &db)
    {
        const T n1 = cnorm<T>(a0), n2 = cnorm<T>(b0);
        const T D11 = (T)1.5 * l111 * n1 * n1 + (T)0.5 * l12 * n2;
        const T D22 = (T)0.5 * l12 * n1 + l22 * n2;
        const cpx<T> w1 = mk<T>(z[0] * ws, z[1] * ws), w2 = mk<T>(z[2] * ws, z[3] * ws);
        const T b11 = sqrt_t(D11);
        cpx<T> b21 = mk<T>(0, 0);
        T b22;
        if (D11 > (T)0) {
            const cpx<T> cab = mk<T>(a0.x * b0.x + a0.y * b0.y, a0.x * b0.y - a0.y * b0.x);   // conj(a0) b0
            b21 = cscale<T>(cab, (T)0.5 * l12 / b11);
            const T r = D22 - (T)0.25 * l12 * l12 * n1 * n2 / D11;
            b22 = sqrt_t(r > (T)0 ? r : (T)0);
        } else {
            b22 = sqrt_t(D22);
        }
        da = cscale<T>(w1, b11);
        db = mk<T>(b21.x * w1.x - b21.y * w1.y + b22 * w2.x, b21.x * w1.y + b21.y * w1.x + b22 * w2.y);
    }
    __device__ __forceinline__ void apply(cpx<T> (&r)[2][N], int y, int z, int traj, long, double *) const
    {
        const TrajParams &p = P[traj >> rec_shift];
        const T half_dt = (T)(0.5 * p.dt);
        const T g11 = (T)p.g11, g12 = (T)p.g12, g22 = (T)p.g22;
        const T l111 = (T)p.l111, l12 = (T)p.l12, l22 = (T)p.l22, det = (T)p.detuning;
        const T scale = (T)(1.0 / (N * N));
        const int itmax = p.itmax;
        const T yy = (T)(-p.ymax) + (T)p.dy * (T)y;
        const T zz = (T)(-p.zmax) + (T)p.dz * (T)z;
        const T vyz = yy * yy + zz * zz * (T)p.inv_lambda2;
        // the loss noise of this row: the step, dW scaled by the x->k share 1/(NX NY)
        uint32_t st = 0;
        T ws = (T)0;
        if constexpr (NOISE) {
            st = *step;
            ws = (T)(sqrt(p.dt / (2.0 * p.dx * p.dy * p.dz)) / (N * N));
        }
        const uint32_t cell0 = (uint32_t)((z * N + y) * N);
        auto add_noise = [&](int x, cpx<T> a0, cpx<T> b0) {
            if constexpr (NOISE) {
                T zn[4];
                loss_noise4<T>(cell0 + (uint32_t)x, (uint32_t)traj, st, group, k0, k1, zn);
                cpx<T> da, db;
                loss_increment(l111, l12, l22, a0, b0, zn, ws, da, db);
                r[0][x] = mk<T>(r[0][x].x + da.x, r[0][x].y + da.y);
                r[1][x] = mk<T>(r[1][x].x + db.x, r[1][x].y + db.y);
            }
        };
        if constexpr (IT1) {
            // itmax = 1: d = exp(p(psi0) dt/2) and the output psi0 d^2, i.e. the
            // phase 2 phi = Im(p) dt and the loss exponent y = Re(p) dt, with dt
            // folded into the (uniform) coefficients
            const double dt = p.dt;
            const T g11d = (T)(p.g11 * dt), g12d = (T)(p.g12 * dt), g22d = (T)(p.g22 * dt);
            const T n111 = (T)(-0.5 * p.l111 * dt), n12 = (T)(-0.5 * p.l12 * dt), n22 = (T)(-0.5 * p.l22 * dt);
            const T detd = (T)(p.detuning * dt);
            static_for<0, N>([&](auto X_) {
                constexpr int x = decltype(X_)::value;
                const T xx = (T)(-p.xmax) + (T)p.dx * (T)x;
                const T Vd = half_dt * fma_t(xx, xx, vyz);     // V dt
                const cpx<T> a0 = r[0][x], b0 = r[1][x];
                const T na = cnorm<T>(a0), nb = cnorm<T>(b0);
                const T pha = fma_t(g11d, na, fma_t(g12d, nb, Vd));
                const T phb = fma_t(g22d, nb, fma_t(g12d, na, Vd - detd));
                const T ya = fma_t(n111, na * na, n12 * nb);
                const T yb = fma_t(n22, nb, n12 * na);
                T sa, ca, sb, cb;
                sincos2_scaled_t(pha, scale, &sa, &ca);
                sincos2_scaled_t(phb, scale, &sb, &cb);
                const T ta = expm1_small_t(ya), tb = expm1_small_t(yb);
                const cpx<T> va = cmul<T>(a0, mk<T>(ca, sa));
                const cpx<T> vb = cmul<T>(b0, mk<T>(cb, sb));
                r[0][x] = mk<T>(fma_t(va.x, ta, va.x), fma_t(va.y, ta, va.y));
                r[1][x] = mk<T>(fma_t(vb.x, tb, vb.x), fma_t(vb.y, tb, vb.y));
                add_noise(x, a0, b0);
            });
            return;
        }
        // itmax = 0 never forms the factor: the output is zero, as the reference's
        // uninitialised factor (quirk Q10) and the oracle give
        const T scale_out = itmax > 0 ? scale : (T)0;
        static_for<0, N>([&](auto X_) {
            constexpr int x = decltype(X_)::value;
            const T xx = (T)(-p.xmax) + (T)p.dx * (T)x;
            const T V = (T)0.5 * (xx * xx + vyz);
            const cpx<T> a0 = r[0][x], b0 = r[1][x];
            // d = m e^{i phi}; the output psi0 d^2 = psi0 e^{2 i phi} (1 + t) with
            // t = m^2 - 1 = expm1(2 Re(p) dt/2) formed directly, so that the small
            // loss increments (|t| ~ 1e-6 per step) keep full relative precision
            cpx<T> ea = mk<T>(1, 0), eb = mk<T>(1, 0);
            T ta = (T)0, tb = (T)0;
            auto factors = [&](cpx<T> ua, cpx<T> ub) {
                const T na = cnorm<T>(ua), nb = cnorm<T>(ub);
                const T pa_re = -(l111 * na * na + l12 * nb) * (T)0.5;
                const T pa_im = V + g11 * na + g12 * nb;
                const T pb_re = -(l22 * nb + l12 * na) * (T)0.5;
                const T pb_im = V + g22 * nb + g12 * na - det;
                T sa, cs, sb, cb2;
                sincos_t(pa_im * half_dt, &sa, &cs);
                sincos_t(pb_im * half_dt, &sb, &cb2);
                ea = mk<T>(cs, sa);
                eb = mk<T>(cb2, sb);
                ta = expm1_t((T)2 * pa_re * half_dt);
                tb = expm1_t((T)2 * pb_re * half_dt);
            };
            for (int it = 0; it < itmax; ++it) {
                // midpoint: psi0 d = psi0 e^{i phi} sqrt(1 + t)
                factors(cscale<T>(cmul<T>(a0, ea), sqrt_t((T)1 + ta)),
                        cscale<T>(cmul<T>(b0, eb), sqrt_t((T)1 + tb)));
            }
            const cpx<T> va = cscale<T>(cmul<T>(a0, cmul<T>(ea, ea)), scale_out);
            const cpx<T> vb = cscale<T>(cmul<T>(b0, cmul<T>(eb, eb)), scale_out);
            r[0][x] = mk<T>(fma_t(va.x, ta, va.x), fma_t(va.y, ta, va.y));
            r[1][x] = mk<T>(fma_t(vb.x, tb, vb.x), fma_t(vb.y, tb, vb.y));
            add_noise(x, a0, b0);
        });
    }
    // The IT1 step of one component c (0: a, 1: b) of a row, given the other
    // component's densities m (xy_split_kernel): the same operations, in the
    // same order, as apply() performs for that component.
    __device__ __forceinline__ void apply_one(cpx<T> (&r)[N], const T (&m)[N], int comp, int y, int z,
                                              int traj) const
    {
        static_assert(IT1 && !NOISE, "the split pass carries the straight-line noiseless step");
        const TrajParams &p = P[traj >> rec_shift];
        const double dt = p.dt;
        const T half_dt = (T)(0.5 * dt);
        const T scale = (T)(1.0 / (N * N));
        const T yy = (T)(-p.ymax) + (T)p.dy * (T)y;
        const T zz = (T)(-p.zmax) + (T)p.dz * (T)z;
        const T vyz = yy * yy + zz * zz * (T)p.inv_lambda2;
        const T gsd = (T)((comp ? p.g22 : p.g11) * dt), g12d = (T)(p.g12 * dt);
        const T nsd = (T)(-0.5 * (comp ? p.l22 : p.l111) * dt), n12 = (T)(-0.5 * p.l12 * dt);
        const T shift = comp ? (T)(p.detuning * dt) : (T)0;
        static_for<0, N>([&](auto X_) {
            constexpr int x = decltype(X_)::value;
            const T xx = (T)(-p.xmax) + (T)p.dx * (T)x;
            const T Vd = half_dt * fma_t(xx, xx, vyz);
            const cpx<T> a0 = r[x];
            const T n = cnorm<T>(a0), o = m[x];
            const T ph = fma_t(gsd, n, fma_t(g12d, o, Vd - shift));
            const T yl = fma_t(nsd, comp ? n : n * n, n12 * o);
            T sn, cs;
            sincos2_scaled_t(ph, scale, &sn, &cs);
            const T tl = expm1_small_t(yl);
            const cpx<T> va = cmul<T>(a0, mk<T>(cs, sn));
            r[x] = mk<T>(fma_t(va.x, tl, va.x), fma_t(va.y, tl, va.y));
        });
    }
    // S sin x and S cos x of the doubled phase (IT1 path), S a power of two (the
    // 1/N^2 of the xy transform pair): exactly the rounded sin and cos times S
    static __device__ __forceinline__ void sincos2_scaled_t(T x, T S, T *s, T *c)
    {
        if constexpr (sizeof(T) == 4) {
            fast_sincos_pi_scaled(x, S, s, c);
        } else {
            sincos(x, s, c);
            *s *= S;
            *c *= S;
        }
    }
    // expm1 for the IT1 path: a cubic (relative truncation error y^3/24 <
    // 2.5e-9 for |y| < 2^-8); only when a lane of the wave has |y| >= 2^-8 (a
    // loss of > 0.4% in one step) does the wave take exp(y) - 1.  The loss
    // exponents of the experiments' parameters are ~1e-4 per step.
    static __device__ __forceinline__ T expm1_small_t(T y)
    {
        if constexpr (sizeof(T) == 4) {
            float p = fmaf(y, 1.0f / 6.0f, 0.5f);
            p = fmaf(y, p, 1.0f);
            float t = y * p;
            const bool big = fabsf(y) >= 0.00390625f;
            if (__any(big)) t = big ? __expf(y) - 1.0f : t;
            return t;
        } else {
            return expm1(y);
        }
    }
    // expm1: a degree-4 Taylor polynomial for |y| < 1/64 (truncation error
    // < y^4/120 < 5e-10 relative), else exp(y) - 1
    static __device__ __forceinline__ T expm1_t(T y)
    {
        if constexpr (sizeof(T) == 4) {
            float p = fmaf(y, 1.0f / 24.0f, 1.0f / 6.0f);
            p = fmaf(y, p, 0.5f);
            p = fmaf(y, p, 1.0f);
            const float big = __expf(y) - 1.0f;
            return fabsf(y) < 0.015625f ? y * p : big;   // a select, no branch
        } else {
            return expm1(y);
        }
    }
    static __device__ __forceinline__ T fma_t(T a, T b, T c)
    {
        if constexpr (sizeof(T) == 4) return fmaf(a, b, c);
        else return fma(a, b, c);
    }
    static __device__ __forceinline__ T sqrt_t(T x)
    {
        if constexpr (sizeof(T) == 4) return sqrtf(x);
        else return sqrt(x);
    }
    // FP32: compact Cody-Waite/Taylor sincos (fastmath.hpp, ~1 ulp); FP64: libm.
    static __device__ __forceinline__ void sincos_t(T x, T *s, T *c)
    {
        if constexpr (sizeof(T) == 4) fast_sincosf(x, s, c);
        else sincos(x, s, c);
    }
};

// Split form of the real-time nonlinear xy pass (FP32, itmax = 1, no noise):
// a workgroup of two wavefronts; wavefront c carries component c of the same
// 64/N planes through the transforms.  The step couples the components only
// through the densities, which the wavefronts exchange in the (then idle)
// transpose tiles.  Twice the wavefronts of xy_kernel, each with half the
// registers and work: for batches too small to fill the chip with xy_kernel.
template <int N, bool NT>
__global__ void __launch_bounds__(128) xy_split_kernel(OpNonlinRT<float, N, true> op, XyGeom g)
{
    using T = float;
    constexpr int PPW = 64 / N;
    __shared__ cpx<T> lds[2 * PPW * N * (N + 1)];
    const int comp = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pl = lane / N, t = lane % N;
    const long blk = g.xcd ? xcd_remap(blockIdx.x, g.nz / PPW) : (long)blockIdx.x;
    const long plane = blk * PPW + pl;
    const int z = (int)(plane % g.nz);
    const int traj = (int)(plane / g.nz);
    cpx<T> *tile = lds + (comp * PPW + pl) * N * (N + 1);
    const cpx<T> *other = lds + ((1 - comp) * PPW + pl) * N * (N + 1);
    const long base = plane * (long)(N * N);
    cpx<T> *arr = comp ? op.b : op.a;

    cpx<T> v[N];
#pragma unroll
    for (int y = 0; y < N; ++y) v[y] = ld<NT, T>(arr + base + y * N + t);
    dft<T, N, -1>(v);
    col2row<T, N>(tile, t, v);
    dft<T, N, -1>(v);

    // densities of row t, rows padded to N + 1 against bank conflicts
    T *dm = (T *)tile + t * (N + 1);
#pragma unroll
    for (int x = 0; x < N; ++x) dm[x] = cnorm<T>(v[x]);
    __syncthreads();
    const T *dn = (const T *)other + t * (N + 1);
    T m[N];
#pragma unroll
    for (int x = 0; x < N; ++x) m[x] = dn[x];
    __syncthreads();   // the tiles are rewritten below
    op.apply_one(v, m, comp, t, z, traj);

    dft<T, N, +1>(v);
    row2col<T, N>(tile, t, v);
    dft<T, N, +1>(v);
#pragma unroll
    for (int y = 0; y < N; ++y) st<NT, T>(arr + base + y * N + t, v[y]);
}

// Read-only sample: x-space sums per trajectory
// {sum|a|^2, sum|b|^2, Re sum conj(a) b, Im sum conj(a) b} in FP64.
template <typename T, int N> struct OpSample : NoSkip {
    static constexpr int NC = 2, SIGN_IN = -1, SIGN_OUT = 0, NSUM = 4;
    static constexpr bool WRITE = false;
    const cpx<T> *a, *b;
    __device__ const cpx<T> *src(int c) const { return c == 0 ? a : b; }
    __device__ cpx<T> *dst(int) const { return nullptr; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[2][N], int, int, int, long, double *s) const
    {
        static_for<0, N>([&](auto X_) {
            constexpr int x = decltype(X_)::value;
            const double ar = r[0][x].x, ai = r[0][x].y, br = r[1][x].x, bi = r[1][x].y;
            s[0] += ar * ar + ai * ai;
            s[1] += br * br + bi * bi;
            s[2] += ar * br + ai * bi;
            s[3] += ar * bi - ai * br;
        });
    }
};

// Plain 2D transform of one component (in -> out), scaled.
template <typename T, int N, int SIGN> struct OpFftXY : NoSkip {
    static constexpr int NC = 1, SIGN_IN = SIGN, SIGN_OUT = 0, NSUM = 0;
    static constexpr bool WRITE = true;
    const cpx<T> *in;
    cpx<T> *out;
    T scale;
    __device__ const cpx<T> *src(int) const { return in; }
    __device__ cpx<T> *dst(int) const { return out; }
    __device__ __forceinline__ void apply(cpx<T> (&r)[1][N], int, int, int, long, double *) const
    {
        if (scale != (T)1)
#pragma unroll
            for (int x = 0; x < N; ++x) r[0][x] = cscale<T>(r[0][x], scale);
    }
};

// ============================================================= z-pencil pass
// A workgroup owns PW consecutive pencils (x,y columns) of one trajectory and
// all NZ = R1 * NW of their z values.  Thread (w, p) holds z = w + NW*j,
// j < R1 (coalesced rows of PW complex values).  Forward transform: R1-point
// DFT in registers, twiddle exp(+-2 pi i w k'/NZ), exchange through LDS so a
// thread owns k' = w*KPT .. w*KPT+KPT-1 for all w, NW-point DFT in registers.
// The thread then holds psi_k at kz = k' + R1*m for m < NW.  The inverse runs
// the same path backwards.

struct ZGeom {
    int nxy;             // pencils per trajectory (NX*NY)
    int tiles_per_traj;  // nxy / PW
    int xcd;             // remap workgroups so a trajectory stays on one XCD
};

// Op interface:
//   static constexpr int NC;
//   const cpx<T>* src(int c); cpx<T>* dst(int c);
//   void kspace(int c, cpx<T> (&X)[KPT][NW], int kk0, int q, int traj)   (k-space values, kz = kk0+i + R1*m)
// blockIdx.y selects the first component (component-split launches of NC = 1).
// HALF = true exchanges real and imaginary parts in two rounds through a tile
// of NZ*PW reals (half the LDS, so more workgroups fit on a CU).
template <typename T, bool HALF> using ZSlot = typename std::conditional<HALF, T, cpx<T>>::type;

// One tile (PW pencils of one trajectory, component cb onward) of the z pass,
// by one workgroup of (NZ / R1) * PW threads, exchanging through `lds`
// (NZ * PW slots).  tid is threadIdx.x, passed as xy_tile's lane is.
template <typename T, int NZ, int R1, int PW, int SIGN_F, bool ROUNDTRIP, class Op, bool NT, bool HALF>
__device__ __forceinline__ void z_tile(const Op &op, const ZGeom &g, const cpx<T> *__restrict__ tw, long blk, int cb,
                                       ZSlot<T, HALF> *lds, int tid);

template <typename T, int NZ, int R1, int PW, int SIGN_F, bool ROUNDTRIP, class Op, bool NT = false,
          bool HALF = false>
__global__ void __launch_bounds__((NZ / R1) * PW) z_kernel(Op op, ZGeom g, const cpx<T> *__restrict__ tw)
{
    __shared__ ZSlot<T, HALF> lds[NZ * PW];
    const long blk = g.xcd ? xcd_remap(blockIdx.x, g.tiles_per_traj) : (long)blockIdx.x;
    z_tile<T, NZ, R1, PW, SIGN_F, ROUNDTRIP, Op, NT, HALF>(op, g, tw, blk, blockIdx.y, lds, threadIdx.x);
}

// The loads of one z tile: thread (p, w) reads z = w + NW j of pencil q, j < R1.
template <typename T, int NZ, int R1, int PW, class Op, bool NT>
__device__ __forceinline__ void z_tile_load(const Op &op, const ZGeom &g, long blk, int cb, int tid,
                                            cpx<T> (&v)[Op::NC][R1])
{
    constexpr int NW = NZ / R1;
    const int p = tid % PW;
    const int w = PW == 64 ? __builtin_amdgcn_readfirstlane(tid / PW) : tid / PW;
    const int traj = (int)(blk / g.tiles_per_traj);
    const int q = (int)(blk % g.tiles_per_traj) * PW + p;
    const long base = (long)traj * NZ * g.nxy + q;
#pragma unroll
    for (int c = 0; c < Op::NC; ++c) {
        const cpx<T> *s = op.src(c + cb);
#pragma unroll
        for (int j = 0; j < R1; ++j) v[c][j] = ld<NT, T>(s + base + (long)(w + NW * j) * g.nxy);
    }
}

// The rest of one z tile (transforms, operator, stores) on the values
// z_tile_load read, with op already begun for the tile's trajectory.
template <typename T, int NZ, int R1, int PW, int SIGN_F, bool ROUNDTRIP, class Op, bool NT, bool HALF>
__device__ __forceinline__ void z_tile_compute(Op &op, const ZGeom &g, const cpx<T> *__restrict__ tw, long blk,
                                               int cb, ZSlot<T, HALF> *lds, int tid, cpx<T> (&v)[Op::NC][R1]);

template <typename T, int NZ, int R1, int PW, int SIGN_F, bool ROUNDTRIP, class Op, bool NT, bool HALF>
__device__ __forceinline__ void z_tile(const Op &op_, const ZGeom &g, const cpx<T> *__restrict__ tw, long blk, int cb,
                                       ZSlot<T, HALF> *lds, int tid)
{
    Op op = op_;
    const int traj = (int)(blk / g.tiles_per_traj);
    op.begin(traj, blk % g.tiles_per_traj == 0 && cb == 0);
    if (op.skip(traj)) return;
    cpx<T> v[Op::NC][R1];
    z_tile_load<T, NZ, R1, PW, Op, NT>(op, g, blk, cb, tid, v);
    z_tile_compute<T, NZ, R1, PW, SIGN_F, ROUNDTRIP, Op, NT, HALF>(op, g, tw, blk, cb, lds, tid, v);
}

template <typename T, int NZ, int R1, int PW, int SIGN_F, bool ROUNDTRIP, class Op, bool NT, bool HALF>
__device__ __forceinline__ void z_tile_compute(Op &op, const ZGeom &g, const cpx<T> *__restrict__ tw, long blk,
                                               int cb, ZSlot<T, HALF> *lds, int tid, cpx<T> (&v)[Op::NC][R1])
{
    constexpr int NW = NZ / R1;
    constexpr int KPT = R1 / NW;
    constexpr int NC = Op::NC;
    static_assert(R1 % NW == 0, "R1 must be a multiple of NW");
    using Slot = ZSlot<T, HALF>;
    const int p = tid % PW;
    // with PW = 64 a wavefront is one z subset: make w provably uniform so the
    // twiddle and kz-table loads that depend only on it become scalar loads
    const int w = PW == 64 ? __builtin_amdgcn_readfirstlane(tid / PW) : tid / PW;
    const int traj = (int)(blk / g.tiles_per_traj);
    const int q = (int)(blk % g.tiles_per_traj) * PW + p;
    const long base = (long)traj * NZ * g.nxy + q;
    const int kk0 = w * KPT;

    // v[k] -> slot (k, w); X[i][w2] <- slot (kk0 + i, w2)   (forward exchange)
    auto to_k = [&](cpx<T> (&v)[R1], cpx<T> (&X)[KPT][NW]) {
        if constexpr (HALF) {
#pragma unroll
            for (int part = 0; part < 2; ++part) {
#pragma unroll
                for (int k = 0; k < R1; ++k) lds[(k * NW + w) * PW + p] = part ? v[k].y : v[k].x;
                __syncthreads();
#pragma unroll
                for (int i = 0; i < KPT; ++i)
#pragma unroll
                    for (int w2 = 0; w2 < NW; ++w2) {
                        const T val = lds[((kk0 + i) * NW + w2) * PW + p];
                        if (part) X[i][w2].y = val;
                        else X[i][w2].x = val;
                    }
                __syncthreads();
            }
        } else {
#pragma unroll
            for (int k = 0; k < R1; ++k) lds[(k * NW + w) * PW + p] = v[k];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < KPT; ++i)
#pragma unroll
                for (int w2 = 0; w2 < NW; ++w2) X[i][w2] = lds[((kk0 + i) * NW + w2) * PW + p];
            __syncthreads();
        }
    };
    // X[i][w2] -> slot (kk0 + i, w2); v[k] <- slot (k, w)   (inverse exchange)
    auto from_k = [&](cpx<T> (&X)[KPT][NW], cpx<T> (&v)[R1]) {
#pragma unroll
        for (int part = 0; part < (HALF ? 2 : 1); ++part) {
#pragma unroll
            for (int i = 0; i < KPT; ++i)
#pragma unroll
                for (int w2 = 0; w2 < NW; ++w2) {
                    Slot val;
                    if constexpr (HALF) val = part ? X[i][w2].y : X[i][w2].x;
                    else val = X[i][w2];
                    lds[((kk0 + i) * NW + w2) * PW + p] = val;
                }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < R1; ++k) {
                const Slot val = lds[(k * NW + w) * PW + p];
                if constexpr (HALF) {
                    if (part) v[k].y = val;
                    else v[k].x = val;
                } else {
                    v[k] = val;
                }
            }
            __syncthreads();
        }
    };

#pragma unroll
    for (int c = 0; c < NC; ++c) {
        // the first stage's twiddles also offset the modulus errors of the NW-point stage
        dft<T, R1, SIGN_F, NW>(v[c]);
#pragma unroll
        for (int k = 1; k < R1; ++k) {
            cpx<T> f = tw[(w * k) % NZ];
            if (SIGN_F < 0) f.y = -f.y;
            v[c][k] = cmul<T>(v[c][k], f);
        }
        cpx<T> X[KPT][NW];
        to_k(v[c], X);
#pragma unroll
        for (int i = 0; i < KPT; ++i) dft<T, NW, SIGN_F>(X[i]);

        op.kspace(c, X, kk0, q, traj);

        if constexpr (ROUNDTRIP) {
#pragma unroll
            for (int i = 0; i < KPT; ++i) {
                dft<T, NW, -SIGN_F>(X[i]);
#pragma unroll
                for (int w2 = 1; w2 < NW; ++w2) {
                    cpx<T> f = tw[(w2 * (kk0 + i)) % NZ];
                    if (SIGN_F > 0) f.y = -f.y;
                    X[i][w2] = cmul<T>(X[i][w2], f);
                }
            }
            from_k(X, v[c]);
            dft<T, R1, -SIGN_F, NW>(v[c]);
        } else {
            cpx<T> *d = op.dst(c + cb);
#pragma unroll
            for (int i = 0; i < KPT; ++i)
#pragma unroll
                for (int m = 0; m < NW; ++m) st<NT, T>(d + base + (long)(kk0 + i + R1 * m) * g.nxy, X[i][m]);
        }
    }
    if constexpr (ROUNDTRIP) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            cpx<T> *d = op.dst(c + cb);
#pragma unroll
            for (int j = 0; j < R1; ++j) st<NT, T>(d + base + (long)(w + NW * j) * g.nxy, v[c][j]);
        }
    }
}

// ------------------------------------------------------------- z operators

// Kinetic phase in full k-space: psi_k *= Kxy[ky,kx] * Kz[kz], the tables
// holding exp(i c dt (kx^2+ky^2)/2) and exp(i c dt kz^2/2)/NZ
// (propagateKSpaceRealTime, wigner_kernel.cu:137-151; separable form).
template <typename T, int NZ, int R1, int NC_> struct OpKinetic : NoSkip {
    static constexpr int NC = NC_;
    static constexpr int NW = NZ / R1, KPT = R1 / NW;
    cpx<T> *a, *b;
    const cpx<T> *kxy;   // [records][nxy]
    const cpx<T> *kz;    // [records][NZ]
    int nxy;
    int rec_shift;
    __device__ const cpx<T> *src(int c) const { return c == 0 ? a : b; }
    __device__ cpx<T> *dst(int c) const { return c == 0 ? a : b; }
    __device__ __forceinline__ void kspace(int, cpx<T> (&X)[KPT][NW], int kk0, int q, int traj) const
    {
        const int rec = traj >> rec_shift;
        const cpx<T> fxy = kxy[(long)rec * nxy + q];
        const cpx<T> *kzr = kz + (long)rec * NZ;
        // the two factors one after the other (the same instruction count as
        // forming their product first): the modulus error a mode sees every step
        // is then that of the two tables' entries, chosen near one, rather than
        // the rounding of their product
#pragma unroll
        for (int i = 0; i < KPT; ++i)
#pragma unroll
            for (int m = 0; m < NW; ++m)
                X[i][m] = cmul_fixed<T>(cmul_fixed<T>(X[i][m], fxy), kzr[kk0 + i + R1 * m]);
    }
};

// Pure transform along z (one component, in -> out), scaled.
template <typename T, int NZ, int R1> struct OpScaleZ : NoSkip {
    static constexpr int NC = 1;
    static constexpr int NW = NZ / R1, KPT = R1 / NW;
    const cpx<T> *in;
    cpx<T> *out;
    T scale;
    __device__ const cpx<T> *src(int) const { return in; }
    __device__ cpx<T> *dst(int) const { return out; }
    __device__ __forceinline__ void kspace(int, cpx<T> (&X)[KPT][NW], int, int, int) const
    {
        if (scale != (T)1)
#pragma unroll
            for (int i = 0; i < KPT; ++i)
#pragma unroll
                for (int m = 0; m < NW; ++m) X[i][m] = cscale<T>(X[i][m], scale);
    }
};

} // namespace bec
