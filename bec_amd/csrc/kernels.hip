// Kernel launchers for gfx950: dispatch from runtime lattice shape and
// precision to the compile-time pass templates of passes.hpp.
#include "launch.hpp"
#include "philox.hpp"

namespace bec {

bool shape_supported(int nx, int ny, int nz)
{
    return nx == ny && (nx == 16 || nx == 32) && (nz == 64 || nz == 128 || nz == 256);
}

size_t elem_size(int prec) { return prec == 0 ? 8 : 16; }

long xy_workgroups(const DevGroup &g, int ntraj) { return (long)ntraj * g.nz / (64 / g.nx); }
int xy_wg_per_traj(const DevGroup &g) { return g.nz / (64 / g.nx); }

int default_pw(int prec, int nx, int nz)
{
    // keep the LDS exchange tile at <= 64 KiB
    int pw = 65536 / (nz * (int)elem_size(prec));
    if (pw > 64) pw = 64;
    if (pw > nx * nx) pw = nx * nx;
    return pw;
}

int default_r1(int nz) { return nz == 64 ? 8 : 16; }

int schedule_candidates(int prec, int nx, int nz, int ntraj, Schedule *out, int max_out, bool all)
{
    int n = 0;
    auto add = [&](Schedule c) {
        if (n < max_out) out[n] = c;
        ++n;
    };
    const int xcds = ntraj % 8 == 0 ? 2 : 1;
    if (prec == 1) {   // FP64: default radix, temporal accesses
        for (int pw : {16, 32, 64})
            for (int split = 0; split < 2; ++split)
                for (int half = 0; half < 2; ++half)
                    for (int xcd = 0; xcd < xcds; ++xcd) {
                        const long lds = (long)nz * pw * (half ? 8 : 16);
                        if (lds > 65536 || pw > nx * nx) continue;
                        add(Schedule{pw, default_r1(nz), split, 0, xcd, 1, half});
                    }
        return n;
    }
    // the z pass is instantiated only for LDS exchange tiles up to 64 KiB and
    // workgroups up to 1024 threads (launch_z)
    auto fits = [&](int pw, int r1, int half) {
        return (long)nz * pw * (half ? 4 : 8) <= 65536 && (nz / r1) * pw <= 1024 && pw <= nx * nx;
    };
    const int r1s[2] = {default_r1(nz), nz == 64 ? 16 : 32};
    for (int pw : {32, 64})
        for (int r1 : r1s)
            for (int nt = 0; nt < 2; ++nt)
                for (int half = 0; half < 2; ++half)
                    for (int xcd = 0; xcd < xcds; ++xcd)
                        if (fits(pw, r1, half)) add(Schedule{pw, r1, 1, nt, xcd, 1, half});
    // the split xy pass (one wavefront per component) on the XCD-aware family,
    // offered only while the two-component pass's wavefronts do not fill the GPU
    // once (4 per SIMD at 16-point rows, 2 at 32): it exists to double them for
    // small batches; at 2048 trajectories it is 0.8% slower per step and timed
    // trials cannot tell the two apart
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;   // MI355X
    const long xy_waves = (long)ntraj * nz / (64 / nx), xy_slots = (long)cus * 4 * (nx == 16 ? 4 : 2);
    if (all || xy_waves < xy_slots)
        for (int pw : {32, 64})
            for (int r1 : r1s)
                for (int nt = 0; nt < 2; ++nt)
                    for (int half = 0; half < 2; ++half)
                        if (fits(pw, r1, half)) add(Schedule{pw, r1, 1, nt, xcds - 1, 1, half, 1});
    // the xy occupancy target on the streaming, XCD-aware family
    for (int pw : {32, 64})
        for (int half = 0; half < 2; ++half)
            if (fits(pw, default_r1(nz), half)) add(Schedule{pw, default_r1(nz), 1, 1, xcds - 1, 5, half});
    // the unsplit z pass for reference
    if (fits(default_pw(prec, nx, nz), default_r1(nz), 0))
        add(Schedule{default_pw(prec, nx, nz), default_r1(nz), 0, 1, xcds - 1, 1, 0});
    return n;
}

// ------------------------------------------------------------------ passes

template <typename T, int NZ, int R1, int NC>
static OpKinetic<T, NZ, R1, NC> make_kinetic(const DevGroup &g, int variant)
{
    const int nxy = g.nx * g.ny;
    OpKinetic<T, NZ, R1, NC> op;
    op.a = (cpx<T> *)g.a;
    op.b = (cpx<T> *)g.b;
    op.kxy = (const cpx<T> *)g.kxy + (long)variant * g.n_records * nxy;
    op.kz = (const cpx<T> *)g.kz + (long)variant * g.n_records * NZ;
    op.nxy = nxy;
    op.rec_shift = g.rec_shift;
    return op;
}

hipError_t run_z_kinetic(const DevGroup &g, int variant, hipStream_t s)
{
    const int pw = g.pw ? g.pw : default_pw(g.prec, g.nx, g.nz);
    const int r1 = g.z_r1 ? g.z_r1 : default_r1(g.nz);
    return dispatch(g, [&](auto Tt, auto, auto NZt) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int NZ = decltype(NZt)::value;
        return with_pw(pw, [&](auto PWt) -> hipError_t {
            constexpr int PW = decltype(PWt)::value;
            return with_bool(g.z_split, [&](auto SPt) -> hipError_t {
                constexpr bool SPLIT = decltype(SPt)::value;
                constexpr int NC = SPLIT ? 1 : 2;
                return with_bool(g.z_half, [&](auto HFt) -> hipError_t {
                    constexpr bool HALF = decltype(HFt)::value;
                    if constexpr (sizeof(T) == 8) {   // FP64: default radix, temporal accesses
                        constexpr int R1 = z_r1<NZ>();
                        return launch_z<T, NZ, PW, +1, true, OpKinetic<T, NZ, R1, NC>, R1, false, HALF>(
                            g, make_kinetic<T, NZ, R1, NC>(g, variant), g.ntraj, s, SPLIT ? 2 : 1, g.xcd);
                    } else if constexpr (!SPLIT) {   // unsplit: default radix, streaming, full tile
                        constexpr int R1 = z_r1<NZ>();
                        if (HALF) return hipErrorInvalidValue;
                        return launch_z<T, NZ, PW, +1, true, OpKinetic<T, NZ, R1, NC>, R1, true, false>(
                            g, make_kinetic<T, NZ, R1, NC>(g, variant), g.ntraj, s, 1, g.xcd);
                    } else {
                        return with_bool(g.nt, [&](auto NTt) -> hipError_t {
                            constexpr bool NT = decltype(NTt)::value;
                            if (r1 == z_r1<NZ>()) {
                                constexpr int R1 = z_r1<NZ>();
                                return launch_z<T, NZ, PW, +1, true, OpKinetic<T, NZ, R1, NC>, R1, NT, HALF>(
                                    g, make_kinetic<T, NZ, R1, NC>(g, variant), g.ntraj, s, 2, g.xcd);
                            }
                            if (r1 == z_r1_alt<NZ>()) {
                                constexpr int R1 = z_r1_alt<NZ>();
                                return launch_z<T, NZ, PW, +1, true, OpKinetic<T, NZ, R1, NC>, R1, NT, HALF>(
                                    g, make_kinetic<T, NZ, R1, NC>(g, variant), g.ntraj, s, 2, g.xcd);
                            }
                            return hipErrorInvalidValue;
                        });
                    }
                });
            });
        });
    });
}

// Advances the loss-noise step counter after the pass that read it (one
// thread, a vector store; captured into step graphs like the passes).
__global__ void noise_tick_kernel(unsigned *step)
{
    if (threadIdx.x == 0) step[0] = step[0] + 1u;
}

template <typename T, int N, bool IT1>
static hipError_t launch_nonlinear(const DevGroup &g, hipStream_t s)
{
    if (g.noise_step) {
        OpNonlinRT<T, N, IT1, true> op;
        op.a = (cpx<T> *)g.a;
        op.b = (cpx<T> *)g.b;
        op.P = g.params;
        op.rec_shift = g.rec_shift;
        op.step = g.noise_step;
        op.k0 = g.noise_k0;
        op.k1 = g.noise_k1;
        op.group = (uint32_t)g.group_index;
        hipError_t e = with_bool(g.nt, [&](auto NTt) -> hipError_t {
            constexpr bool NT = decltype(NTt)::value;
            return launch_xy<T, N, OpNonlinRT<T, N, IT1, true>, NT, 1>(g, op, g.ntraj, nullptr, s, g.xcd);
        });
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(noise_tick_kernel, dim3(1), dim3(64), 0, s, g.noise_step);
        return hipGetLastError();
    }
    OpNonlinRT<T, N, IT1> op;
    op.a = (cpx<T> *)g.a;
    op.b = (cpx<T> *)g.b;
    op.P = g.params;
    op.rec_shift = g.rec_shift;
    return with_bool(g.nt, [&](auto NTt) -> hipError_t {
        constexpr bool NT = decltype(NTt)::value;
        if constexpr (sizeof(T) == 4 && IT1) {
            if (g.xy_split) {
                constexpr int PPW = 64 / N;
                XyGeom geo{g.nz, (long)g.ntraj * g.nz, (g.xcd && g.ntraj % 8 == 0) ? 1 : 0};
                if (geo.n_planes % PPW) return hipErrorInvalidValue;
                hipLaunchKernelGGL((xy_split_kernel<N, NT>), dim3((unsigned)(geo.n_planes / PPW)), dim3(128), 0, s,
                                   op, geo);
                return hipGetLastError();
            }
            if (g.xy_minw == 5) return launch_xy<T, N, OpNonlinRT<T, N, IT1>, NT, 5>(g, op, g.ntraj, nullptr, s, g.xcd);
        }
        return launch_xy<T, N, OpNonlinRT<T, N, IT1>, NT, 1>(g, op, g.ntraj, nullptr, s, g.xcd);
    });
}

hipError_t run_xy_nonlinear(const DevGroup &g, hipStream_t s)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        if (g.itmax1) return launch_nonlinear<T, N, true>(g, s);
        return launch_nonlinear<T, N, false>(g, s);
    });
}

// calculateAverages (wigner.cu:305-337) per cell: one thread per cell walks the
// trajectories twice (mean, then squared deviations, as kernelAverage,
// wigner_kernel.cu:383-392) in FP64; out [cells][6] = {sum|a|^2, sum|b|^2,
// dev Re a, dev Im a, dev Re b, dev Im b}.
template <typename T>
__global__ __launch_bounds__(256) void averages_kernel(const cpx<T> *a, const cpx<T> *b, long cells, int ntraj,
                                                       double *out)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= cells) return;
    double m[4] = {0, 0, 0, 0}, sa = 0.0, sb = 0.0;
    for (int j = 0; j < ntraj; ++j) {
        const cpx<T> x = a[(long)j * cells + i], y = b[(long)j * cells + i];
        m[0] += (double)x.x;
        m[1] += (double)x.y;
        m[2] += (double)y.x;
        m[3] += (double)y.y;
        sa += (double)x.x * x.x + (double)x.y * x.y;
        sb += (double)y.x * y.x + (double)y.y * y.y;
    }
    for (int k = 0; k < 4; ++k) m[k] /= ntraj;
    double d[4] = {0, 0, 0, 0};
    for (int j = 0; j < ntraj; ++j) {
        const cpx<T> x = a[(long)j * cells + i], y = b[(long)j * cells + i];
        const double v[4] = {(double)x.x, (double)x.y, (double)y.x, (double)y.y};
        for (int k = 0; k < 4; ++k) d[k] += (v[k] - m[k]) * (v[k] - m[k]);
    }
    double *o = out + i * 6;
    o[0] = sa;
    o[1] = sb;
    for (int k = 0; k < 4; ++k) o[2 + k] = d[k];
}

hipError_t run_averages(const DevGroup &g, const void *ax, const void *bx, double *out, hipStream_t s)
{
    const long cells = (long)g.nx * g.ny * g.nz;
    const dim3 grid((unsigned)((cells + 255) / 256));
    if (g.prec == 0)
        hipLaunchKernelGGL(averages_kernel<float>, grid, dim3(256), 0, s, (const cpx<float> *)ax,
                           (const cpx<float> *)bx, cells, g.ntraj, out);
    else
        hipLaunchKernelGGL(averages_kernel<double>, grid, dim3(256), 0, s, (const cpx<double> *)ax,
                           (const cpx<double> *)bx, cells, g.ntraj, out);
    return hipGetLastError();
}

hipError_t run_xy_sample(const DevGroup &g, double *partial, hipStream_t s)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        OpSample<T, N> op;
        op.a = (const cpx<T> *)g.a;
        op.b = (const cpx<T> *)g.b;
        return launch_xy<T, N>(g, op, g.ntraj, partial, s, g.xcd);
    });
}

hipError_t run_xy_fft(const DevGroup &g, const void *in, void *out, int ntraj, int sign, double scale,
                      hipStream_t s)
{
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        if (sign < 0) {
            OpFftXY<T, N, -1> op;
            op.in = (const cpx<T> *)in;
            op.out = (cpx<T> *)out;
            op.scale = (T)scale;
            return launch_xy<T, N>(g, op, ntraj, nullptr, s);
        }
        OpFftXY<T, N, +1> op;
        op.in = (const cpx<T> *)in;
        op.out = (cpx<T> *)out;
        op.scale = (T)scale;
        return launch_xy<T, N>(g, op, ntraj, nullptr, s);
    });
}

hipError_t run_z_fft(const DevGroup &g, const void *in, void *out, int ntraj, int sign, double scale,
                     hipStream_t s)
{
    int pw = g.pw ? g.pw : default_pw(g.prec, g.nx, g.nz);
    return dispatch(g, [&](auto Tt, auto, auto NZt) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int NZ = decltype(NZt)::value;
        constexpr int R1 = z_r1<NZ>();
        OpScaleZ<T, NZ, R1> op;
        op.in = (const cpx<T> *)in;
        op.out = (cpx<T> *)out;
        op.scale = (T)scale;
        return with_pw(pw, [&](auto PWt) {
            constexpr int PW = decltype(PWt)::value;
            if (sign < 0) return launch_z<T, NZ, PW, -1, false>(g, op, ntraj, s);
            return launch_z<T, NZ, PW, +1, false>(g, op, ntraj, s);
        });
    });
}

// ------------------------------------------------------------- reductions

__global__ void reduce_partials_kernel(const double *__restrict__ partial, int per_traj, int ntraj,
                                       int nsum, double *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntraj * nsum) return;
    int traj = i / nsum, k = i % nsum;
    const double *p = partial + (long)traj * per_traj * nsum + k;
    double s = 0.0;
    for (int j = 0; j < per_traj; ++j) s += p[(long)j * nsum];
    out[i] = s;
}

hipError_t run_reduce_partials(const double *partial, int per_traj, int ntraj, int nsum, double *out,
                               hipStream_t s)
{
    int n = ntraj * nsum;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((n + 255) / 256), dim3(256), 0, s, partial, per_traj,
                       ntraj, nsum, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ noise init
// Philox noise (philox.hpp) with counter {cell, traj, 0, 0} and key = seed;
// initializeEnsembles (wigner_kernel.cu:295-317).

template <typename T>
__global__ void noise_init_kernel(const cpx<T> *__restrict__ gs, cpx<T> *__restrict__ a, cpx<T> *__restrict__ b,
                                  long cells, long n, uint32_t k0, uint32_t k1, double Va, double Vb,
                                  double coeff)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long cell = i % cells;
    const long traj = i / cells;
    double eta[4];
    noise4((uint32_t)cell, (uint32_t)traj, 0u, k0, k1, eta);
    const cpx<T> g = gs[cell];
    a[i] = mk<T>((T)((double)g.x + Va * eta[0] * coeff), (T)((double)g.y + Va * eta[1] * coeff));
    b[i] = mk<T>((T)(Vb * eta[2] * coeff), (T)(Vb * eta[3] * coeff));
}

hipError_t run_noise_init(const DevGroup &g, const void *gs, uint64_t seed, double Va, double Vb,
                          double coeff, hipStream_t s)
{
    long cells = (long)g.nx * g.ny * g.nz, n = cells * g.ntraj;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (g.prec == 0)
        hipLaunchKernelGGL(noise_init_kernel<float>, grid, block, 0, s, (const float2 *)gs, (float2 *)g.a,
                           (float2 *)g.b, cells, n, k0, k1, Va, Vb, coeff);
    else
        hipLaunchKernelGGL(noise_init_kernel<double>, grid, block, 0, s, (const double2 *)gs, (double2 *)g.a,
                           (double2 *)g.b, cells, n, k0, k1, Va, Vb, coeff);
    return hipGetLastError();
}

// ----------------------------------------------------------------- pulses
// applyHalfPiPulse / applyPiPulse (wigner_kernel.cu:320-341); pointwise and
// linear, so valid in any linear representation of the state.
template <typename T>
__global__ void pulse_kernel(cpx<T> *__restrict__ a, cpx<T> *__restrict__ b, long n, int kind)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cpx<T> a0 = a[i], b0 = b[i];
    if (kind == 0) {
        const T r = (T)0.70710678118654752440;
        a[i] = mk<T>((a0.x + b0.y) * r, (a0.y - b0.x) * r);   // (a - i b)/sqrt2
        b[i] = mk<T>((a0.y + b0.x) * r, (b0.y - a0.x) * r);   // (-i a + b)/sqrt2
    } else {
        a[i] = mk<T>(b0.y, -b0.x);                            // -i b
        b[i] = mk<T>(a0.y, -a0.x);                            // -i a
    }
}

// A resonant pulse: (a, b) <- U (a, b) with U = [[u00, u01], [u10, u11]].
template <typename T>
__global__ void rotation_kernel(cpx<T> *__restrict__ a, cpx<T> *__restrict__ b, long n, double2 u00, double2 u01,
                                double2 u10, double2 u11)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cpx<T> a0 = a[i], b0 = b[i];
    const cpx<T> m00 = mk<T>((T)u00.x, (T)u00.y), m01 = mk<T>((T)u01.x, (T)u01.y);
    const cpx<T> m10 = mk<T>((T)u10.x, (T)u10.y), m11 = mk<T>((T)u11.x, (T)u11.y);
    a[i] = cadd<T>(cmul<T>(m00, a0), cmul<T>(m01, b0));
    b[i] = cadd<T>(cmul<T>(m10, a0), cmul<T>(m11, b0));
}

hipError_t run_rotation(const DevGroup &g, const double u[8], hipStream_t s)
{
    long n = (long)g.nx * g.ny * g.nz * g.ntraj;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    const double2 u00 = make_double2(u[0], u[1]), u01 = make_double2(u[2], u[3]);
    const double2 u10 = make_double2(u[4], u[5]), u11 = make_double2(u[6], u[7]);
    if (g.prec == 0)
        hipLaunchKernelGGL(rotation_kernel<float>, grid, block, 0, s, (float2 *)g.a, (float2 *)g.b, n, u00, u01, u10,
                           u11);
    else
        hipLaunchKernelGGL(rotation_kernel<double>, grid, block, 0, s, (double2 *)g.a, (double2 *)g.b, n, u00, u01,
                           u10, u11);
    return hipGetLastError();
}

hipError_t run_pulse(const DevGroup &g, int kind, hipStream_t s)
{
    long n = (long)g.nx * g.ny * g.nz * g.ntraj;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (g.prec == 0)
        hipLaunchKernelGGL(pulse_kernel<float>, grid, block, 0, s, (float2 *)g.a, (float2 *)g.b, n, kind);
    else
        hipLaunchKernelGGL(pulse_kernel<double>, grid, block, 0, s, (double2 *)g.a, (double2 *)g.b, n, kind);
    return hipGetLastError();
}

// ----------------------------------------------------- density projections
// Each workgroup owns 64/N planes at fixed z and loops over the trajectories,
// accumulating the vacuum-subtracted densities after halfPiRotate(alpha)
// (wigner_kernel.cu:283-286,344-360) in FP64; deterministic order.
template <typename T, int N>
__global__ void __launch_bounds__(64) cell_sums_kernel(const cpx<T> *__restrict__ a, const cpx<T> *__restrict__ b,
                                                       int nz, int ntraj, double ca, double sa, double vac,
                                                       double *__restrict__ out)
{
    constexpr int PPW = 64 / N;
    __shared__ cpx<T> lds[PPW * N * (N + 1)];
    const int lane = threadIdx.x, pl = lane / N, t = lane % N;
    const int z = blockIdx.x * PPW + pl;
    cpx<T> *tile = lds + pl * N * (N + 1);
    const double r = 0.70710678118654752440;
    double acc[2][N];
#pragma unroll
    for (int x = 0; x < N; ++x) acc[0][x] = acc[1][x] = 0.0;
    for (int traj = 0; traj < ntraj; ++traj) {
        const long base = ((long)traj * nz + z) * (N * N);
        cpx<T> v[2][N];
#pragma unroll
        for (int y = 0; y < N; ++y) {
            v[0][y] = a[base + y * N + t];
            v[1][y] = b[base + y * N + t];
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            dft<T, N, -1>(v[c]);
            col2row<T, N>(tile, t, v[c]);
            dft<T, N, -1>(v[c]);
        }
#pragma unroll
        for (int x = 0; x < N; ++x) {
            const double ar = v[0][x].x, ai = v[0][x].y, br = v[1][x].x, bi = v[1][x].y;
            // a' = (a + b (sin - i cos))/sqrt2, b' = (a (-sin - i cos) + b)/sqrt2
            const double anr = (ar + br * sa + bi * ca) * r, ani = (ai + bi * sa - br * ca) * r;
            const double bnr = (-ar * sa + ai * ca + br) * r, bni = (-ai * sa - ar * ca + bi) * r;
            acc[0][x] += anr * anr + ani * ani - vac;
            acc[1][x] += bnr * bnr + bni * bni - vac;
        }
    }
#pragma unroll
    for (int x = 0; x < N; ++x) {
        const long cell = ((long)z * N + t) * N + x;
        out[cell * 2] = acc[0][x];
        out[cell * 2 + 1] = acc[1][x];
    }
}

hipError_t run_cell_density_sums(const DevGroup &g, double alpha, double vac, double *out, hipStream_t s)
{
    double ca = cos(alpha), sa = sin(alpha);
    return dispatch(g, [&](auto Tt, auto Nt, auto) -> hipError_t {
        using T = typename decltype(Tt)::type;
        constexpr int N = decltype(Nt)::value;
        constexpr int PPW = 64 / N;
        if (g.nz % PPW) return hipErrorInvalidValue;
        hipLaunchKernelGGL((cell_sums_kernel<T, N>), dim3(g.nz / PPW), dim3(64), 0, s, (const cpx<T> *)g.a,
                           (const cpx<T> *)g.b, g.nz, g.ntraj, ca, sa, vac, out);
        return hipGetLastError();
    });
}

// Projections of the per-cell sums cs[cell*2 + c] (run_cell_density_sums):
// out = {xy[2][ny][nx] (sum over z), zy[2][ny][nz] (sum over x),
// axial[2][nz] (sum over y, then x)}, one thread per output value, each sum in
// increasing index order (wigner.cu:279-303,388-410).
__global__ void __launch_bounds__(256) projection_sums_kernel(const double *__restrict__ cs, int nx, int ny, int nz,
                                                             double *__restrict__ out)
{
    const int nxy = nx * ny, nyz = ny * nz;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n_out = 2 * (nxy + nyz + nz);
    if (i >= n_out) return;
    double s = 0.0;
    if (i < 2 * nxy) {
        const int c = i / nxy, q = i - c * nxy;   // q = y nx + x
        for (int z = 0; z < nz; ++z) s += cs[((long)z * nxy + q) * 2 + c];
    } else if (i < 2 * (nxy + nyz)) {
        const int r = i - 2 * nxy, c = r / nyz, q = r - c * nyz, y = q / nz, z = q - y * nz;
        for (int x = 0; x < nx; ++x) s += cs[(((long)z * ny + y) * nx + x) * 2 + c];
    } else {
        const int r = i - 2 * (nxy + nyz), c = r / nz, z = r - c * nz;
        for (int q = 0; q < nxy; ++q) s += cs[((long)z * nxy + q) * 2 + c];
    }
    out[i] = s;
}

long projection_doubles(const DevGroup &g)
{
    const long cells = (long)g.nx * g.ny * g.nz;
    return 2 * cells + 2 * ((long)g.nx * g.ny + (long)g.ny * g.nz + g.nz);
}

hipError_t run_projections(const DevGroup &g, double alpha, double vac, double *buf, hipStream_t s)
{
    hipError_t e = run_cell_density_sums(g, alpha, vac, buf, s);
    if (e != hipSuccess) return e;
    const long cells = (long)g.nx * g.ny * g.nz;
    const int n_out = (int)(projection_doubles(g) - 2 * cells);
    hipLaunchKernelGGL(projection_sums_kernel, dim3((n_out + 255) / 256), dim3(256), 0, s, buf, g.nx, g.ny, g.nz,
                       buf + 2 * cells);
    return hipGetLastError();
}

// ------------------------------------------------------------------- copy
// A streaming copy for the achievable-bandwidth reference: one 16-byte
// non-temporal load and store per thread over a grid that covers the buffer.
// Measured on 2 GiB buffers it moves 6.57 TB/s (read + write), above every
// variant with 2 or 4 loads per thread in flight (5.80-5.88 TB/s), with
// temporal accesses (5.57-6.12 TB/s), or with 512- and 1024-thread
// workgroups (5.80-6.02 TB/s).
typedef float v4f_t __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy_kernel(const v4f_t *__restrict__ src, v4f_t *__restrict__ dst, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

hipError_t run_copy(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    const long n = (long)(bytes / 16);
    const long blocks = (n + 255) / 256;
    if (blocks < 1 || blocks > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const v4f_t *)src, (v4f_t *)dst, n);
    return hipGetLastError();
}

// ------------------------------------------------------------- conversion
template <typename A, typename B>
__global__ void convert_kernel(const A *__restrict__ in, B *__restrict__ out, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (B)in[i];
}

hipError_t run_convert(int prec_from, const void *in, int prec_to, void *out, long n_complex, hipStream_t s)
{
    long n = 2 * n_complex;
    dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (prec_from == 1 && prec_to == 0)
        hipLaunchKernelGGL((convert_kernel<double, float>), grid, block, 0, s, (const double *)in, (float *)out, n);
    else if (prec_from == 0 && prec_to == 1)
        hipLaunchKernelGGL((convert_kernel<float, double>), grid, block, 0, s, (const float *)in, (double *)out, n);
    else
        return hipMemcpyAsync(out, in, n * (prec_to == 0 ? 4 : 8), hipMemcpyDeviceToDevice, s);
    return hipGetLastError();
}

} // namespace bec
