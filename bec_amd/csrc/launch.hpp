// Compile-time dispatch helpers shared by the kernel launcher files.
#pragma once

#include <type_traits>

#include "backend.hpp"
#include "passes.hpp"

namespace bec {

template <typename T> struct tag { using type = T; };
template <int V> using ic = std::integral_constant<int, V>;

// z-pass radix split: NZ = R1 * NW with R1 values per thread in registers
template <int NZ> constexpr int z_r1() { return NZ == 64 ? 8 : 16; }
template <int NZ> constexpr int z_r1_alt() { return NZ == 64 ? 16 : 32; }

template <class F> inline hipError_t dispatch(const DevGroup &g, F &&f)
{
    auto with_nz = [&](auto Tt, auto Nt) -> hipError_t {
        switch (g.nz) {
        case 64: return f(Tt, Nt, ic<64>{});
        case 128: return f(Tt, Nt, ic<128>{});
        case 256: return f(Tt, Nt, ic<256>{});
        default: return hipErrorInvalidValue;
        }
    };
    auto with_n = [&](auto Tt) -> hipError_t {
        if (g.nx != g.ny) return hipErrorInvalidValue;
        switch (g.nx) {
        case 16: return with_nz(Tt, ic<16>{});
        case 32: return with_nz(Tt, ic<32>{});
        default: return hipErrorInvalidValue;
        }
    };
    return g.prec == 0 ? with_n(tag<float>{}) : with_n(tag<double>{});
}

// Runtime PW (pencils per workgroup) -> compile-time instantiation
template <class F> inline hipError_t with_pw(int pw, F &&f)
{
    switch (pw) {
    case 16: return f(ic<16>{});
    case 32: return f(ic<32>{});
    case 64: return f(ic<64>{});
    default: return hipErrorInvalidValue;
    }
}

template <class F> inline hipError_t with_bool(int v, F &&f)
{
    return v ? f(std::true_type{}) : f(std::false_type{});
}

// ncomp_grid > 1: component-split launch (grid.y = ncomp_grid, Op::NC == 1)
template <typename T, int NZ, int PW, int SIGN_F, bool RT, class Op, int R1 = z_r1<NZ>(), bool NT = false,
          bool HALF = false>
inline hipError_t launch_z(const DevGroup &g, const Op &op, int ntraj, hipStream_t s, int ncomp_grid = 1,
                           int xcd = 0)
{
    if constexpr ((long)NZ * PW * (HALF ? sizeof(T) : sizeof(cpx<T>)) > 65536 || (NZ / R1) * PW > 1024) {
        return hipErrorInvalidValue;   // LDS exchange tile above 64 KiB: not instantiated
    } else {
        constexpr int NW = NZ / R1;
        const int nxy = g.nx * g.ny;
        if (nxy % PW) return hipErrorInvalidValue;
        ZGeom geo{nxy, nxy / PW, (xcd && ntraj % 8 == 0) ? 1 : 0};
        dim3 grid((unsigned)((long)ntraj * geo.tiles_per_traj), (unsigned)ncomp_grid), block(NW * PW);
        hipLaunchKernelGGL((z_kernel<T, NZ, R1, PW, SIGN_F, RT, Op, NT, HALF>), grid, block, 0, s, op, geo,
                           (const cpx<T> *)g.tw);
        return hipGetLastError();
    }
}

template <typename T, int N, class Op, bool NT = false, int MINW = 1>
inline hipError_t launch_xy(const DevGroup &g, const Op &op, int ntraj, double *partial, hipStream_t s,
                            int xcd = 0)
{
    constexpr int PPW = 64 / N;
    XyGeom geo{g.nz, (long)ntraj * g.nz, (xcd && ntraj % 8 == 0) ? 1 : 0};
    if (geo.n_planes % PPW) return hipErrorInvalidValue;
    dim3 grid((unsigned)(geo.n_planes / PPW)), block(64);
    hipLaunchKernelGGL((xy_kernel<T, N, Op, NT, MINW>), grid, block, 0, s, op, geo, partial);
    return hipGetLastError();
}

} // namespace bec
