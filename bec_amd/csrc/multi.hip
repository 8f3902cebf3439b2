// One launch per hot pass over every group of a context: the uneven grids and
// batches of one run (several lattices and trajectory counts) stepped
// together, replacing the per-frame launches of calculateEvolution
// (wigner.cu:340-411), one per grid.
//
// The launch arguments carry a table of the groups' tile ranges.  The grid is
// the number of workgroups the GPU keeps resident, and each workgroup strides
// through the table's tiles (group 0's first): no workgroup waits on another,
// and every one exits when the table is done.  The z pass of every lattice
// runs 512-thread workgroups (R1 = z_r1<NZ>, PW = 512 / (NZ / R1) pencils,
// the LDS exchange in two real rounds: 32 KiB for every NZ); the xy pass runs
// 64-thread workgroups.  The tile bodies are those of the per-group kernels,
// so both launch paths compute the same values.  FP32 states only.
#include <algorithm>

#include "launch.hpp"
#include "multi.hpp"

namespace bec {

namespace {

constexpr int ZMT = 512;   // threads of a z workgroup

template <int NZ> struct ZShape {
    static constexpr int R1 = z_r1<NZ>();
    static constexpr int NW = NZ / R1;
    static constexpr int PW = ZMT / NW;
};
constexpr int Z_LDS_SLOTS = 8192;   // NZ * PW for NZ = 64 (4096), 128 and 256 (8192)
static_assert(64 * ZShape<64>::PW <= Z_LDS_SLOTS && 128 * ZShape<128>::PW <= Z_LDS_SLOTS &&
                  256 * ZShape<256>::PW <= Z_LDS_SLOTS,
              "z exchange tile");

int multi_pw(int nz) { return ZMT / (nz / (nz == 64 ? 8 : 16)); }

template <int NZ>
__device__ __forceinline__ void z_group_tile(const MultiZGroup &G, int variant, long blk, int cb, float *lds,
                                             int tid)
{
    using S = ZShape<NZ>;
    using Op = OpKinetic<float, NZ, S::R1, 1>;
    Op op;
    op.a = (cpx<float> *)G.a;
    op.b = (cpx<float> *)G.b;
    op.kxy = (const cpx<float> *)G.kxy + (long)variant * G.kxy_stride;
    op.kz = (const cpx<float> *)G.kz + (long)variant * G.kz_stride;
    op.nxy = G.nxy;
    op.rec_shift = G.rec_shift;
    const ZGeom geo{G.nxy, G.nxy / S::PW, 0};
    z_tile<float, NZ, S::R1, S::PW, +1, true, Op, true, true>(op, geo, (const cpx<float> *)G.tw, blk, cb, lds, tid);
}

// At most 128 VGPRs (4 waves per SIMD), so that a CU holds two of the 512-thread
// workgroups; above that it holds one and the pass slows by about a tenth.
__global__ void __launch_bounds__(ZMT) __attribute__((amdgpu_waves_per_eu(4, 8)))
z_multi_kernel(MultiZArgs args, int variant)
{
    __shared__ float lds[Z_LDS_SLOTS];
    for (long t = blockIdx.x; t < args.total_tiles; t += gridDim.x) {
        int gi = 0;
        while (gi + 1 < args.n_groups && t >= args.g[gi + 1].tile0) ++gi;
        const MultiZGroup &G = args.g[gi];
        const long local = t - G.tile0;   // tile of the group * 2 + component
        const int cb = (int)(local & 1);
        const long blk = local >> 1;
        const int tid = opaque_index(threadIdx.x);
        switch (G.nz) {
        case 64: z_group_tile<64>(G, variant, blk, cb, lds, tid); break;
        case 128: z_group_tile<128>(G, variant, blk, cb, lds, tid); break;
        default: z_group_tile<256>(G, variant, blk, cb, lds, tid); break;
        }
        __syncthreads();   // the next tile rewrites the exchange slots
    }
}

template <int N, bool IT1, bool NOISE>
__device__ __forceinline__ void xy_group_tile(const MultiXyGroup &G, long blk, cpx<float> *lds, int lane)
{
    using Op = OpNonlinRT<float, N, IT1, NOISE>;
    Op op;
    op.a = (cpx<float> *)G.a;
    op.b = (cpx<float> *)G.b;
    op.P = G.params;
    op.rec_shift = G.rec_shift;
    op.step = G.step;
    op.k0 = G.k0;
    op.k1 = G.k1;
    op.group = G.group;
    const XyGeom geo{G.nz, (long)G.ntraj * G.nz, 0};
    xy_tile<float, N, Op, true>(op, geo, nullptr, blk, lds, lane);
}

// NMAX: the largest plane edge among the groups (sizes the LDS)
template <int NMAX, bool IT1, bool NOISE>
__global__ void __launch_bounds__(64) xy_multi_kernel(MultiXyArgs args)
{
    __shared__ cpx<float> lds[(64 / NMAX) * NMAX * (NMAX + 1)];
    for (long t = blockIdx.x; t < args.total_tiles; t += gridDim.x) {
        int gi = 0;
        while (gi + 1 < args.n_groups && t >= args.g[gi + 1].tile0) ++gi;
        const MultiXyGroup &G = args.g[gi];
        const long blk = t - G.tile0;
        const int lane = opaque_index(threadIdx.x);
        if (G.n == 16) xy_group_tile<16, IT1, NOISE>(G, blk, lds, lane);
        else if constexpr (NMAX == 32) xy_group_tile<32, IT1, NOISE>(G, blk, lds, lane);
        __syncthreads();   // the next tile rewrites the transpose tiles
    }
}

// advances every group's loss-noise step counter after the xy pass read it
__global__ void noise_tick_multi_kernel(MultiXyArgs args)
{
    const int i = threadIdx.x;
    if (i < args.n_groups && args.g[i].step) args.g[i].step[0] = args.g[i].step[0] + 1u;
}

// workgroups of `threads` threads the GPU keeps resident at once (the
// occupancy of the kernel on the current device times its CUs)
template <class K> int resident_grid(K kernel, int threads)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return cus * per_cu;
}

} // namespace

bool multi_plan(const DevGroup *const *groups, int n, MultiPlan *plan)
{
    if (n < 1 || n > MULTI_MAX_GROUPS || !plan) return false;
    const DevGroup &g0 = *groups[0];
    MultiPlan P{};
    P.n_groups = n;
    P.nmax = 16;
    P.itmax1 = g0.itmax1;
    P.noise = g0.noise_step != nullptr;
    long zt = 0, xt = 0;
    for (int i = 0; i < n; ++i) {
        const DevGroup &d = *groups[i];
        if (d.prec != 0 || !shape_supported(d.nx, d.ny, d.nz)) return false;
        if (d.itmax1 != P.itmax1 || (d.noise_step != nullptr) != P.noise) return false;
        const int nxy = d.nx * d.ny;
        const int pw = multi_pw(d.nz);
        if (nxy % pw || d.nz % (64 / d.nx)) return false;
        MultiZGroup &Z = P.z.g[i];
        Z.a = d.a;
        Z.b = d.b;
        Z.kxy = d.kxy;
        Z.kz = d.kz;
        Z.kxy_stride = (long)d.n_records * nxy;
        Z.kz_stride = (long)d.n_records * d.nz;
        Z.tw = d.tw;
        Z.tile0 = zt;
        Z.nxy = nxy;
        Z.nz = d.nz;
        Z.rec_shift = d.rec_shift;
        zt += 2L * d.ntraj * (nxy / pw);
        MultiXyGroup &X = P.xy.g[i];
        X.a = d.a;
        X.b = d.b;
        X.params = d.params;
        X.step = d.noise_step;
        X.tile0 = xt;
        X.k0 = d.noise_k0;
        X.k1 = d.noise_k1;
        X.group = (uint32_t)d.group_index;
        X.n = d.nx;
        X.nz = d.nz;
        X.ntraj = d.ntraj;
        X.rec_shift = d.rec_shift;
        xt += (long)d.ntraj * d.nz / (64 / d.nx);
        P.nmax = std::max(P.nmax, d.nx);
    }
    P.z.n_groups = P.xy.n_groups = n;
    P.z.total_tiles = zt;
    P.xy.total_tiles = xt;
    if (zt > 0x7fffffffL || xt > 0x7fffffffL) return false;
    P.z_grid = (int)std::min<long>(zt, resident_grid(z_multi_kernel, ZMT));
    with_bool(P.itmax1, [&](auto IT1t) {
        return with_bool(P.noise, [&](auto NOISEt) {
            constexpr bool IT1 = decltype(IT1t)::value, NOISE = decltype(NOISEt)::value;
            const int r = P.nmax == 16 ? resident_grid(xy_multi_kernel<16, IT1, NOISE>, 64)
                                       : resident_grid(xy_multi_kernel<32, IT1, NOISE>, 64);
            P.xy_grid = (int)std::min<long>(xt, r);
            return hipSuccess;
        });
    });
    *plan = P;
    return true;
}

hipError_t run_z_multi(const MultiPlan &P, int variant, hipStream_t s)
{
    if (P.z_grid < 1 || variant < 0 || variant >= KT_COUNT) return hipErrorInvalidValue;
    hipLaunchKernelGGL(z_multi_kernel, dim3(P.z_grid), dim3(ZMT), 0, s, P.z, variant);
    return hipGetLastError();
}

hipError_t run_xy_multi(const MultiPlan &P, hipStream_t s)
{
    if (P.xy_grid < 1) return hipErrorInvalidValue;
    const dim3 grid(P.xy_grid), block(64);
    auto go = [&](auto IT1t, auto NOISEt) {
        constexpr bool IT1 = decltype(IT1t)::value, NOISE = decltype(NOISEt)::value;
        if (P.nmax == 16) hipLaunchKernelGGL((xy_multi_kernel<16, IT1, NOISE>), grid, block, 0, s, P.xy);
        else hipLaunchKernelGGL((xy_multi_kernel<32, IT1, NOISE>), grid, block, 0, s, P.xy);
        return hipSuccess;
    };
    with_bool(P.itmax1, [&](auto IT1t) { return with_bool(P.noise, [&](auto Nt) { return go(IT1t, Nt); }); });
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && P.noise) {
        hipLaunchKernelGGL(noise_tick_multi_kernel, dim3(1), dim3(64), 0, s, P.xy);
        e = hipGetLastError();
    }
    return e;
}

} // namespace bec
