// Internal interface between the host orchestration (bec_api.cpp) and the
// kernel launchers (kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "common.hpp"

namespace bec {

// One group of trajectories sharing a lattice shape.  Trajectory `traj` uses
// parameter record traj >> rec_shift.
struct DevGroup {
    int prec;            // 0: complex64 state, 1: complex128 state
    int nx, ny, nz;
    int ntraj;
    int rec_shift;
    int n_records;
    void *a, *b;         // state arrays, ntraj * cells complex each
    TrajParams *params;  // device records [n_records]
    void *kxy;           // kinetic tables [KT_COUNT][n_records][nx*ny] complex
    void *kz;            // [KT_COUNT][n_records][nz] complex
    void *tw;            // z twiddles exp(+2 pi i j / nz), j < nz, complex
    // launch schedule of the two hot passes (chosen by the autotuner; 0 = default)
    int pw;              // pencils per z-pass workgroup: 16, 32, 64
    int z_r1;            // z values per thread in registers (radix of the first z stage)
    int z_split;         // 1: one component per z-pass workgroup (grid.y = 2)
    int nt;              // 1: non-temporal loads/stores in both passes
    int xcd;             // 1: XCD-aware workgroup remap (needs ntraj % 8 == 0)
    int xy_minw;         // xy pass: minimum waves per SIMD for register allocation (1 or 5)
    int z_half;          // 1: z-pass LDS exchange in two real-valued rounds (half the LDS)
    int xy_split;        // 1: real-time xy pass with one wavefront per component (FP32, itmax 1, no noise)
    int itmax1;          // every record has itmax == 1: straight-line nonlinear step
    // loss-induced noise (bec_set_loss_noise): the device step counter read by
    // the nonlinear pass and advanced after it; null when the noise is off
    unsigned *noise_step;
    uint32_t noise_k0, noise_k1;   // Philox key
    int group_index;               // keeps the groups' noise streams apart
};

// One schedule candidate of the hot passes.
struct Schedule {
    int pw, z_r1, z_split, nt, xcd, xy_minw, z_half, xy_split;
};

enum { KT_RT_HALF = 0, KT_RT_FULL = 1, KT_IM_HALF = 2, KT_COUNT = 3 };

bool shape_supported(int nx, int ny, int nz);
size_t elem_size(int prec);                       // bytes of one complex value
long xy_workgroups(const DevGroup &g, int ntraj); // partial-sum slots of an xy pass
int xy_wg_per_traj(const DevGroup &g);
int default_pw(int prec, int nx, int nz);
int default_r1(int nz);
// the schedule candidates the hot-pass launchers are instantiated for; unless
// `all`, without the forms that cannot win at this batch (the ones the autotuner
// times and bec_schedule_candidates lists; bec_set_schedule accepts all)
int schedule_candidates(int prec, int nx, int nz, int ntraj, Schedule *out, int max_out, bool all = false);

hipError_t run_z_kinetic(const DevGroup &g, int variant, hipStream_t s);
hipError_t run_xy_nonlinear(const DevGroup &g, hipStream_t s);
hipError_t run_xy_sample(const DevGroup &g, double *partial, hipStream_t s);
// per-cell ensemble statistics of x-space states ax, bx: out [cells][6]
hipError_t run_averages(const DevGroup &g, const void *ax, const void *bx, double *out, hipStream_t s);
hipError_t run_reduce_partials(const double *partial, int per_traj, int ntraj, int nsum,
                               double *out, hipStream_t s);
// 2D transform of every plane of `ntraj` trajectories (in -> out), times scale
hipError_t run_xy_fft(const DevGroup &g, const void *in, void *out, int ntraj, int sign,
                      double scale, hipStream_t s);
// transform along z of every pencil (in -> out), times scale
hipError_t run_z_fft(const DevGroup &g, const void *in, void *out, int ntraj, int sign,
                     double scale, hipStream_t s);
// x-space a, b = psi_gs + Va eta_a / sqrt(dV), Vb eta_b / sqrt(dV)
hipError_t run_noise_init(const DevGroup &g, const void *gs, uint64_t seed, double Va, double Vb,
                          double inv_sqrt_dv, hipStream_t s);
hipError_t run_pulse(const DevGroup &g, int kind, hipStream_t s);
// (a, b) <- U (a, b) on every cell; u = {u00, u01, u10, u11} as (re, im) pairs
hipError_t run_rotation(const DevGroup &g, const double u[8], hipStream_t s);
// per-cell ensemble sums of vacuum-subtracted densities after halfPiRotate(alpha):
// out[cell*2 + c], x-space
hipError_t run_cell_density_sums(const DevGroup &g, double alpha, double vac, double *out,
                                 hipStream_t s);
// doubles of the projection workspace of a group: the per-cell sums [cells][2],
// then xy [2][ny][nx], zy [2][ny][nz], axial [2][nz]
long projection_doubles(const DevGroup &g);
// the density projections after halfPiRotate(alpha) into buf (projection_doubles(g) doubles)
hipError_t run_projections(const DevGroup &g, double alpha, double vac, double *buf, hipStream_t s);
// plain streaming copy (HBM counter calibration)
hipError_t run_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
hipError_t run_convert(int prec_from, const void *in, int prec_to, void *out, long n, hipStream_t s);

} // namespace bec
