// The 1D program on the GPU (kernels1d.hip): launchers used by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/bec_api.h"

namespace bec {

using W1dParams = bec_1d_params;

constexpr int W1D_MAX_NV = 512;   // eight cells per lane of one wavefront

// ground state (x-space) into gs_out[nv]; ntot[ntG] (or null); rep_out = {N0, residual, mu}
hipError_t run_w1d_ground_state(const W1dParams &p, int nv, double2 *gs_out, double *ntot, double *rep_out,
                                hipStream_t s);
// one wavefront per trajectory: meas [ne][ntW][3] = {N1, N2, visibility}; avn [ne][nv]
hipError_t run_w1d_ensemble(const W1dParams &p, int nv, const double2 *gs, double *meas, double *avn,
                            hipStream_t s);
// sums over trajectories in trajectory order: sums [5][ntW] = {N1, N2, N1^2, N2^2, vis}; avn_sum [nv]
hipError_t run_w1d_reduce(const W1dParams &p, int nv, const double *meas, const double *avn, double *sums,
                          double *avn_sum, hipStream_t s);

} // namespace bec
