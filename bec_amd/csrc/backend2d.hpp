// The 2D program on the GPU (kernels2d.hip): launchers used by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/bec_api.h"

namespace bec {

using W2dParams = bec_2d_params;

constexpr int W2D_MAX_N = 64;   // lattice side limit: both components resident in LDS

// ground state (x-space) into gs_out[cells]; rep_out = {N0, iterations, residual, mu};
// pv [cells]: the gradient flow's previous iterate (gs_rule 2; may be null otherwise)
hipError_t run_w2d_ground_state(const W2dParams &p, double2 *gs_out, double2 *pv, double *rep_out, hipStream_t s);
// one workgroup per trajectory; meas [ne][n_meas][3] = {N1, N2, visibility}; avn [ne][cells] or null
hipError_t run_w2d_ensemble(const W2dParams &p, const double2 *gs, double *meas, double *avn, hipStream_t s);

} // namespace bec
