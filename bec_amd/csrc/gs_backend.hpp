// Ground-state kernels: interface between gs.cpp and kernels_gs.hip.
#pragma once

#include "backend.hpp"

namespace bec {

// Per-trajectory scalar state of a batched ground-state solve (device).
struct GsScalars {
    double *s;        // pending renormalisation factor
    double *mu;       // Rayleigh quotient of the current iterate
    double *e_prev;   // previous reference energy integral
    double *res;      // last residual
    double *e_ref;    // last reference energy integral (not per particle)
    int *iters;       // iterations done
    int *done;        // converged flag
    const double *target;  // per-trajectory atom number (nullptr: the record's N0)
    int pair;         // 1: trajectories 2j, 2j+1 are the components of one coupled set
};

// Energy tables: e_xy[r][q] = (kx^2 + ky^2)/2, e_z[r][kz] = kz^2/2 (double)
struct GsTables {
    const double *exy;
    const double *ez;
};

// ---- split step (calculateSteadyState, wigner.cu:84-175)
// z-pass with the imaginary-time half step (and the pending scale); when
// kcopy != nullptr it also stores the normalised k-space values.
hipError_t gs_z_imag(const DevGroup &g, void *psi, const GsScalars &sc, void *kcopy, hipStream_t s);
hipError_t gs_xy_nonlinear_imag(const DevGroup &g, void *psi, const GsScalars &sc, hipStream_t s);
// partials: {sum n, sum n V, sum n^2, sum Re(psi_x psi_k k^2/2)} per workgroup
hipError_t gs_xy_measure_ref(const DevGroup &g, const void *psi, const void *kcopy, const GsScalars &sc,
                             const GsTables &t, double *partial, hipStream_t s);
// reference rule: s = sqrt(N0/N), E = integral of the renormalised state,
// done when |(E - E_prev)/E| < tol
hipError_t gs_scalar_ref(const DevGroup &g, const double *partial, const GsScalars &sc, double tol,
                         int use_rule, hipStream_t s);

// The residual test of scalar_residual_kernel, evaluated inside the next pass
// (every workgroup from the partials; the first records res and done).
struct GsResidualTest {
    const double *ph;   // Hamiltonian-pass partials
    const double *pr;   // residual-pass partials
    double tol;
    int check;
};

// ---- gradient flow / residual machinery (x-space iterate psi)
// pn != nullptr: first renormalise with the scale of the previous update's
// partials (the work of gs_scalar_norm, counting the iteration)
hipError_t gs_xy_forward(const DevGroup &g, const void *psi, void *w, const GsScalars &sc, hipStream_t s,
                         const double *pn = nullptr);
// variant 0: multiply by k^2/2 / nz; 1: by 1 / ((alpha + k^2/2) nz), alpha
// taken per record from alpha_rec when it is not nullptr; test != nullptr:
// run the residual test first and skip the trajectories it stops
hipError_t gs_z_diag(const DevGroup &g, void *w, const GsScalars &sc, const GsTables &t, int variant,
                     double alpha, const double *alpha_rec, hipStream_t s, const GsResidualTest *test = nullptr);
// h = K psi + (V + g11 |psi|^2) psi; partials {Re<psi,h>, <psi,psi>, <psi,K psi>, sum V n, sum n^2}
hipError_t gs_xy_hamiltonian(const DevGroup &g, const void *psi, void *w, const GsScalars &sc,
                             double *partial, hipStream_t s);
hipError_t gs_scalar_mu(const DevGroup &g, const double *partial, const GsScalars &sc, hipStream_t s);
// r = h - mu psi; partial {|r|^2}; w <- F_xy(r); ph != nullptr: mu from the
// Hamiltonian-pass partials first (the work of gs_scalar_mu, recorded in sc.mu)
hipError_t gs_xy_residual(const DevGroup &g, const void *psi, void *w, const GsScalars &sc, double *partial,
                          hipStream_t s, const double *ph = nullptr);
hipError_t gs_scalar_residual(const DevGroup &g, const double *partial_h, const double *partial_r,
                              const GsScalars &sc, double tol, int check, hipStream_t s);
// psi_new = s psi - tau d + beta (s psi - prev); prev <- s psi; partial {|psi_new|^2}
hipError_t gs_xy_update(const DevGroup &g, void *psi, void *prev, const void *w, const GsScalars &sc,
                        double tau, double beta, double *partial, hipStream_t s);
hipError_t gs_scalar_norm(const DevGroup &g, const double *partial, const GsScalars &sc, hipStream_t s);
// s <- sqrt(N0 / (dV sum |psi|^2)) of an x-space state (no iteration count)
hipError_t gs_norm_only(const DevGroup &g, const void *psi, const GsScalars &sc, double *partial, hipStream_t s);
// psi <- s psi (x-space), s <- 1
hipError_t gs_apply_scale(const DevGroup &g, void *psi, const GsScalars &sc, hipStream_t s);
// Thomas-Fermi start (fillWithTFGroundState, wigner_kernel.cu:116-125), per record mu
hipError_t gs_tf_start(const DevGroup &g, void *psi, const double *mu_tf, hipStream_t s);
// Two-component start from the surrogate network's output net [sets][cells][2]
// (channels s and w d): a = ((s + d) r) u_j, b = ((s - d) r) u_j with
// d = net1 / w and r = 1/sqrt(2), real, into trajectories 2j and 2j + 1 (FP64
// arithmetic in the order of the host conversion, so both give the same bits)
hipError_t gs_net_start(const DevGroup &g, void *psi, const float *net, const double *unit, double spin_weight,
                        double r, hipStream_t s);

} // namespace bec
