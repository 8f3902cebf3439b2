/*
 * FP64 CPU oracle for the split-step Wigner GPE propagator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (programs/bec_wigner_3d.cuda, programs/bec_wigner_2d.cpp/wigner.cpp
 * and wigner_1d.cpp of fjarri-attic/bec), plus the loss-induced noise its
 * report derives but its code does not integrate.  It is the checker that
 * the HIP path is compared against and the timed CPU baseline.  Nothing in
 * the product library links to it.  Every function is pinned by closed forms,
 * known answers or exact quantum dynamics in tests/test_oracle_*.py; none is
 * parity-unpinned.
 *
 * Conventions (identical to the reference):
 *   - cell index = z*(nx*ny) + y*nx + x, x fastest; trajectories outermost
 *     (programs/bec_wigner_3d.cuda/wigner_kernel.cu:99-113, defines.h:160)
 *   - FFT sign -1 == CUFFT_FORWARD (exp(-i...)) is used k->x, sign +1 ==
 *     CUFFT_INVERSE followed by 1/cells is used x->k
 *     (programs/bec_wigner_3d.cuda/wigner.cu:106-107,121,129-130)
 */
#ifndef BEC_ORACLE_H
#define BEC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_c;

/* ModelParameters, SI units (programs/bec_wigner_3d.cuda/defines.h:25-54) */
typedef struct {
    double a11, a12, a22;      /* Bohr radii */
    double m;                  /* kg */
    double fx, fy, fz;         /* Hz */
    int N;
    double detuning;           /* Hz */
    double gamma111, gamma12, gamma22; /* cm^6/s, cm^3/s, cm^3/s */
    double Va, Vb;
    int nvx, nvy, nvz;
    int itmax;
    double dt_steady, t_equilib, dt_evo; /* s */
    int ne;
} orc_model;

/* CalculationParameters, natural units (programs/bec_wigner_3d.cuda/defines.h:57-112) */
typedef struct {
    int N;
    double g11, g12, g22, mu;
    int cells;
    double Va, Vb, V;
    double lambda, l_rho, t_rho;
    double detuning, l111, l12, l22;
    int nvx, nvy, nvz;
    double xmax, ymax, zmax, dx, dy, dz, dkx, dky, dkz;
    int itmax;
    double dt_steady, t_equilib, dt_evo;
    int ne;
} orc_calc;

/* fillDerivedParameters (programs/bec_wigner_3d.cuda/wigner.cpp:119-176) */
void orc_derive(const orc_model *m, orc_calc *c);

/* k^2/2 table (wigner_kernel.cu:15-55) and trap potential (wigner_kernel.cu:99-113) */
void orc_ktable(const orc_calc *c, double *k2half);
double orc_potential(const orc_calc *c, int cell);

/* Unnormalised batched C2C FFTs.  sign = -1: exp(-2 pi i jk/n), +1: exp(+...). */
void orc_fft1d(orc_c *data, int n, int sign);
void orc_fft3d(orc_c *data, int nz, int ny, int nx, int batch, int sign);
void orc_fft_axes(orc_c *data, int nz, int ny, int nx, int batch, int sign, int axes_mask);
void orc_dft3d_bruteforce(const orc_c *in, orc_c *out, int nz, int ny, int nx, int sign);

/* ---- ground state (calculateSteadyState, wigner.cu:84-175) ---- */
typedef struct {
    int iterations;
    int converged;
    double E_ref;      /* the reference's "energy" integral (quirk Q3), per particle */
    double mu_ref;     /* the reference's "mu" integral, per particle */
    double E;          /* true <psi|H|psi>/N */
    double mu;         /* true <psi|H|psi>/<psi|psi> + interaction term */
    double N;          /* dV * sum |psi|^2 */
    double residual;   /* ||H psi - mu psi|| / (|mu| ||psi||) */
} orc_gs_report;

#define ORC_GS_RULE_REFERENCE 0   /* stop at |(E_new-E)/E_new| < tol (wigner.cu:155) */
#define ORC_GS_RULE_RESIDUAL  1   /* stop at residual < tol, checked every check_every iterations */

int orc_ground_state(const orc_calc *c, const orc_c *guess_x, int rule, double tol,
                     int max_iter, int check_every, orc_c *out_x, orc_gs_report *rep);
void orc_tf_state(const orc_calc *c, orc_c *out_x);
double orc_state_integral_ref(const orc_calc *c, const orc_c *psi_x, int energy);
void orc_energy_mu(const orc_calc *c, const orc_c *psi_x, double *E, double *mu, double *N);
double orc_residual(const orc_calc *c, const orc_c *psi_x, double *mu_out);
void orc_gs_iterate(const orc_calc *c, orc_c *a_k, int n_iter, double *E_ref_out);

/* ---- two-component ground state (docs/wigner_simulation.report/report.tex:200-211) ----
 * mu_a a = (K + V + g11|a|^2 + g12|b|^2) a, mu_b b = (K + V + g22|b|^2 + g12|a|^2) b
 * with dV sum |a|^2 = frac_a N and dV sum |b|^2 = (1 - frac_a) N.  The constant
 * hyperfine offset of |2> shifts mu_b only and is left out. */
typedef struct {
    int iterations;
    int converged;
    double E;          /* total energy per particle */
    double mu_a, mu_b; /* per-component Rayleigh quotients */
    double N_a, N_b;
    double residual;   /* max over components of ||H_c c - mu_c c|| / (|mu_c| ||c||) */
} orc_gs2_report;

/* H_a a and H_b b (x-space, one pair) */
void orc_apply_h2(const orc_calc *c, const orc_c *a, const orc_c *b, orc_c *ha, orc_c *hb);
double orc_residual2(const orc_calc *c, const orc_c *a, const orc_c *b, double *mu_a, double *mu_b);
double orc_energy2(const orc_calc *c, const orc_c *a, const orc_c *b);
/* Thomas-Fermi start: the one-component TF shape of N atoms (g11), scaled to
 * frac_a N and (1 - frac_a) N */
void orc_tf_state2(const orc_calc *c, double frac_a, orc_c *a, orc_c *b);
/* Preconditioned imaginary-time gradient flow with momentum:
 *   c <- c - tau P (H_c - mu_c) c + beta (c - c_prev), then each component
 *   renormalised, P = (alpha + k^2/2)^-1; stops at residual < tol (checked
 *   every iteration).  guesses may be NULL (TF start). */
int orc_ground_state2(const orc_calc *c, double frac_a, const orc_c *guess_a, const orc_c *guess_b,
                      double tol, int max_iter, double alpha, double tau, double beta,
                      orc_c *out_a, orc_c *out_b, orc_gs2_report *rep);

/* ---- noise + ensembles ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_noise_cell(uint64_t seed, uint32_t traj, uint32_t cell, uint32_t stream, double out[4]);
void orc_init_ensembles(const orc_calc *c, const orc_c *gs_x, uint64_t seed, orc_c *a_x, orc_c *b_x);
int orc_equilibration_steps(const orc_calc *c);
void orc_to_kspace(const orc_calc *c, orc_c *data, int ne);
void orc_to_xspace(const orc_calc *c, orc_c *data, int ne);
void orc_init_evolution(const orc_calc *c, const orc_c *gs_x, uint64_t seed, int equilibrate,
                        orc_c *a_k, orc_c *b_k);

/* ---- real-time evolution (propagate, wigner.cu:178-199) ---- */
void orc_propagate(const orc_calc *c, orc_c *a_k, orc_c *b_k, int ne, double dt, int n_steps);
void orc_kinetic_rt(const orc_calc *c, const double *k2half, orc_c *a, orc_c *b, int ne, double dt);
void orc_nonlinear_rt(const orc_calc *c, orc_c *a_x, orc_c *b_x, int ne, double dt);
void orc_half_pi_pulse(orc_c *a, orc_c *b, long n);
void orc_pi_pulse(orc_c *a, orc_c *b, long n);

/* ---- observables, computed the reference way: rotate per cell, then sum ---- */
/* halfPiRotate + density (wigner_kernel.cu:283-286,344-360); returns vacuum-subtracted sums */
void orc_rotate_sums(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne, double alpha,
                     double *N1, double *N2);
/* per-trajectory plane sums: [ne][4] = sum|a|^2, sum|b|^2, Re sum conj(a) b, Im sum conj(a) b */
void orc_traj_sums(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne, double *sums);
/* per-cell ensemble sums of vacuum-subtracted densities after halfPiRotate(alpha) */
void orc_cell_density_sums(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne,
                           double alpha, double *dens_a, double *dens_b);
/* getVisibility: max over alpha = 0, 0.5, ... < 2 pi of |N1-N2|/(N1+N2), with N = |sum| */
double orc_visibility_grid(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne);

/* ---- 2D CPU program (programs/bec_wigner_2d.cpp/wigner.cpp) ---- */
typedef struct {
    int nvx, nvy;
    double xmax, ymax;
    double g11, g12, g22, E, mu;
    double Va, Vb;
    double dtG; int ntG; int itmax;
    double dtW; int ntW;          /* equilibration: ntW steps including the pulse step */
    int n_meas;                   /* samples after the pulse (reference: 1) */
    int ne;
    double n0;
    uint64_t seed;
    double noise_dev;             /* reference: 1/sqrt(dx*dy) */
    int gs_rule;                  /* 0: fixed ntG-1 steps (reference); 1: to residual gs_tol;
                                     2: gradient flow renormalised to n_atoms, to residual gs_tol */
    double gs_tol; int gs_max_iter;
    double n_atoms;               /* gs_rule 2: atoms of the ground state */
    double gs_tau, gs_beta;       /* gs_rule 2: step and momentum of the flow */
} orc2d_params;

typedef struct {
    double N0;
    int gs_iterations;
    double gs_residual;
} orc2d_report;

/* m, var: [2][n_meas]; vis: [n_meas]; gs_x, avn: nvx*nvy (may be NULL) */
int orc2d_run(const orc2d_params *p, double *m, double *var, double *vis, orc_c *gs_x,
              double *avn, orc2d_report *rep);
double orc2d_residual(const orc2d_params *p, const orc_c *psi_x, double *mu_out);

/* resonant pulse of area theta and phase phi on n cells (U(pi/2,0), U(pi,0) = the reference pulses) */
void orc_rotation(orc_c *a, orc_c *b, long n, double theta, double phi);

/* calculateAverages (wigner.cu:305-337) on x-space states: {N, norm, var Re a, var Im a, var Re b, var Im b} */
void orc_averages(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne, double out[6]);

/* ---- loss-induced Wigner noise during the evolution (derived in
   docs/wigner_simulation.report/report.tex:624-656, not implemented by the reference) ---- */
/* four standard normal deviates: Philox counter {cell, traj, step, 1 + group} */
void orc_loss_noise_cell(uint64_t seed, uint32_t traj, uint32_t cell, uint32_t step, uint32_t group, double out[4]);
/* orc_nonlinear_rt plus the Ito increment B dW of the loss diffusion */
void orc_nonlinear_rt_noise(const orc_calc *c, orc_c *a_x, orc_c *b_x, int ne, double dt, uint64_t seed,
                            uint32_t step, uint32_t group);
void orc_propagate_noise(const orc_calc *c, orc_c *a, orc_c *b, int ne, double dt, int n_steps, uint64_t seed,
                         uint32_t step0, uint32_t group);

/* ---- 1D CPU program (programs/bec_wigner_2d.cpp/wigner_1d.cpp; the same scheme as the Scilab
   prototype programs/bec_wigner_1d.scilab/WignerSplitFFT.sce) ---- */
typedef struct {
    double zmax, dz;              /* nv = 1 + 2 (int)(zmax / dz) cells at z = -zmax + dz j */
    double g11, g12, g22, E, mu;
    double Va, Vb;
    double dt; int ntG; int itmax;
    int ntW;                      /* ntW - 1 equilibration steps, then ntW steps each sampled */
    int ne;
    double n0;
    uint64_t seed;
    double noise_dev;             /* reference: sqrt(1/dz)/2 per quadrature */
} orc1d_params;

typedef struct {
    double N0;
    int nv;
    double gs_residual;
    double gs_mu;
} orc1d_report;

/* DFT of any length by its definition: d[k] = sum_j d[j] exp(sign 2 pi i j k / n) */
void orc_dft1d(orc_c *d, int n, int sign);
int orc1d_nv(const orc1d_params *p);
/* m, var: [2][ntW]; vis: [ntW]; gs_x, avn: nv; ntot: ntG (each may be NULL except m, var) */
int orc1d_run(const orc1d_params *p, double *m, double *var, double *vis, orc_c *gs_x, double *avn,
              double *ntot, orc1d_report *rep);
double orc1d_residual(const orc1d_params *p, const orc_c *psi_x, double *mu_out);

int orc_num_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
