/*
 * bec_api.h -- C ABI of the MI355X (gfx950) split-step Wigner GPE propagator.
 *
 * Plain C: opaque handles, int status codes (never exit()), caller-owned host
 * buffers, device memory owned by the context.  All GPU work is issued on the
 * stream given at creation.
 *
 * Reference interface replaced (paths relative to the fjarri-attic/bec tree):
 *   programs/bec_wigner_3d.cuda/wigner.cpp:33-40   host functions of the solver
 *   programs/bec_wigner_3d.cuda/batchfft.h:65-108  batched FFT sub-API
 *   programs/bec_wigner_3d.cuda/defines.h:25-210   parameter structs, state
 *
 * Host-side complex arrays are interleaved (re, im) doubles in the reference
 * layout: trajectory outermost, then z, y, x with x fastest
 * (wigner_kernel.cu:99-113, defines.h:160).
 */
#ifndef BEC_API_H
#define BEC_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
enum {
    BEC_OK = 0,
    BEC_ERR_INVALID = 1,        /* bad argument / NULL pointer */
    BEC_ERR_SHAPE = 2,          /* unsupported grid shape or batch */
    BEC_ERR_ALLOC = 3,          /* device or host allocation failed */
    BEC_ERR_HIP = 4,            /* a HIP runtime call failed */
    BEC_ERR_NOT_CONVERGED = 5,  /* ground state hit max_iter before tol */
    BEC_ERR_NONFINITE = 6,      /* NaN/Inf detected by an FP64 norm check */
    BEC_ERR_STATE = 7,          /* call out of order (e.g. step before init) */
    BEC_ERR_UNSUPPORTED = 8     /* option not available in this build */
};

enum { BEC_FP32 = 0, BEC_FP64 = 1 };          /* state / FFT / nonlinear precision */
/* Representation for get/set state: x-space, k-space, or the stored hybrid
 * H(z, ky, kx) -- the state transformed along x and y only, psi_x = F_xy(-) H,
 * psi_k = F_z(+) H / nz -- whose round trip is exact (a bit-exact checkpoint). */
enum { BEC_SPACE_X = 0, BEC_SPACE_K = 1, BEC_SPACE_HYBRID = 2 };
enum { BEC_PULSE_HALF_PI = 0, BEC_PULSE_PI = 1 };
enum { BEC_FFT_FORWARD = -1, BEC_FFT_INVERSE = 1 };  /* = CUFFT_FORWARD / CUFFT_INVERSE */

/* Ground-state methods */
enum {
    BEC_GS_SPLIT_STEP = 0,  /* imaginary-time split step of wigner.cu:84-175 */
    BEC_GS_GRADIENT = 1     /* preconditioned imaginary-time gradient flow (fixed point: H psi = mu psi) */
};
/* Ground-state stopping rules */
enum {
    BEC_GS_RULE_REFERENCE = 0,  /* |(E_new - E)/E_new| < tol on the reference integral (wigner.cu:155) */
    BEC_GS_RULE_RESIDUAL = 1    /* ||H psi - mu psi|| / (|mu| ||psi||) < tol, evaluated in FP64 */
};

/* ModelParameters, SI units (defines.h:25-54). */
typedef struct bec_model_params {
    double a11, a12, a22;               /* scattering lengths, Bohr radii */
    double m;                           /* atom mass, kg */
    double fx, fy, fz;                  /* trap frequencies, Hz (fx radial, fz axial) */
    int N;                              /* atom number */
    double detuning;                    /* Hz */
    double gamma111, gamma12, gamma22;  /* loss rates: cm^6/s, cm^3/s, cm^3/s */
    double Va, Vb;                      /* vacuum-noise switches */
    int nvx, nvy, nvz;                  /* lattice (powers of two) */
    int itmax;                          /* midpoint iterations */
    double dt_steady, t_equilib, dt_evo;/* s */
    int ne;                             /* trajectories */
} bec_model_params;

/* CalculationParameters, natural units (defines.h:57-112). */
typedef struct bec_calc_params {
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
} bec_calc_params;

/* fillCalculationParameters defaults (wigner.cpp:43-98). */
int bec_default_model_params(bec_model_params *out);
/* fillDerivedParameters (wigner.cpp:119-176), evaluated in double. */
int bec_fill_derived_params(const bec_model_params *in, bec_calc_params *out);

/* ----------------------------------------------------------- context */
typedef struct bec_ctx bec_ctx;

/* initConstants + initWaveVectors + EvolutionState::init
 * (wigner_kernel.cu:25-55,395-399; defines.h:157-184).  One grid group. */
int bec_create(const bec_calc_params *p, int precision, void *hip_stream, bec_ctx **out);

/* Several grid groups (uneven grids / batches) stepped together.  Each group
 * has its own parameters, lattice and trajectory count. */
int bec_create_groups(const bec_calc_params *groups, int n_groups, int precision,
                      void *hip_stream, bec_ctx **out);

/* EvolutionState::release + releaseWaveVectors (defines.h:186-205, wigner_kernel.cu:57-61). */
int bec_destroy(bec_ctx *ctx);
const char *bec_strerror(int status);
int bec_synchronize(bec_ctx *ctx);
int bec_group_count(const bec_ctx *ctx);

/* ------------------------------------------------------ ground state */
typedef struct bec_gs_options {
    int method;         /* BEC_GS_SPLIT_STEP | BEC_GS_GRADIENT */
    int rule;           /* BEC_GS_RULE_REFERENCE | BEC_GS_RULE_RESIDUAL */
    double tol;         /* stopping tolerance of the rule */
    int max_iter;       /* iteration cap (the reference has none) */
    int check_every;    /* residual evaluation period, iterations */
    int precision;      /* BEC_FP32 | BEC_FP64 for the ground-state FFTs */
    double alpha;       /* gradient flow: preconditioner shift (0: mu_TF) */
    double beta;        /* gradient flow: momentum */
    double tau;         /* gradient flow: step */
} bec_gs_options;

typedef struct bec_gs_report {
    int iterations;
    int converged;
    double E_ref;       /* reference "energy" integral / N (wigner.cu:168-169, quirk Q3) */
    double mu_ref;      /* reference "mu" integral / N (wigner.cu:171-172) */
    double E;           /* <psi|H|psi>/N with the mean-field energy g n^2/2 */
    double mu;          /* chemical potential <psi|H|psi>/<psi|psi> */
    double N;           /* dV sum |psi|^2 */
    double residual;    /* ||H psi - mu psi|| / (|mu| ||psi||), FP64 */
    double elapsed_ms;  /* device time of the solve (HIP events) */
} bec_gs_report;

int bec_default_gs_options(bec_gs_options *out);

/* Step and momentum of the two-component gradient flow (the one-component
 * defaults of bec_default_gs_options are 0.5 and 0.6). */
#define BEC_GS2_DEFAULT_TAU 1.0
#define BEC_GS2_DEFAULT_BETA 0.85

/* calculateSteadyState (wigner.cu:84-175) for group `group`: imaginary-time
 * ground state of component |1>, returned in x-space (cells complex values).
 * guess_x (x-space, may be NULL) replaces the Thomas-Fermi start (K1). */
int bec_calculate_steady_state(bec_ctx *ctx, int group, const double *guess_x,
                               const bec_gs_options *opt, double *out_x, bec_gs_report *rep);

/* Batched ground states: `batch` independent parameter sets that share one
 * lattice shape (nvx, nvy, nvz taken from params[0]).  guesses (x-space,
 * batch*cells complex) may be NULL; out_x receives batch*cells values;
 * reps receives batch reports. */
int bec_ground_states_batched(const bec_calc_params *params, int batch, const double *guesses,
                              const bec_gs_options *opt, void *hip_stream, double *out_x,
                              bec_gs_report *reps);

/* Two-component ground states: the coupled imaginary-time GPEs that the
 * reference describes but does not implement (docs/wigner_simulation.report/
 * report.tex:200-211; its solver, wigner.cu:84-175, relaxes |1> only):
 *   mu_a a = (K + V + g11|a|^2 + g12|b|^2) a,  mu_b b = (K + V + g22|b|^2 + g12|a|^2) b
 * with dV sum|a|^2 = frac_a N and dV sum|b|^2 = (1 - frac_a) N, solved by the
 * preconditioned gradient flow (method BEC_GS_GRADIENT, rule
 * BEC_GS_RULE_RESIDUAL; opt NULL selects them, with the coupled flow's own
 * step and momentum BEC_GS2_DEFAULT_TAU / _BETA).  frac_a may be NULL (1/2 each).
 * guesses and out_x hold batch * 2 * cells complex values, x-space, laid out
 * [set][component][cell]; guesses NULL gives the Thomas-Fermi start (the
 * one-component TF shape of N atoms, scaled to each component's number). */
typedef struct bec_gs2_report {
    int iterations;
    int converged;
    double E;           /* total energy per particle */
    double mu_a, mu_b;  /* per-component chemical potentials */
    double N_a, N_b;
    double residual;    /* max over components of ||H_c c - mu_c c|| / (|mu_c| ||c||) */
    double elapsed_ms;  /* device time of the iterations (HIP events) */
} bec_gs2_report;

int bec_ground_states_two_component(const bec_calc_params *params, int batch, const double *frac_a,
                                    const double *guesses, const bec_gs_options *opt, void *hip_stream,
                                    double *out_x, bec_gs2_report *reps);

/* The same solve started from the learned start's network output where the
 * network left it, in device memory (no host round trip; no counterpart in the
 * reference).  d_net holds batch x cells x 2 floats, the network's output
 * channels s and spin_weight * d of each set; the start is
 * a = ((s + d)/sqrt2) u, b = ((s - d)/sqrt2) u with u = sqrt(mu / g11) of the
 * set, in FP64 in that order (the same bits as building the guesses on the host
 * and passing them to bec_ground_states_two_component).  d_net must be
 * complete on hip_stream when it is given, or on the null stream when
 * hip_stream is NULL. */
int bec_ground_states_two_component_net(const bec_calc_params *params, int batch, const double *frac_a,
                                        const float *d_net, double spin_weight, const bec_gs_options *opt,
                                        void *hip_stream, double *out_x, bec_gs2_report *reps);

/* Residual ||H psi - mu psi||/(|mu| ||psi||) and mu of an x-space state, FP64 on the GPU. */
int bec_state_residual(bec_ctx *ctx, int group, const double *psi_x, double *residual, double *mu);

/* ------------------------------------------------------- the 2D program
 * programs/bec_wigner_2d.cpp/wigner.cpp on the GPU, FP64: calculate_steady_state
 * (:196-275) then main_loop (:278-477) for ne trajectories, one workgroup per
 * trajectory with its state resident in LDS.  Field meanings are the program's
 * file-scope constants (:106-136) plus the extensions of the measurement loop:
 * n_meas samples after the pulse (the program takes 1), gs_rule = 1 to stop
 * the ground state at residual < gs_tol instead of after ntG - 1 steps, and
 * gs_rule = 2 for a ground state of n_atoms atoms: the preconditioned gradient
 * flow of the 3D solver (step gs_tau, momentum gs_beta, preconditioner
 * 1/(mu + k^2)) from the Thomas-Fermi profile, renormalised after every update,
 * to residual < gs_tol (the program's loop fixes mu instead of N).  The
 * noise is Philox (counter {cell, trajectory, 0, 0}, key seed) with deviation
 * noise_dev (the program: 1/sqrt(dx dy)).  Lattices: nvx, nvy powers of two
 * from 4 to 64 (the program's 128 x 128 does not fit a workgroup). */
typedef struct bec_2d_params {
    int nvx, nvy;
    double xmax, ymax;
    double g11, g12, g22, E, mu;
    double Va, Vb;
    double dtG;
    int ntG, itmax;
    double dtW;
    int ntW;          /* equilibration steps, the pulse step included */
    int n_meas;
    int ne;
    double n0;        /* uniform start of the ground state */
    uint64_t seed;
    double noise_dev;
    int gs_rule;
    double gs_tol;
    int gs_max_iter;
    double n_atoms;   /* gs_rule 2: atoms of the ground state */
    double gs_tau;    /* gs_rule 2: step of the flow */
    double gs_beta;   /* gs_rule 2: momentum of the flow */
} bec_2d_params;

/* The 2D program's parameters for the two-component cloud of a 3D parameter
 * set (SI units, bec_model_params): the program's isotropic plane is the
 * trap's radial plane (frequency fx), with the axial direction (fz) frozen in
 * its oscillator ground state.  Units are those of the program's Hamiltonian
 * -lap + r^2/4 (wigner.cpp:154-194): energy hbar w, time 1/w, length
 * sqrt(hbar / (2 m w)), w = 2 pi fx.  Then
 *   g_ij = 4 sqrt(2 pi) a_ij / a_z,  a_z = sqrt(hbar / (m 2 pi fz)),
 *   mu = sqrt(g11 N / (2 pi))  (Thomas-Fermi, radius 2 sqrt(mu)),
 *   xmax = ymax = 1.2 * 2 sqrt(mu),  lattice nvx x nvy,
 *   dtG = w dt_steady, dtW = w dt_evo, E = 2 pi detuning / w,
 *   noise deviation: the 3D one, 1/(2 dx dy) per mode,
 *   ground state gs_rule 2 with n_atoms = N, gs_tol 1e-8, tau 0.5, beta 0.6;
 * one pulse step (ntW = 1) and one sample (n_meas = 1); ne, Va, Vb and the
 * itmax of the model.  The caller sets n_meas and the seed. */
int bec_fill_2d_params(const bec_model_params *m, bec_2d_params *out);

typedef struct bec_2d_report {
    double N0;            /* atoms of the ground state */
    int gs_iterations;
    double gs_residual;   /* ||H psi - mu psi|| / (|mu| ||psi||) of the ground state */
    double gs_mu;
    double elapsed_ms;    /* device time of ground state + ensemble */
} bec_2d_report;

/* m, var: [2][n_meas] ensemble means of N1, N2 and their standard errors
 * (wigner.cpp:431-443); vis: [n_meas] mean visibility; gs_x: [cells] complex
 * ground state (may be NULL); avn: [cells] mean x-space density minus the
 * vacuum term after the last step (may be NULL). */
int bec_run_2d(const bec_2d_params *p, double *m, double *var, double *vis, double *gs_x, double *avn,
               bec_2d_report *rep);

/* ------------------------------------------------------- the 1D program
 * programs/bec_wigner_2d.cpp/wigner_1d.cpp on the GPU, FP64 (the scheme of the
 * Scilab prototype programs/bec_wigner_1d.scilab/WignerSplitFFT.sce):
 * calculate_steady_state (:166-243) then main_loop (:245-445) for ne
 * trajectories, one wavefront per trajectory.  Field meanings are the
 * program's file-scope constants (:93-130): the lattice has
 * nv = 1 + 2 (int)(zmax / dz) cells at z = -zmax + dz j (any nv from 1 to
 * 512, the program's is 41), ntG counts the ground-state loop including its
 * initial pass (ntG - 1 steps), and the ensemble loop takes ntW - 1
 * equilibration steps, the pulse, then ntW steps each followed by a sample.
 * The noise is Philox (counter {cell, trajectory, 0, 0}, key seed) with
 * deviation noise_dev per quadrature (the program: sqrt(1/dz)/2). */
typedef struct bec_1d_params {
    double zmax, dz;
    double g11, g12, g22, E, mu;
    double Va, Vb;
    double dt;
    int ntG, itmax;
    int ntW;
    int ne;
    double n0;        /* uniform start of the ground state */
    uint64_t seed;
    double noise_dev;
} bec_1d_params;

typedef struct bec_1d_report {
    double N0;            /* atoms of the ground state: the last entry of its N series */
    int nv;               /* lattice cells */
    double gs_residual;   /* ||H psi - mu psi|| / (|mu| ||psi||) of the ground state */
    double gs_mu;
    double elapsed_ms;    /* device time of ground state + ensemble + reduction */
} bec_1d_report;

/* Lattice cells nv of these parameters, or minus a status code when they are invalid. */
int bec_1d_cells(const bec_1d_params *p);

/* m, var: [2][ntW] ensemble means of N1, N2 and their standard errors computed
 * as wigner_1d.cpp:405-410 writes them (sqrt(var / (ne - 1)) with no absolute
 * value: NaN for ne = 1); vis: [ntW] mean visibility 2|<a|b>| / (Na + Nb) before
 * the second pulse (may be NULL); gs_x: [nv] complex ground state (may be NULL);
 * avn: [nv] mean x-space density minus the vacuum term after the last step
 * (may be NULL); ntot: [ntG] the ground-state loop's N series (may be NULL). */
int bec_run_1d(const bec_1d_params *p, double *m, double *var, double *vis, double *gs_x, double *avn,
               double *ntot, bec_1d_report *rep);

/* ---------------------------------------------------------- evolution */
/* initEvolution (wigner.cu:202-227): ground state + Philox vacuum noise of
 * variance 1/(2 dV) per mode (initializeEnsembles, wigner_kernel.cu:295-317),
 * to k-space, equilibration steps (one at least when equilibrate != 0, quirk
 * Q5), pi/2 pulse.  gs_x: cells complex values (x-space). */
int bec_init_evolution(bec_ctx *ctx, int group, const double *gs_x, uint64_t seed, int equilibrate);

/* The hot path: n_steps of propagate() (wigner.cu:178-199) with step dt
 * (natural units; dt <= 0 selects dt_evo) on every group.  When
 * sample_every > 0, after every sample_every-th step the per-trajectory sums
 * {sum|a|^2, sum|b|^2, Re sum conj(a) b, Im sum conj(a) b} of the x-space
 * state (FP64) are recorded; sums_out (may be NULL) receives
 * [n_steps/sample_every][total trajectories][4] doubles, groups in order. */
int bec_propagate(bec_ctx *ctx, int n_steps, double dt, int sample_every, double *sums_out);

/* Captures (without running them) the step graphs that bec_propagate with these
 * arguments replays: each segment of up to 64 steps between samples is one
 * graph launch, longer ones replay 32-step blocks.  A later bec_propagate then
 * spends no time on capture.  No-op when graphs are off (bec_set_graphs). */
int bec_propagate_prepare(bec_ctx *ctx, int n_steps, double dt, int sample_every);

/* Same, captured once into a hipGraph of `steps_per_graph` steps and
 * replayed (launch-overhead-free inner loop).  n_steps must be a multiple of
 * steps_per_graph. */
int bec_propagate_graph(bec_ctx *ctx, int n_steps, double dt, int steps_per_graph);

/* Observables of the current state (calculateEvolution tail, wigner.cu:345-410). */
typedef struct bec_obs {
    double t;                   /* natural units */
    double Na, Nb;              /* sums of vacuum-subtracted densities after the virtual pi/2 pulse (alpha = 0), times dV */
    double N_total;             /* (Na + Nb) / ne: the per-step printout of wigner.cu:369 */
    double visibility_grid;     /* getVisibility: max over alpha in {0, 0.5, .., 6} (wigner.cu:267-277) */
    double visibility;          /* 2 |<psi1|psi2>| / (N1 + N2), vacuum-subtracted */
    double phase;               /* arg sum conj(a) b: mean relative phase */
    double N1, N2;              /* vacuum-subtracted populations without the virtual pulse, dV units */
    double Na_stderr, Nb_stderr; /* standard error of the ensemble mean of the per-trajectory Na, Nb
                                    (sqrt(var / (ne - 1)), the error bars of 2d/wigner.cpp:438-443); 0 for ne = 1 */
} bec_obs;

int bec_observables(bec_ctx *ctx, int group, bec_obs *out);

/* calculateEvolution (wigner.cu:340-411): one propagate(dt) and the observables. */
int bec_calculate_evolution(bec_ctx *ctx, double dt, bec_obs *out_per_group);

/* Per-trajectory sums of the current x-space state: [ne][4] as in bec_propagate. */
int bec_trajectory_sums(bec_ctx *ctx, int group, double *sums);

/* Density projections after the virtual pulse halfPiRotate(alpha), ensemble
 * summed and vacuum subtracted (wigner.cu:279-303,384-410).
 * xy: [nvy][nvx] (sum over z), zy: [nvy][nvz] (sum over x), axial: [nvz]. */
int bec_projections(bec_ctx *ctx, int group, double alpha, double *xy_a, double *xy_b,
                    double *zy_a, double *zy_b, double *axial_a, double *axial_b);

/* bec_projections with no device allocation on any call: its workspace is a
 * fixed device array of the library (4.25 MiB per GPU, held until the process
 * ends; every supported lattice fits, a larger workspace gives BEC_ERR_SHAPE).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Calls that overlap on one GPU share the array and race; each call writes all
 * of the workspace it reads, so nothing carries over from a previous call. */
int bec_projections_static(bec_ctx *ctx, int group, double alpha, double *xy_a, double *xy_b,
                           double *zy_a, double *zy_b, double *axial_a, double *axial_b);

/* applyHalfPiPulse / applyPiPulse (wigner_kernel.cu:320-341) on every group. */
int bec_pulse(bec_ctx *ctx, int kind);

/* A resonant pulse of area theta and phase phi on every group:
 *   (a, b) <- U (a, b),  U = [[cos(theta/2), -i e^{-i phi} sin(theta/2)],
 *                             [-i e^{+i phi} sin(theta/2), cos(theta/2)]].
 * U(pi/2, 0) and U(pi, 0) are applyHalfPiPulse and applyPiPulse
 * (wigner_kernel.cu:320-341); U(pi/2, -alpha) is the measurement rotation
 * halfPiRotate(alpha) (wigner_kernel.cu:344-360).  The pulse is local, so it
 * acts on the stored state in any representation. */
int bec_pulse_rotation(bec_ctx *ctx, double theta, double phi);

/* Checkpoint: read / write the state of a group in x- or k-space
 * (k-space: psi_k = (1/cells) sum_x psi_x exp(+i k.x), the reference's), or in
 * the stored hybrid representation (BEC_SPACE_HYBRID), which restores the
 * state bit for bit (the host copies of wigner.cu:166,207).  A resumed run
 * equals the uninterrupted one bit for bit when it also restores the time
 * (bec_set_time) and, with loss noise on, the noise step
 * (bec_get_loss_noise_step, then bec_set_loss_noise with that first step). */
int bec_get_state(bec_ctx *ctx, int group, int space, double *a, double *b);
int bec_set_state(bec_ctx *ctx, int group, int space, const double *a, const double *b);
double bec_get_time(const bec_ctx *ctx);
int bec_set_time(bec_ctx *ctx, double t);

/* Device time of each pass kernel averaged over `iters` launches, measured
 * with HIP events on the context stream: ms[0] z-pass, ms[1] xy-pass
 * (nonlinear), ms[2] sample pass.  Leaves the state unchanged in value
 * class (it runs real steps). */
int bec_time_passes(bec_ctx *ctx, int group, int iters, double *ms);

/* Launch schedule of the two hot passes of a group: pencils per z-pass
 * workgroup (pw), z values per thread in registers (z_r1, the radix of the
 * first z stage), one component per z workgroup (z_split), non-temporal
 * loads/stores (nt), XCD-aware workgroup placement (xcd), minimum waves per
 * SIMD the xy pass's register allocation must allow (xy_minw: 1 or 5), z LDS
 * exchange in two real-valued rounds through half the LDS (z_half), one
 * wavefront per component in the real-time xy pass (xy_split; FP32, itmax = 1,
 * loss noise off). */
typedef struct bec_schedule {
    int pw, z_r1, z_split, nt, xcd, xy_minw, z_half, xy_split;
    double ms_per_step;   /* filled by bec_autotune */
} bec_schedule;

int bec_set_schedule(bec_ctx *ctx, int group, const bec_schedule *sch);
int bec_get_schedule(bec_ctx *ctx, int group, bec_schedule *sch);
/* Candidate schedules for the group's lattice, precision and batch (*count of
 * them): the ones bec_autotune times.  The split xy pass is among them only
 * while the two-component pass's wavefronts do not fill the GPU once;
 * bec_set_schedule accepts it at any batch. */
int bec_schedule_candidates(bec_ctx *ctx, int group, bec_schedule *out, int max_out, int *count);
/* Times every candidate with `steps` alternating z/xy passes on the group's
 * own state (saved and restored), replayed from a captured graph when the
 * context uses graphs, four times (the first warms up, the fastest of the
 * other three counts), and keeps the fastest candidate; results receives
 * per-candidate ms_per_step, *n_results the number of candidates. */
int bec_autotune(bec_ctx *ctx, int group, int steps, bec_schedule *results, int max_results,
                 int *n_results, bec_schedule *best);

/* calculateAverages (wigner.cu:305-337), the reference's debug report of the
 * ensemble, on the x-space state of group `group` (its per-cell ensemble sums
 * in fixed order, quirk Q6 fixed):
 *   N     = (sum|a|^2 + sum|b|^2) dV / ne, the atoms per trajectory it prints;
 *   norm  = N_model / ((sum|a|^2 + sum|b|^2) / ne);
 *   var_* = the ensemble variance (1/ne) sum_j (x_j - <x>)^2 of each quadrature
 *           per cell (kernelAverage, wigner_kernel.cu:383-392), averaged over
 *           cells and multiplied by norm, as it prints them. */
typedef struct bec_avg_report {
    double N, norm;
    double var_re_a, var_im_a, var_re_b, var_im_b;
} bec_avg_report;
int bec_calculate_averages(bec_ctx *ctx, int group, bec_avg_report *out);

/* Enable (default) or disable the replay of hipGraph-captured step blocks
 * inside bec_propagate (disabled: one launch per pass). */
int bec_set_graphs(bec_ctx *ctx, int enable);

/* Step every group of the context together: enable != 0 makes each pass of
 * bec_propagate one launch over all groups' tiles (a table of tile ranges
 * walked by the workgroups the GPU keeps resident), in place of one launch
 * per group, or one graph branch per group.  Replaces the reference's
 * per-frame, per-grid launches (calculateEvolution, wigner.cu:340-411).
 * Needs FP32 states, at most 8 groups, and the same itmax == 1 and
 * loss-noise setting in every group: BEC_ERR_UNSUPPORTED otherwise (the
 * setting then stays off).  The launch uses fixed schedules (512-thread z
 * workgroups, 64-thread xy workgroups), not the groups' own (bec_set_schedule).
 * enable == 0 (the default) restores the per-group launches. */
int bec_set_multi_launch(bec_ctx *ctx, int enable);

/* Loss-induced Wigner noise during the evolution.  No counterpart in the
 * reference's code: its report derives the diffusion of the loss terms
 * (docs/wigner_simulation.report/report.tex:624-656; the poster
 * docs/icap2010.poster/poster.tex:245-290) but its kernel
 * (wigner_kernel.cu:184-229) integrates only the drift.  enable != 0 adds to
 * every nonlinear step of every group (bec_propagate, the equilibration of
 * bec_init_evolution, bec_autotune and bec_time_passes steps alike) the Ito
 * increment of the leading-order diffusion, at the order of the drift the
 * kernel integrates:
 *   D11 = 3/2 l111 |a|^4 + 1/2 l12 |b|^2,  D22 = 1/2 l12 |a|^2 + l22 |b|^2,
 *   D12 = 1/2 l12 a conj(b),  <d psi_i d psi_j*> = D_ij dt / dV per cell,
 * from the state at the start of the step, with normal deviates from the
 * Philox stream {cell, trajectory, step, 1 + group}, key seed.  The step
 * counter of each group starts at first_step and advances once per nonlinear
 * step, so a run resumed from a checkpoint continues the same stream when it
 * passes the step read by bec_get_loss_noise_step.  enable = 0 (the default)
 * restores the reference's noiseless step. */
int bec_set_loss_noise(bec_ctx *ctx, int enable, uint64_t seed, unsigned first_step);
/* The next step index of group `group`'s loss-noise stream (BEC_ERR_STATE when off). */
int bec_get_loss_noise_step(bec_ctx *ctx, int group, unsigned *step);

/* ------------------------------------------------- batched FFT sub-API */
/* batchfftPlan3d / batchfftExecute / batchfftDestroy (batchfft.h:83-108).
 * Unnormalised C2C over batch contiguous nz x ny x nx arrays in device
 * memory; sign = BEC_FFT_FORWARD (exp(-i)) or BEC_FFT_INVERSE (exp(+i)).
 * in == out performs the transform in place. */
typedef struct bec_fft_plan bec_fft_plan;
int bec_fft_plan3d(bec_fft_plan **plan, int nz, int ny, int nx, int precision, int batch);
int bec_fft_execute(bec_fft_plan *plan, const void *d_in, void *d_out, int sign, void *hip_stream);
int bec_fft_destroy(bec_fft_plan *plan);

/* ------------------------------------------------------ device memory */
/* Small helpers so callers without a device-memory library can drive the
 * FFT sub-API. */
int bec_device_alloc(void **d_ptr, size_t bytes);
int bec_device_free(void *d_ptr);
int bec_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
int bec_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
int bec_device_count(int *count);
/* Plain streaming device-to-device copy kernel (bytes % 16 == 0); used to
 * calibrate hardware byte counters against a known traffic. */
int bec_device_copy(void *d_dst, const void *d_src, size_t bytes, void *hip_stream);
/* Device time of the same copy kernel on `bytes` (read + write), averaged over
 * iters launches with HIP events: the achievable-bandwidth reference. */
int bec_time_copy(size_t bytes, int iters, double *ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* BEC_API_H */
