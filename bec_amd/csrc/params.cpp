// Physical parameters: the SI -> natural-unit derivation of
// programs/bec_wigner_3d.cuda/wigner.cpp:119-176 (evaluated in double).
#include <math.h>
#include <string.h>

#include "../../include/bec_api.h"

extern "C" int bec_default_model_params(bec_model_params *m)
{
    if (!m) return BEC_ERR_INVALID;
    memset(m, 0, sizeof(*m));
    // fillCalculationParameters (wigner.cpp:43-98)
    m->N = 150000;
    m->m = 1.443160648e-25;
    m->a11 = 100.4;
    m->a22 = 95.0;
    m->a12 = 97.66;
    m->fx = 97.6;
    m->fy = 97.6;
    m->fz = 11.96;
    m->detuning = -41;
    m->gamma111 = 0;
    m->gamma12 = 0;
    m->gamma22 = 0;
    m->Va = 1;
    m->Vb = 1;
    m->nvx = 16;
    m->nvy = 16;
    m->nvz = 128;
    m->itmax = 1;
    m->dt_steady = 1e-5;
    m->t_equilib = 0;
    m->dt_evo = 4e-5;
    m->ne = 4;
    return BEC_OK;
}

static bool pow2(int n) { return n > 1 && (n & (n - 1)) == 0; }

extern "C" int bec_fill_derived_params(const bec_model_params *m, bec_calc_params *c)
{
    if (!m || !c) return BEC_ERR_INVALID;
    if (!(m->m > 0) || !(m->fx > 0) || !(m->fz > 0) || m->N <= 0 || m->ne <= 0) return BEC_ERR_INVALID;
    if (!pow2(m->nvx) || !pow2(m->nvy) || !pow2(m->nvz)) return BEC_ERR_SHAPE;
    const double h_bar = 1.054571628e-34;  // wigner.cpp:121
    const double a0 = 5.2917720859e-11;    // wigner.cpp:122
    memset(c, 0, sizeof(*c));
    const double w_rho = 2.0 * M_PI * m->fx;
    c->l_rho = sqrt(h_bar / (m->m * w_rho));
    c->t_rho = 1.0 / w_rho;
    c->lambda = m->fx / m->fz;
    c->nvx = m->nvx;
    c->nvy = m->nvy;
    c->nvz = m->nvz;
    c->cells = m->nvx * m->nvy * m->nvz;
    c->Va = m->Va;
    c->Vb = m->Vb;
    c->V = (m->Va + m->Vb) / 2.0;
    c->detuning = 2.0 * M_PI * m->detuning / w_rho;
    c->l111 = (m->gamma111 / 1e12) / (pow(c->l_rho, 6) * w_rho);
    c->l12 = (m->gamma12 / 1e6) / (pow(c->l_rho, 3) * w_rho);
    c->l22 = (m->gamma22 / 1e6) / (pow(c->l_rho, 3) * w_rho);
    c->g11 = 4.0 * M_PI * m->a11 * a0 / c->l_rho;
    c->g12 = 4.0 * M_PI * m->a12 * a0 / c->l_rho;
    c->g22 = 4.0 * M_PI * m->a22 * a0 / c->l_rho;
    c->N = m->N;
    c->mu = pow(15.0 * c->N * c->g11 / (16.0 * M_PI * c->lambda * sqrt(2.0)), 0.4);
    c->xmax = 1.2 * sqrt(2.0 * c->mu);
    c->ymax = 1.2 * sqrt(2.0 * c->mu);
    c->zmax = 1.2 * c->lambda * sqrt(2.0 * c->mu);
    c->dx = 2.0 * c->xmax / (c->nvx - 1);
    c->dy = 2.0 * c->ymax / (c->nvy - 1);
    c->dz = 2.0 * c->zmax / (c->nvz - 1);
    c->dkx = M_PI / c->xmax;   // pi / xmax, as the reference (not 2 pi / (n dx))
    c->dky = M_PI / c->ymax;
    c->dkz = M_PI / c->zmax;
    c->itmax = m->itmax;
    c->dt_steady = m->dt_steady / c->t_rho;
    c->t_equilib = m->t_equilib / c->t_rho;
    c->dt_evo = m->dt_evo / c->t_rho;
    c->ne = m->ne;
    return BEC_OK;
}

// The 2D program's parameters (programs/bec_wigner_2d.cpp/wigner.cpp:106-136
// meanings) for the cloud of a 3D parameter set: the radial plane of the trap
// with the axial direction in its oscillator ground state.  In the program's
// units (energy hbar w, time 1/w, length l = sqrt(hbar / (2 m w)), w = 2 pi fx)
// the plane's Hamiltonian is -lap + r^2/4 + g |psi|^2 with
// dx dy sum |psi|^2 = N, where g = g2D / (hbar w l^2) and
// g2D = (4 pi hbar^2 a / m) / (sqrt(2 pi) a_z): g = 4 sqrt(2 pi) a / a_z.
extern "C" int bec_fill_2d_params(const bec_model_params *m, bec_2d_params *p)
{
    if (!m || !p) return BEC_ERR_INVALID;
    if (!(m->m > 0) || !(m->fx > 0) || !(m->fz > 0) || m->N <= 0 || m->ne <= 0) return BEC_ERR_INVALID;
    if (!pow2(m->nvx) || !pow2(m->nvy)) return BEC_ERR_SHAPE;
    const double h_bar = 1.054571628e-34;  // wigner.cpp:121
    const double a0 = 5.2917720859e-11;    // wigner.cpp:122
    memset(p, 0, sizeof(*p));
    const double w = 2.0 * M_PI * m->fx;
    const double a_z = sqrt(h_bar / (m->m * 2.0 * M_PI * m->fz));
    const double k = 4.0 * sqrt(2.0 * M_PI) * a0 / a_z;
    p->g11 = k * m->a11;
    p->g12 = k * m->a12;
    p->g22 = k * m->a22;
    p->mu = sqrt(p->g11 * m->N / (2.0 * M_PI));
    p->nvx = m->nvx;
    p->nvy = m->nvy;
    p->xmax = p->ymax = 1.2 * 2.0 * sqrt(p->mu);
    p->E = 2.0 * M_PI * m->detuning / w;
    p->Va = m->Va;
    p->Vb = m->Vb;
    p->dtG = w * m->dt_steady;
    p->ntG = 1;
    p->itmax = m->itmax;
    p->dtW = w * m->dt_evo;
    p->ntW = 1;
    p->n_meas = 1;
    p->ne = m->ne;
    p->n0 = 1.0;
    p->seed = 0x5EED;
    const double dx = 2.0 * p->xmax / (p->nvx - 1), dy = 2.0 * p->ymax / (p->nvy - 1);
    p->noise_dev = 0.5 / sqrt(dx * dy);
    p->gs_rule = 2;
    p->gs_tol = 1e-8;
    p->gs_max_iter = 100000;
    p->n_atoms = m->N;
    p->gs_tau = 0.5;
    p->gs_beta = 0.6;
    return BEC_OK;
}
