/*
 * FP64 CPU oracle for the split-step Wigner GPE propagator.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Every function restates the reference algorithm it cites
 * (paths relative to the fjarri-attic/bec tree).  The FFT is our own
 * radix-2 transform (cuFFT/FFTW are the reference's libraries and are not
 * available here); it is pinned by a brute-force DFT in the tests.
 */
#include "bec_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static inline orc_c cmk(double re, double im) { orc_c r = {re, im}; return r; }
static inline orc_c cmul(orc_c a, orc_c b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline orc_c cadd(orc_c a, orc_c b) { return cmk(a.re + b.re, a.im + b.im); }
static inline orc_c csub(orc_c a, orc_c b) { return cmk(a.re - b.re, a.im - b.im); }
static inline orc_c cscale(orc_c a, double s) { return cmk(a.re * s, a.im * s); }
static inline double cnorm(orc_c a) { return a.re * a.re + a.im * a.im; }
static inline orc_c cexp_(orc_c a) { double m = exp(a.re); return cmk(m * cos(a.im), m * sin(a.im)); }

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* Parameters: fillDerivedParameters (bec_wigner_3d.cuda/wigner.cpp:119-176),
 * evaluated in double (the reference evaluates in float, quirk Q4).       */
void orc_derive(const orc_model *m, orc_calc *c)
{
    const double h_bar = 1.054571628e-34;   /* wigner.cpp:121 */
    const double a0 = 5.2917720859e-11;     /* wigner.cpp:122 */
    double w_rho = 2.0 * M_PI * m->fx;
    memset(c, 0, sizeof(*c));
    c->l_rho = sqrt(h_bar / (m->m * w_rho));
    c->t_rho = 1.0 / w_rho;
    c->lambda = m->fx / m->fz;
    c->nvx = m->nvx; c->nvy = m->nvy; c->nvz = m->nvz;
    c->cells = m->nvx * m->nvy * m->nvz;
    c->Va = m->Va; c->Vb = m->Vb; c->V = (m->Va + m->Vb) / 2.0;
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
    c->dkx = M_PI / c->xmax;   /* quirk Q2: pi/xmax, not 2 pi/(n dx) */
    c->dky = M_PI / c->ymax;
    c->dkz = M_PI / c->zmax;
    c->itmax = m->itmax;
    c->dt_steady = m->dt_steady / c->t_rho;
    c->t_equilib = m->t_equilib / c->t_rho;
    c->dt_evo = m->dt_evo / c->t_rho;
    c->ne = m->ne;
}

/* getK (wigner_kernel.cu:15-21): Nyquist index stays positive */
static double getk(int i, double dk, int n) { return (2 * i > n) ? dk * (i - n) : dk * i; }

/* initWaveVectors (wigner_kernel.cu:25-55) */
void orc_ktable(const orc_calc *c, double *k2half)
{
    for (int z = 0; z < c->nvz; z++)
        for (int y = 0; y < c->nvy; y++)
            for (int x = 0; x < c->nvx; x++) {
                double kx = getk(x, c->dkx, c->nvx), ky = getk(y, c->dky, c->nvy), kz = getk(z, c->dkz, c->nvz);
                k2half[(z * c->nvy + y) * c->nvx + x] = (kx * kx + ky * ky + kz * kz) / 2.0;
            }
}

/* potential (wigner_kernel.cu:99-113) */
double orc_potential(const orc_calc *c, int cell)
{
    int x = cell % c->nvx, y = (cell / c->nvx) % c->nvy, z = cell / (c->nvx * c->nvy);
    double xx = -c->xmax + c->dx * x, yy = -c->ymax + c->dy * y, zz = -c->zmax + c->dz * z;
    return (xx * xx + yy * yy + zz * zz / (c->lambda * c->lambda)) / 2.0;
}

/* ------------------------------------------------------------------ */
/* FFT: iterative radix-2 Cooley-Tukey, in place, power-of-two n.       */
void orc_fft1d(orc_c *d, int n, int sign)
{
    if (n <= 1) return;
    for (int i = 1, j = 0; i < n; i++) {           /* bit reversal */
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { orc_c t = d[i]; d[i] = d[j]; d[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; k++) {
            double ang = sign * 2.0 * M_PI * (double)k / (double)len;
            orc_c w = cmk(cos(ang), sin(ang));
            for (int i = k; i < n; i += len) {
                orc_c u = d[i], v = cmul(d[i + half], w);
                d[i] = cadd(u, v);
                d[i + half] = csub(u, v);
            }
        }
    }
}

/* axes_mask: bit0 = x, bit1 = y, bit2 = z */
void orc_fft_axes(orc_c *data, int nz, int ny, int nx, int batch, int sign, int axes_mask)
{
    long cells = (long)nx * ny * nz;
    if (axes_mask & 1) {
        long lines = (long)batch * nz * ny;
#pragma omp parallel for schedule(static)
        for (long l = 0; l < lines; l++) orc_fft1d(data + l * nx, nx, sign);
    }
    if (axes_mask & 2) {
        long lines = (long)batch * nz * nx;
#pragma omp parallel for schedule(static)
        for (long l = 0; l < lines; l++) {
            orc_c buf[4096];
            long e = l / ((long)nz * nx), r = l % ((long)nz * nx);
            long z = r / nx, x = r % nx;
            orc_c *base = data + e * cells + z * nx * ny + x;
            for (int y = 0; y < ny; y++) buf[y] = base[(long)y * nx];
            orc_fft1d(buf, ny, sign);
            for (int y = 0; y < ny; y++) base[(long)y * nx] = buf[y];
        }
    }
    if (axes_mask & 4) {
        long lines = (long)batch * ny * nx;
#pragma omp parallel for schedule(static)
        for (long l = 0; l < lines; l++) {
            orc_c buf[4096];
            long e = l / ((long)ny * nx), r = l % ((long)ny * nx);
            orc_c *base = data + e * cells + r;
            for (int z = 0; z < nz; z++) buf[z] = base[(long)z * nx * ny];
            orc_fft1d(buf, nz, sign);
            for (int z = 0; z < nz; z++) base[(long)z * nx * ny] = buf[z];
        }
    }
}

void orc_fft3d(orc_c *data, int nz, int ny, int nx, int batch, int sign)
{
    orc_fft_axes(data, nz, ny, nx, batch, sign, 7);
}

/* O(N^2) DFT, for pinning the FFT on tiny sizes */
void orc_dft3d_bruteforce(const orc_c *in, orc_c *out, int nz, int ny, int nx, int sign)
{
    for (int kz = 0; kz < nz; kz++)
        for (int ky = 0; ky < ny; ky++)
            for (int kx = 0; kx < nx; kx++) {
                double sr = 0, si = 0;
                for (int z = 0; z < nz; z++)
                    for (int y = 0; y < ny; y++)
                        for (int x = 0; x < nx; x++) {
                            double ph = sign * 2.0 * M_PI *
                                        ((double)(kz * z % nz) / nz + (double)(ky * y % ny) / ny +
                                         (double)(kx * x % nx) / nx);
                            orc_c v = in[(z * ny + y) * nx + x];
                            sr += v.re * cos(ph) - v.im * sin(ph);
                            si += v.re * sin(ph) + v.im * cos(ph);
                        }
                out[(kz * ny + ky) * nx + kx] = cmk(sr, si);
            }
}

void orc_to_kspace(const orc_calc *c, orc_c *data, int ne)
{
    orc_fft3d(data, c->nvz, c->nvy, c->nvx, ne, +1);
    long n = (long)c->cells * ne;
    double s = 1.0 / c->cells;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) data[i] = cscale(data[i], s);
}

void orc_to_xspace(const orc_calc *c, orc_c *data, int ne)
{
    orc_fft3d(data, c->nvz, c->nvy, c->nvx, ne, -1);
}

/* ------------------------------------------------------------------ */
/* Ground state                                                         */

/* fillWithTFGroundState (wigner_kernel.cu:116-125) */
void orc_tf_state(const orc_calc *c, orc_c *out)
{
    for (int i = 0; i < c->cells; i++) {
        double e = c->mu - orc_potential(c, i);
        out[i] = e > 0 ? cmk(sqrt(e / c->g11), 0.0) : cmk(0.0, 0.0);
    }
}

static double particles(const orc_calc *c, const orc_c *x)
{
    double s = 0;
    for (int i = 0; i < c->cells; i++) s += cnorm(x[i]);
    return s * c->dx * c->dy * c->dz;
}

/* calculateStateIntegral (wigner.cu:38-67) with the kernels of
 * wigner_kernel.cu:232-258: x-space psi times k-space psi at the same index
 * and no conjugate (quirk Q3).  Returns the integral, not per particle. */
double orc_state_integral_ref(const orc_calc *c, const orc_c *psi_x, int energy)
{
    int n = c->cells;
    orc_c *k = malloc(sizeof(orc_c) * n);
    double *k2 = malloc(sizeof(double) * n);
    memcpy(k, psi_x, sizeof(orc_c) * n);
    orc_to_kspace(c, k, 1);
    orc_ktable(c, k2);
    double s = 0;
    for (int i = 0; i < n; i++) {
        double na = cnorm(psi_x[i]);
        double v = orc_potential(c, i);
        double r = energy ? na * (v + c->g11 * na / 2.0) : na * (v + c->g11 * na);
        orc_c t = cmul(psi_x[i], cscale(k[i], k2[i]));
        s += r + t.re;
    }
    free(k); free(k2);
    return s * c->dx * c->dy * c->dz;
}

/* H psi with H = k^2/2 (dk table) + V + g11 |psi|^2 */
static void apply_h(const orc_calc *c, const orc_c *psi, orc_c *out)
{
    int n = c->cells;
    double *k2 = malloc(sizeof(double) * n);
    orc_ktable(c, k2);
    memcpy(out, psi, sizeof(orc_c) * n);
    orc_to_kspace(c, out, 1);
    for (int i = 0; i < n; i++) out[i] = cscale(out[i], k2[i]);
    orc_to_xspace(c, out, 1);
    for (int i = 0; i < n; i++) {
        double v = orc_potential(c, i) + c->g11 * cnorm(psi[i]);
        out[i] = cadd(out[i], cscale(psi[i], v));
    }
    free(k2);
}

void orc_energy_mu(const orc_calc *c, const orc_c *psi_x, double *E, double *mu, double *N)
{
    int n = c->cells;
    orc_c *k = malloc(sizeof(orc_c) * n);
    double *k2 = malloc(sizeof(double) * n);
    memcpy(k, psi_x, sizeof(orc_c) * n);
    orc_to_kspace(c, k, 1);
    orc_ktable(c, k2);
    double kin = 0, pot = 0, inter = 0, norm = 0;
    for (int i = 0; i < n; i++) {
        double na = cnorm(psi_x[i]);
        kin += k2[i] * cnorm(k[i]);
        pot += orc_potential(c, i) * na;
        inter += c->g11 * na * na;
        norm += na;
    }
    double dV = c->dx * c->dy * c->dz;
    kin *= (double)n;   /* Parseval: sum_x |f_x|^2 = cells * sum_k |f_k|^2 */
    double Nn = norm * dV;
    if (E) *E = (kin + pot + inter / 2.0) * dV / Nn;
    if (mu) *mu = (kin + pot + inter) * dV / Nn;
    if (N) *N = Nn;
    free(k); free(k2);
}

double orc_residual(const orc_calc *c, const orc_c *psi_x, double *mu_out)
{
    int n = c->cells;
    orc_c *h = malloc(sizeof(orc_c) * n);
    apply_h(c, psi_x, h);
    double num = 0, den = 0;
    for (int i = 0; i < n; i++) {
        num += psi_x[i].re * h[i].re + psi_x[i].im * h[i].im;
        den += cnorm(psi_x[i]);
    }
    double mu = num / den;
    double r = 0;
    for (int i = 0; i < n; i++) r += cnorm(csub(h[i], cscale(psi_x[i], mu)));
    free(h);
    if (mu_out) *mu_out = mu;
    return sqrt(r / den) / fabs(mu);
}

/* propagateKSpaceImaginaryTime (wigner_kernel.cu:128-134) */
static void kinetic_im(const orc_calc *c, const double *k2, orc_c *a)
{
    for (int i = 0; i < c->cells; i++) a[i] = cscale(a[i], exp(-c->dt_steady / 2.0 * k2[i]));
}

/* propagateXSpaceOneComponent (wigner_kernel.cu:154-181); itmax = 0 gives
 * zero output in the reference (uninitialised da, quirk Q10) */
static void nonlinear_im(const orc_calc *c, orc_c *a)
{
    for (int i = 0; i < c->cells; i++) {
        orc_c a0 = a[i], cur = a0;
        double v = orc_potential(c, i), da = 0.0;
        for (int it = 0; it < c->itmax; it++) {
            da = exp(c->dt_steady / 2.0 * (-v - c->g11 * cnorm(cur)));
            cur = cscale(a0, da);
        }
        a[i] = cscale(cur, da);
    }
}

/* One iteration of the loop body of calculateSteadyState (wigner.cu:114-153),
 * on a k-space state; returns the x-space renormalised state's E_ref.      */
static double gs_iteration(const orc_calc *c, const double *k2, orc_c *a, orc_c *x_out)
{
    kinetic_im(c, k2, a);
    orc_to_xspace(c, a, 1);
    nonlinear_im(c, a);
    orc_to_kspace(c, a, 1);
    kinetic_im(c, k2, a);
    orc_to_xspace(c, a, 1);
    double N = particles(c, a);
    double s = sqrt((double)c->N / N);
    for (int i = 0; i < c->cells; i++) a[i] = cscale(a[i], s);
    double E = orc_state_integral_ref(c, a, 1);
    if (x_out) memcpy(x_out, a, sizeof(orc_c) * c->cells);
    orc_to_kspace(c, a, 1);
    return E;
}

void orc_gs_iterate(const orc_calc *c, orc_c *a_k, int n_iter, double *E_ref_out)
{
    double *k2 = malloc(sizeof(double) * c->cells);
    orc_ktable(c, k2);
    double E = 0;
    for (int i = 0; i < n_iter; i++) E = gs_iteration(c, k2, a_k, NULL);
    if (E_ref_out) *E_ref_out = E;
    free(k2);
}

/* calculateSteadyState (wigner.cu:84-175).  Returns the x-space state. */
int orc_ground_state(const orc_calc *c, const orc_c *guess_x, int rule, double tol,
                     int max_iter, int check_every, orc_c *out_x, orc_gs_report *rep)
{
    int n = c->cells;
    orc_c *a = malloc(sizeof(orc_c) * n), *x = malloc(sizeof(orc_c) * n);
    double *k2 = malloc(sizeof(double) * n);
    orc_ktable(c, k2);
    if (guess_x) memcpy(a, guess_x, sizeof(orc_c) * n);
    else orc_tf_state(c, a);
    double N = particles(c, a);
    double s = sqrt((double)c->N / N);
    for (int i = 0; i < n; i++) a[i] = cscale(a[i], s);
    orc_to_kspace(c, a, 1);
    if (check_every < 1) check_every = 1;

    double E = 0, res = -1;
    int it = 0, conv = 0;
    while (it < max_iter) {
        double newE = gs_iteration(c, k2, a, x);
        it++;
        if (rule == ORC_GS_RULE_REFERENCE) {
            if (fabs((newE - E) / newE) < tol) { conv = 1; break; }
        } else if (it % check_every == 0) {
            res = orc_residual(c, x, NULL);
            if (res < tol) { conv = 1; break; }
        }
        E = newE;
    }
    orc_to_xspace(c, a, 1);
    if (out_x) memcpy(out_x, a, sizeof(orc_c) * n);
    if (rep) {
        rep->iterations = it;
        rep->converged = conv;
        rep->E_ref = orc_state_integral_ref(c, a, 1) / c->N;
        rep->mu_ref = orc_state_integral_ref(c, a, 0) / c->N;
        orc_energy_mu(c, a, &rep->E, &rep->mu, &rep->N);
        rep->residual = orc_residual(c, a, NULL);
    }
    free(a); free(x); free(k2);
    return conv ? 0 : 1;
}

/* ------------------------------------------------------------------ */
/* Two-component ground state: the coupled GPEs of
 * docs/wigner_simulation.report/report.tex:200-211 in imaginary time. */

static void kinetic_apply(const orc_calc *c, const double *k2, const orc_c *in, orc_c *out)
{
    memcpy(out, in, sizeof(orc_c) * c->cells);
    orc_to_kspace(c, out, 1);
    for (int i = 0; i < c->cells; i++) out[i] = cscale(out[i], k2[i]);
    orc_to_xspace(c, out, 1);
}

void orc_apply_h2(const orc_calc *c, const orc_c *a, const orc_c *b, orc_c *ha, orc_c *hb)
{
    double *k2 = malloc(sizeof(double) * c->cells);
    orc_ktable(c, k2);
    kinetic_apply(c, k2, a, ha);
    kinetic_apply(c, k2, b, hb);
    for (int i = 0; i < c->cells; i++) {
        double v = orc_potential(c, i), na = cnorm(a[i]), nb = cnorm(b[i]);
        ha[i] = cadd(ha[i], cscale(a[i], v + c->g11 * na + c->g12 * nb));
        hb[i] = cadd(hb[i], cscale(b[i], v + c->g22 * nb + c->g12 * na));
    }
    free(k2);
}

static double comp_residual(const orc_calc *c, const orc_c *x, const orc_c *h, double *mu_out)
{
    double num = 0, den = 0, r = 0;
    for (int i = 0; i < c->cells; i++) {
        num += x[i].re * h[i].re + x[i].im * h[i].im;
        den += cnorm(x[i]);
    }
    double mu = num / den;
    for (int i = 0; i < c->cells; i++) r += cnorm(csub(h[i], cscale(x[i], mu)));
    if (mu_out) *mu_out = mu;
    return sqrt(r / den) / fabs(mu);
}

double orc_residual2(const orc_calc *c, const orc_c *a, const orc_c *b, double *mu_a, double *mu_b)
{
    orc_c *ha = malloc(sizeof(orc_c) * c->cells), *hb = malloc(sizeof(orc_c) * c->cells);
    orc_apply_h2(c, a, b, ha, hb);
    double ra = comp_residual(c, a, ha, mu_a), rb = comp_residual(c, b, hb, mu_b);
    free(ha); free(hb);
    return ra > rb ? ra : rb;
}

double orc_energy2(const orc_calc *c, const orc_c *a, const orc_c *b)
{
    int n = c->cells;
    orc_c *ka = malloc(sizeof(orc_c) * n), *kb = malloc(sizeof(orc_c) * n);
    double *k2 = malloc(sizeof(double) * n);
    orc_ktable(c, k2);
    memcpy(ka, a, sizeof(orc_c) * n);
    memcpy(kb, b, sizeof(orc_c) * n);
    orc_to_kspace(c, ka, 1);
    orc_to_kspace(c, kb, 1);
    double kin = 0, rest = 0, norm = 0;
    for (int i = 0; i < n; i++) {
        double na = cnorm(a[i]), nb = cnorm(b[i]);
        kin += k2[i] * (cnorm(ka[i]) + cnorm(kb[i]));
        rest += orc_potential(c, i) * (na + nb) + c->g11 * na * na / 2.0 + c->g22 * nb * nb / 2.0
                + c->g12 * na * nb;
        norm += na + nb;
    }
    free(ka); free(kb); free(k2);
    return (kin * (double)n + rest) / norm;   /* Parseval as in orc_energy_mu */
}

void orc_tf_state2(const orc_calc *c, double frac_a, orc_c *a, orc_c *b)
{
    orc_tf_state(c, a);
    double s = particles(c, a);
    double fa = sqrt(frac_a * c->N / s), fb = sqrt((1.0 - frac_a) * c->N / s);
    for (int i = 0; i < c->cells; i++) {
        b[i] = cscale(a[i], fb);
        a[i] = cscale(a[i], fa);
    }
}

static void renorm(const orc_calc *c, orc_c *x, double target)
{
    double s = sqrt(target / particles(c, x));
    for (int i = 0; i < c->cells; i++) x[i] = cscale(x[i], s);
}

int orc_ground_state2(const orc_calc *c, double frac_a, const orc_c *guess_a, const orc_c *guess_b,
                      double tol, int max_iter, double alpha, double tau, double beta,
                      orc_c *out_a, orc_c *out_b, orc_gs2_report *rep)
{
    int n = c->cells;
    size_t bytes = sizeof(orc_c) * n;
    orc_c *x[2], *pv[2], *h[2];
    double *pre = malloc(sizeof(double) * n);
    orc_ktable(c, pre);
    for (int i = 0; i < n; i++) pre[i] = 1.0 / (alpha + pre[i]);
    for (int k = 0; k < 2; k++) {
        x[k] = malloc(bytes); pv[k] = malloc(bytes); h[k] = malloc(bytes);
    }
    const double target[2] = {frac_a * c->N, (1.0 - frac_a) * c->N};
    if (guess_a && guess_b) {
        memcpy(x[0], guess_a, bytes);
        memcpy(x[1], guess_b, bytes);
    } else {
        orc_tf_state2(c, frac_a, x[0], x[1]);
    }
    for (int k = 0; k < 2; k++) {
        renorm(c, x[k], target[k]);
        memcpy(pv[k], x[k], bytes);
    }
    int it = 0, conv = 0;
    double res = 0, mu[2] = {0, 0};
    for (;;) {
        orc_apply_h2(c, x[0], x[1], h[0], h[1]);
        double r0 = comp_residual(c, x[0], h[0], &mu[0]), r1 = comp_residual(c, x[1], h[1], &mu[1]);
        res = r0 > r1 ? r0 : r1;
        if (!(res == res)) break;
        if (res < tol) { conv = 1; break; }
        if (it >= max_iter) break;
        for (int k = 0; k < 2; k++) {
            /* d = P (H - mu) x, applied in k-space */
            for (int i = 0; i < n; i++) h[k][i] = csub(h[k][i], cscale(x[k][i], mu[k]));
            orc_to_kspace(c, h[k], 1);
            for (int i = 0; i < n; i++) h[k][i] = cscale(h[k][i], pre[i]);
            orc_to_xspace(c, h[k], 1);
            for (int i = 0; i < n; i++) {
                orc_c cur = x[k][i];
                orc_c nw = cadd(csub(cur, cscale(h[k][i], tau)), cscale(csub(cur, pv[k][i]), beta));
                pv[k][i] = cur;
                x[k][i] = nw;
            }
            renorm(c, x[k], target[k]);
        }
        it++;
    }
    if (out_a) memcpy(out_a, x[0], bytes);
    if (out_b) memcpy(out_b, x[1], bytes);
    if (rep) {
        rep->iterations = it;
        rep->converged = conv;
        rep->E = orc_energy2(c, x[0], x[1]);
        rep->mu_a = mu[0];
        rep->mu_b = mu[1];
        rep->N_a = particles(c, x[0]);
        rep->N_b = particles(c, x[1]);
        rep->residual = res;
    }
    for (int k = 0; k < 2; k++) { free(x[k]); free(pv[k]); free(h[k]); }
    free(pre);
    return conv ? 0 : 1;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), restated from the
 * published round function.                                           */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Four N(0, 0.5^2) numbers for one cell of one trajectory: the Box-Muller
 * transform of fillWithNormalDistribution (wigner.cu:15-34, dev = 0.5) fed by
 * Philox with counter {cell, traj, stream, 0} and key = seed.           */
void orc_noise_cell(uint64_t seed, uint32_t traj, uint32_t cell, uint32_t stream, double out[4])
{
    uint32_t ctr[4] = {cell, traj, stream, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    for (int p = 0; p < 2; p++) {
        double u1 = ((double)r[2 * p] + 1.0) * (1.0 / 4294967296.0);
        double u2 = ((double)r[2 * p + 1] + 1.0) * (1.0 / 4294967296.0);
        double rad = 0.5 * sqrt(-2.0 * log(u1));
        out[2 * p] = rad * cos(2.0 * M_PI * u2);
        out[2 * p + 1] = rad * sin(2.0 * M_PI * u2);
    }
}

/* initializeEnsembles (wigner_kernel.cu:295-317): a = psi_GS + Va eta_a / sqrt(dV),
 * b = Vb eta_b / sqrt(dV) */
void orc_init_ensembles(const orc_calc *c, const orc_c *gs_x, uint64_t seed, orc_c *a_x, orc_c *b_x)
{
    double coeff = 1.0 / sqrt(c->dx * c->dy * c->dz);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < c->ne; e++)
        for (int i = 0; i < c->cells; i++) {
            double nz[4];
            orc_noise_cell(seed, (uint32_t)e, (uint32_t)i, 0u, nz);
            long idx = (long)e * c->cells + i;
            a_x[idx] = cmk(gs_x[i].re + c->Va * nz[0] * coeff, gs_x[i].im + c->Va * nz[1] * coeff);
            b_x[idx] = cmk(c->Vb * nz[2] * coeff, c->Vb * nz[3] * coeff);
        }
}

/* The equilibration loop of initEvolution (wigner.cu:222-223) runs in
 * float: for(t = 0; t <= t_equilib; t += dt_evo); at least one step (Q5). */
int orc_equilibration_steps(const orc_calc *c)
{
    float te = (float)c->t_equilib, dt = (float)c->dt_evo;
    int n = 0;
    if (!(dt > 0.0f)) return 1;
    for (float t = 0.0f; t <= te; t += dt) n++;
    return n;
}

/* propagateKSpaceRealTime (wigner_kernel.cu:137-151): exp(+i k^2/2 dt/2) (Q1) */
void orc_kinetic_rt(const orc_calc *c, const double *k2, orc_c *a, orc_c *b, int ne, double dt)
{
    long n = (long)c->cells * ne;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double ang = k2[i % c->cells] * dt / 2.0;
        orc_c p = cmk(cos(ang), sin(ang));
        a[i] = cmul(a[i], p);
        b[i] = cmul(b[i], p);
    }
}

/* propagateXSpaceTwoComponent (wigner_kernel.cu:184-229) */
void orc_nonlinear_rt(const orc_calc *c, orc_c *aa, orc_c *bb, int ne, double dt)
{
    long n = (long)c->cells * ne;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double V = orc_potential(c, (int)(i % c->cells));
        orc_c a = aa[i], b = bb[i], a0 = a, b0 = b;
        orc_c da = cmk(0, 0), db = cmk(0, 0);
        for (int it = 0; it < c->itmax; it++) {
            double na = cnorm(a), nb = cnorm(b);
            orc_c pa = cmk(-(c->l111 * na * na + c->l12 * nb) / 2.0, V + c->g11 * na + c->g12 * nb);
            orc_c pb = cmk(-(c->l22 * nb + c->l12 * na) / 2.0, V + c->g22 * nb + c->g12 * na - c->detuning);
            da = cexp_(cscale(pa, dt / 2.0));
            db = cexp_(cscale(pb, dt / 2.0));
            a = cmul(a0, da);
            b = cmul(b0, db);
        }
        aa[i] = cmul(a, da);
        bb[i] = cmul(b, db);
    }
}

/* propagate (wigner.cu:178-199), state in k-space */
void orc_propagate(const orc_calc *c, orc_c *a, orc_c *b, int ne, double dt, int n_steps)
{
    double *k2 = malloc(sizeof(double) * c->cells);
    orc_ktable(c, k2);
    for (int s = 0; s < n_steps; s++) {
        orc_kinetic_rt(c, k2, a, b, ne, dt);
        orc_to_xspace(c, a, ne);
        orc_to_xspace(c, b, ne);
        orc_nonlinear_rt(c, a, b, ne, dt);
        orc_to_kspace(c, a, ne);
        orc_to_kspace(c, b, ne);
        orc_kinetic_rt(c, k2, a, b, ne, dt);
    }
    free(k2);
}

/* applyHalfPiPulse (wigner_kernel.cu:320-329) */
void orc_half_pi_pulse(orc_c *a, orc_c *b, long n)
{
    double r = 1.0 / sqrt(2.0);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        orc_c a0 = a[i], b0 = b[i];
        a[i] = cscale(cadd(a0, cmul(b0, cmk(0, -1))), r);
        b[i] = cscale(cadd(cmul(a0, cmk(0, -1)), b0), r);
    }
}

/* applyPiPulse (wigner_kernel.cu:332-341) */
void orc_pi_pulse(orc_c *a, orc_c *b, long n)
{
    for (long i = 0; i < n; i++) {
        orc_c a0 = a[i], b0 = b[i];
        a[i] = cmul(b0, cmk(0, -1));
        b[i] = cmul(a0, cmk(0, -1));
    }
}

/* initEvolution (wigner.cu:202-227): returns the k-space state after the pulse */
void orc_init_evolution(const orc_calc *c, const orc_c *gs_x, uint64_t seed, int equilibrate,
                        orc_c *a, orc_c *b)
{
    orc_init_ensembles(c, gs_x, seed, a, b);
    orc_to_kspace(c, a, c->ne);
    orc_to_kspace(c, b, c->ne);
    if (equilibrate) orc_propagate(c, a, b, c->ne, c->dt_evo, orc_equilibration_steps(c));
    orc_half_pi_pulse(a, b, (long)c->cells * c->ne);
}

/* ------------------------------------------------------------------ */
/* Observables                                                          */

/* halfPiRotate (wigner_kernel.cu:344-360) followed by reduce: sums of the
 * vacuum-subtracted densities (density(), wigner_kernel.cu:283-286).  */
void orc_rotate_sums(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne, double alpha,
                     double *N1, double *N2)
{
    long n = (long)c->cells * ne;
    double r = 1.0 / sqrt(2.0), ca = cos(alpha), sa = sin(alpha);
    double vac = c->V / (2.0 * c->dx * c->dy * c->dz);
    double s1 = 0, s2 = 0;
#pragma omp parallel for reduction(+:s1,s2) schedule(static)
    for (long i = 0; i < n; i++) {
        orc_c a0 = a_x[i], b0 = b_x[i];
        orc_c an = cscale(cadd(a0, cmul(b0, cmk(sa, -ca))), r);
        orc_c bn = cscale(cadd(cmul(a0, cmk(-sa, -ca)), b0), r);
        s1 += cnorm(an) - vac;
        s2 += cnorm(bn) - vac;
    }
    *N1 = s1; *N2 = s2;
}

void orc_traj_sums(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne, double *sums)
{
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; e++) {
        double saa = 0, sbb = 0, sre = 0, sim = 0;
        for (int i = 0; i < c->cells; i++) {
            long idx = (long)e * c->cells + i;
            orc_c a = a_x[idx], b = b_x[idx];
            saa += cnorm(a);
            sbb += cnorm(b);
            sre += a.re * b.re + a.im * b.im;   /* Re conj(a) b */
            sim += a.re * b.im - a.im * b.re;   /* Im conj(a) b */
        }
        sums[4 * e] = saa; sums[4 * e + 1] = sbb; sums[4 * e + 2] = sre; sums[4 * e + 3] = sim;
    }
}

void orc_cell_density_sums(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne,
                           double alpha, double *dens_a, double *dens_b)
{
    double r = 1.0 / sqrt(2.0), ca = cos(alpha), sa = sin(alpha);
    double vac = c->V / (2.0 * c->dx * c->dy * c->dz);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c->cells; i++) {
        double s1 = 0, s2 = 0;
        for (int e = 0; e < ne; e++) {
            long idx = (long)e * c->cells + i;
            orc_c a0 = a_x[idx], b0 = b_x[idx];
            orc_c an = cscale(cadd(a0, cmul(b0, cmk(sa, -ca))), r);
            orc_c bn = cscale(cadd(cmul(a0, cmk(-sa, -ca)), b0), r);
            s1 += cnorm(an) - vac;
            s2 += cnorm(bn) - vac;
        }
        dens_a[i] = s1; dens_b[i] = s2;
    }
}

/* getVisibility / getComponentRatio (wigner.cu:257-277) */
double orc_visibility_grid(const orc_calc *c, const orc_c *a_x, const orc_c *b_x, int ne)
{
    double vmax = 0;
    for (float alpha = 0.0f; alpha < 2.0f * (float)M_PI; alpha += 0.5f) {
        double n1, n2;
        orc_rotate_sums(c, a_x, b_x, ne, (double)alpha, &n1, &n2);
        n1 = fabs(n1); n2 = fabs(n2);
        double ratio = (n1 - n2) / (n1 + n2);
        if (fabs(ratio) > vmax) vmax = fabs(ratio);
    }
    return vmax;
}

/* ------------------------------------------------------------------ */
/* 2D CPU program (bec_wigner_2d.cpp/wigner.cpp)                        */

typedef struct {
    int nx, ny, n;
    double dx, dy;
    double *prop, *pG, *pa, *pb;
    orc_c *propG, *propW;
} tables2d;

/* fill_propagation_matrices (wigner.cpp:154-194) */
static void tables2d_init(const orc2d_params *p, tables2d *t)
{
    t->nx = p->nvx; t->ny = p->nvy; t->n = p->nvx * p->nvy;
    t->dx = 2.0 * p->xmax / (p->nvx - 1);
    t->dy = 2.0 * p->ymax / (p->nvy - 1);
    double dkx = M_PI / p->xmax, dky = M_PI / p->ymax;
    t->prop = malloc(sizeof(double) * t->n); t->pG = malloc(sizeof(double) * t->n);
    t->pa = malloc(sizeof(double) * t->n); t->pb = malloc(sizeof(double) * t->n);
    t->propG = malloc(sizeof(orc_c) * t->n); t->propW = malloc(sizeof(orc_c) * t->n);
    for (int i = 0; i < p->nvx; i++) {
        double x = -p->xmax + t->dx * i;
        double kx = (2 * i > p->nvx) ? dkx * (i - p->nvx) : dkx * i;
        for (int j = 0; j < p->nvy; j++) {
            double y = -p->ymax + t->dy * j;
            double ky = (2 * j > p->nvy) ? dky * (j - p->nvy) : dky * j;
            int idx = j * p->nvx + i;
            double r2 = x * x + y * y;
            t->prop[idx] = -kx * kx - ky * ky;
            t->pG[idx] = p->mu - r2 / 4.0;
            t->pa[idx] = p->mu - r2 / 4.0 + 0.5 * p->E;
            t->pb[idx] = p->mu - r2 / 4.0 - 0.5 * p->E;
            t->propG[idx] = cmk(exp(t->prop[idx] * p->dtG / 2.0), 0.0);
            /* exp(prop * (-i dtW/2)) */
            double ang = -t->prop[idx] * p->dtW / 2.0;
            t->propW[idx] = cmk(cos(ang), sin(ang));
        }
    }
}

static void tables2d_free(tables2d *t)
{
    free(t->prop); free(t->pG); free(t->pa); free(t->pb); free(t->propG); free(t->propW);
}

static void fft2d(orc_c *d, int nx, int ny, int sign) { orc_fft_axes(d, 1, ny, nx, 1, sign, 3); }

static void to_k2d(orc_c *d, int nx, int ny)
{
    fft2d(d, nx, ny, +1);
    for (int s = 0; s < nx * ny; s++) d[s] = cscale(d[s], 1.0 / (nx * ny));
}

/* residual of (-lap + r^2/4 + g11|psi|^2) psi = mu psi, relative to mu */
double orc2d_residual(const orc2d_params *p, const orc_c *psi, double *mu_out)
{
    tables2d t;
    tables2d_init(p, &t);
    orc_c *h = malloc(sizeof(orc_c) * t.n);
    memcpy(h, psi, sizeof(orc_c) * t.n);
    to_k2d(h, t.nx, t.ny);
    for (int s = 0; s < t.n; s++) h[s] = cscale(h[s], -t.prop[s]);
    fft2d(h, t.nx, t.ny, -1);
    double num = 0, den = 0;
    for (int s = 0; s < t.n; s++) {
        double v = (p->mu - t.pG[s]) + p->g11 * cnorm(psi[s]);
        h[s] = cadd(h[s], cscale(psi[s], v));
        num += psi[s].re * h[s].re + psi[s].im * h[s].im;
        den += cnorm(psi[s]);
    }
    double mu = num / den, r = 0;
    for (int s = 0; s < t.n; s++) r += cnorm(csub(h[s], cscale(psi[s], mu)));
    free(h);
    tables2d_free(&t);
    if (mu_out) *mu_out = mu;
    return sqrt(r / den) / fabs(mu);
}

/* Ground state of the 2D program's Hamiltonian H = -lap + r^2/4 + g11 |psi|^2
 * with dx dy sum |psi|^2 = n_atoms (the extension gs_rule = 2; the program's
 * own loop fixes mu instead, wigner.cpp:196-275): the preconditioned
 * imaginary-time gradient flow of the 3D solver,
 *   psi <- psi - tau P (H - mu) psi + beta (psi - psi_prev),  P = 1 / (alpha + k^2),
 * renormalised after every update, from the Thomas-Fermi profile
 * sqrt(max(mu - r^2/4, 0) / g11) (uniform when that is empty), alpha = mu,
 * until the residual ||H psi - mu psi|| / (|mu| ||psi||) is below gs_tol.
 * Writes the x-space state into a0; returns the iterations taken. */
static int gs2d_flow(const orc2d_params *p, const tables2d *t, orc_c *a0, double *res_out)
{
    int n = t->n, nx = t->nx, ny = t->ny;
    orc_c *pv = malloc(sizeof(orc_c) * n), *h = malloc(sizeof(orc_c) * n);
    const double alpha = p->mu > 0 ? p->mu : 1.0;
    int any = 0;
    for (int s = 0; s < n; s++) {
        double v = t->pG[s];   /* mu - r^2/4 */
        v = v > 0 ? sqrt(v / p->g11) : 0.0;
        any |= v > 0;
        a0[s] = cmk(v, 0.0);
    }
    if (!any)
        for (int s = 0; s < n; s++) a0[s] = cmk(1.0, 0.0);
    double nn = 0;
    for (int s = 0; s < n; s++) nn += cnorm(a0[s]);
    double sc = sqrt(p->n_atoms / (nn * t->dx * t->dy));
    for (int s = 0; s < n; s++) { a0[s] = cscale(a0[s], sc); pv[s] = a0[s]; }
    int it = 0;
    double res = 0;
    for (;;) {
        /* h = H a0 */
        memcpy(h, a0, sizeof(orc_c) * n);
        to_k2d(h, nx, ny);
        for (int s = 0; s < n; s++) h[s] = cscale(h[s], -t->prop[s]);
        fft2d(h, nx, ny, -1);
        double num = 0, den = 0;
        for (int s = 0; s < n; s++) {
            double v = (p->mu - t->pG[s]) + p->g11 * cnorm(a0[s]);
            h[s] = cadd(h[s], cscale(a0[s], v));
            num += a0[s].re * h[s].re + a0[s].im * h[s].im;
            den += cnorm(a0[s]);
        }
        double mu = num / den, r = 0;
        for (int s = 0; s < n; s++) {
            h[s] = csub(h[s], cscale(a0[s], mu));
            r += cnorm(h[s]);
        }
        res = sqrt(r / den) / fabs(mu);
        if (!(res == res) || res < p->gs_tol || it >= p->gs_max_iter) break;
        /* d = P (H - mu) a0 */
        to_k2d(h, nx, ny);
        for (int s = 0; s < n; s++) h[s] = cscale(h[s], 1.0 / (alpha - t->prop[s]));
        fft2d(h, nx, ny, -1);
        nn = 0;
        for (int s = 0; s < n; s++) {
            orc_c cur = a0[s];
            orc_c nw = cadd(csub(cur, cscale(h[s], p->gs_tau)), cscale(csub(cur, pv[s]), p->gs_beta));
            pv[s] = cur;
            a0[s] = nw;
            nn += cnorm(nw);
        }
        sc = sqrt(p->n_atoms / (nn * t->dx * t->dy));
        for (int s = 0; s < n; s++) a0[s] = cscale(a0[s], sc);
        it++;
    }
    free(pv); free(h);
    if (res_out) *res_out = res;
    return it;
}

/* calculate_steady_state (wigner.cpp:196-275) + main_loop (wigner.cpp:278-477) */
int orc2d_run(const orc2d_params *p, double *m, double *var, double *vis, orc_c *gs_out,
              double *avn, orc2d_report *rep)
{
    tables2d t;
    tables2d_init(p, &t);
    int n = t.n, nx = t.nx, ny = t.ny;
    orc_c *a = malloc(sizeof(orc_c) * n), *b = malloc(sizeof(orc_c) * n);
    orc_c *a1 = malloc(sizeof(orc_c) * n), *a0 = malloc(sizeof(orc_c) * n);
    orc_c *prev = malloc(sizeof(orc_c) * n);
    double *da = malloc(sizeof(double) * n);

    /* ground state: uniform n0 start, FFT to k-space */
    for (int s = 0; s < n; s++) a[s] = cmk(p->n0, 0.0);
    to_k2d(a, nx, ny);
    int iters = 0;
    double res = -1;
    int total = (p->gs_rule == 0) ? p->ntG - 1 : p->gs_rule == 1 ? p->gs_max_iter : 0;
    if (p->gs_rule == 2) {   /* the flow works in x-space; a is brought back to k-space for the shared tail */
        iters = gs2d_flow(p, &t, a, &res);
        to_k2d(a, nx, ny);
    }
    for (int it = 0; it < total; it++) {
        for (int s = 0; s < n; s++) a[s] = cmul(a[s], t.propG[s]);
        fft2d(a, nx, ny, -1);
        memcpy(a1, a, sizeof(orc_c) * n);
        for (int iter = 0; iter < p->itmax; iter++)
            for (int s = 0; s < n; s++) {
                da[s] = exp(p->dtG / 2.0 * (t.pG[s] - p->g11 * cnorm(a[s])));
                a[s] = cscale(a1[s], da[s]);
            }
        for (int s = 0; s < n; s++) a[s] = cscale(a[s], da[s]);
        to_k2d(a, nx, ny);
        for (int s = 0; s < n; s++) a[s] = cmul(a[s], t.propG[s]);
        iters++;
        if (p->gs_rule == 1) {
            memcpy(prev, a, sizeof(orc_c) * n);
            fft2d(prev, nx, ny, -1);
            res = orc2d_residual(p, prev, NULL);
            if (res < p->gs_tol) break;
        }
    }
    double ntot = 0;
    for (int s = 0; s < n; s++) ntot += cnorm(a[s]);
    ntot *= t.dx * t.dy * nx * ny;
    fft2d(a, nx, ny, -1);
    memcpy(a0, a, sizeof(orc_c) * n);
    if (gs_out) memcpy(gs_out, a0, sizeof(orc_c) * n);
    if (rep) {
        rep->N0 = ntot;
        rep->gs_iterations = iters;
        rep->gs_residual = orc2d_residual(p, a0, NULL);
    }

    int nm = p->n_meas;
    double V = (p->Va + p->Vb) / 2.0, r2 = sqrt(2.0);
    for (int i = 0; i < 2 * nm; i++) { m[i] = 0; var[i] = 0; }
    if (vis) for (int i = 0; i < nm; i++) vis[i] = 0;
    if (avn) for (int s = 0; s < n; s++) avn[s] = 0;
    double nd = p->noise_dev / 0.5;
    for (int j = 0; j < p->ne; j++) {
        for (int s = 0; s < n; s++) {
            double z[4];
            orc_noise_cell(p->seed, (uint32_t)j, (uint32_t)s, 0u, z);
            a[s] = cmk(a0[s].re + p->Va * nd * z[0], a0[s].im + p->Va * nd * z[1]);
            b[s] = cmk(p->Vb * nd * z[2], p->Vb * nd * z[3]);
        }
        to_k2d(a, nx, ny);
        to_k2d(b, nx, ny);
        for (int nn = -p->ntW; nn < nm; nn++) {
            if (nn > -p->ntW) {
                if (nn == 0) {   /* first pi/2 pulse (wigner.cpp:341-353) */
                    for (int s = 0; s < n; s++) {
                        orc_c bo = b[s];
                        b[s] = cscale(csub(b[s], cmul(cmk(0, 1), a[s])), 1.0 / r2);
                        a[s] = cscale(csub(a[s], cmul(cmk(0, 1), bo)), 1.0 / r2);
                    }
                }
                for (int s = 0; s < n; s++) { a[s] = cmul(a[s], t.propW[s]); b[s] = cmul(b[s], t.propW[s]); }
                fft2d(a, nx, ny, -1);
                fft2d(b, nx, ny, -1);
                for (int s = 0; s < n; s++) {
                    double na = cnorm(a[s]), nb = cnorm(b[s]);
                    double dA = -p->g11 * na - p->g12 * nb + t.pa[s];
                    double dB = -p->g22 * nb - p->g12 * na + t.pb[s];
                    a[s] = cmul(a[s], cexp_(cmk(0, -p->dtW * dA)));
                    b[s] = cmul(b[s], cexp_(cmk(0, -p->dtW * dB)));
                }
                to_k2d(a, nx, ny);
                to_k2d(b, nx, ny);
                for (int s = 0; s < n; s++) { a[s] = cmul(a[s], t.propW[s]); b[s] = cmul(b[s], t.propW[s]); }
            }
            if (nn >= 0) {   /* second pulse U^dagger and measurement (wigner.cpp:393-425) */
                double sa = 0, sb = 0, cre = 0, cim = 0;
                for (int s = 0; s < n; s++) {
                    orc_c b1 = cscale(cadd(b[s], cmul(cmk(0, 1), a[s])), 1.0 / r2);
                    orc_c aa = cscale(cadd(a[s], cmul(cmk(0, 1), b[s])), 1.0 / r2);
                    sa += cnorm(aa);
                    sb += cnorm(b1);
                    cre += a[s].re * b[s].re + a[s].im * b[s].im;
                    cim += a[s].re * b[s].im - a[s].im * b[s].re;
                }
                double N1 = n * (sa * t.dx * t.dy - 0.5 * V);
                double N2 = n * (sb * t.dx * t.dy - 0.5 * V);
                m[nn] += N1; m[nn + nm] += N2;
                var[nn] += N1 * N1; var[nn + nm] += N2 * N2;
                if (vis) {
                    double na = 0, nb = 0;
                    for (int s = 0; s < n; s++) { na += cnorm(a[s]); nb += cnorm(b[s]); }
                    vis[nn] += 2.0 * sqrt(cre * cre + cim * cim) / (na + nb);
                }
            }
        }
        if (avn) {
            memcpy(a1, a, sizeof(orc_c) * n);
            fft2d(a1, nx, ny, -1);
            for (int s = 0; s < n; s++) avn[s] += cnorm(a1[s]) - 0.5 * V / t.dx / t.dy;
        }
    }
    for (int i = 0; i < 2 * nm; i++) {
        m[i] /= p->ne;
        var[i] = var[i] / p->ne - m[i] * m[i];
        var[i] = p->ne > 1 ? sqrt(fabs(var[i]) / (p->ne - 1)) : 0.0;
    }
    if (vis) for (int i = 0; i < nm; i++) vis[i] /= p->ne;
    if (avn) for (int s = 0; s < n; s++) avn[s] /= p->ne;

    free(a); free(b); free(a1); free(a0); free(prev); free(da);
    tables2d_free(&t);
    return 0;
}

/* ------------------------------------------------------------------ */
/* 1D CPU program (bec_wigner_2d.cpp/wigner_1d.cpp)                     */

/* The transform FFTW's plans compute (fftw_plan_dft_1d, any length), by its
   definition: exponent index (j k) mod n, accumulated over j in order. */
void orc_dft1d(orc_c *d, int n, int sign)
{
    orc_c *tw = malloc(sizeof(orc_c) * n), *out = malloc(sizeof(orc_c) * n);
    for (int m = 0; m < n; m++) {
        double ang = 2.0 * M_PI * (double)m / (double)n;
        tw[m] = cmk(cos(ang), sign * sin(ang));
    }
    for (int k = 0; k < n; k++) {
        orc_c acc = cmk(0.0, 0.0);
        long idx = 0;
        for (int j = 0; j < n; j++) {
            acc = cadd(acc, cmul(d[j], tw[idx]));
            idx += k;
            if (idx >= n) idx -= n;
        }
        out[k] = acc;
    }
    memcpy(d, out, sizeof(orc_c) * n);
    free(tw);
    free(out);
}

/* nv = 1 + 2 (int)(zmax / dz) (wigner_1d.cpp:118) */
int orc1d_nv(const orc1d_params *p) { return 1 + 2 * (int)(p->zmax / p->dz); }

typedef struct {
    int n;
    double *z, *k2, *pG, *pa, *pb, *propG;
    orc_c *propW;
} tables1d;

/* fill_propagation_matrices (wigner_1d.cpp:136-164) */
static void tables1d_init(const orc1d_params *p, tables1d *t)
{
    int n = orc1d_nv(p);
    double dk = M_PI / p->zmax;
    t->n = n;
    t->z = malloc(sizeof(double) * n); t->k2 = malloc(sizeof(double) * n);
    t->pG = malloc(sizeof(double) * n); t->pa = malloc(sizeof(double) * n);
    t->pb = malloc(sizeof(double) * n); t->propG = malloc(sizeof(double) * n);
    t->propW = malloc(sizeof(orc_c) * n);
    for (int j = 0; j < n; j++) {
        double z = -p->zmax + p->dz * j;
        double k = (2 * j > n) ? dk * (j - n) : dk * j;
        t->z[j] = z;
        t->k2[j] = k * k;
        t->pG[j] = p->mu - z * z / 4.0;
        t->pa[j] = p->mu - z * z / 4.0 + 0.5 * p->E;
        t->pb[j] = p->mu - z * z / 4.0 - 0.5 * p->E;
        t->propG[j] = exp(-k * k * p->dt / 2.0);
        /* exp(prop (-i dt/2)) with prop = -k^2 */
        t->propW[j] = cmk(cos(k * k * p->dt / 2.0), sin(k * k * p->dt / 2.0));
    }
}

static void tables1d_free(tables1d *t)
{
    free(t->z); free(t->k2); free(t->pG); free(t->pa); free(t->pb); free(t->propG); free(t->propW);
}

/* x -> k: FFTW_BACKWARD (+i) and 1/n */
static void to_k1d(orc_c *d, int n)
{
    orc_dft1d(d, n, +1);
    for (int j = 0; j < n; j++) d[j] = cscale(d[j], 1.0 / n);
}

/* residual of (-d^2/dz^2 + z^2/4 + g11|psi|^2) psi = mu psi, relative to mu */
double orc1d_residual(const orc1d_params *p, const orc_c *psi, double *mu_out)
{
    tables1d t;
    tables1d_init(p, &t);
    int n = t.n;
    orc_c *h = malloc(sizeof(orc_c) * n);
    memcpy(h, psi, sizeof(orc_c) * n);
    to_k1d(h, n);
    for (int j = 0; j < n; j++) h[j] = cscale(h[j], t.k2[j]);
    orc_dft1d(h, n, -1);
    double num = 0, den = 0;
    for (int j = 0; j < n; j++) {
        double v = (p->mu - t.pG[j]) + p->g11 * cnorm(psi[j]);
        h[j] = cadd(h[j], cscale(psi[j], v));
        num += psi[j].re * h[j].re + psi[j].im * h[j].im;
        den += cnorm(psi[j]);
    }
    double mu = num / den, r = 0;
    for (int j = 0; j < n; j++) r += cnorm(csub(h[j], cscale(psi[j], mu)));
    free(h);
    tables1d_free(&t);
    if (mu_out) *mu_out = mu;
    return sqrt(r / den) / fabs(mu);
}

/* calculate_steady_state (wigner_1d.cpp:166-243) + main_loop (wigner_1d.cpp:245-445).
   The program's rand() Box-Muller deviates are replaced by the Philox stream of
   orc_noise_cell (trajectory j, cell i): (z0, z1) for a, (z2, z3) for b. */
int orc1d_run(const orc1d_params *p, double *m, double *var, double *vis, orc_c *gs_out, double *avn,
              double *ntot, orc1d_report *rep)
{
    tables1d t;
    tables1d_init(p, &t);
    int n = t.n, ntW = p->ntW;
    orc_c *a = malloc(sizeof(orc_c) * n), *b = malloc(sizeof(orc_c) * n);
    orc_c *a1 = malloc(sizeof(orc_c) * n), *a0 = malloc(sizeof(orc_c) * n);
    double *da = malloc(sizeof(double) * n);

    /* ground state: uniform n0 in x-space, to k-space (:183-187) */
    for (int j = 0; j < n; j++) a[j] = cmk(p->n0, 0.0);
    to_k1d(a, n);
    double N0 = 0;
    for (int it = 0; it < p->ntG; it++) {
        if (it > 0) {
            for (int j = 0; j < n; j++) a[j] = cscale(a[j], t.propG[j]);
            orc_dft1d(a, n, -1);
            memcpy(a1, a, sizeof(orc_c) * n);
            for (int j = 0; j < n; j++) da[j] = 0.0;
            for (int iter = 0; iter < p->itmax; iter++)
                for (int j = 0; j < n; j++) {
                    da[j] = exp(p->dt / 2.0 * (t.pG[j] - p->g11 * cnorm(a[j])));
                    a[j] = cscale(a1[j], da[j]);
                }
            for (int j = 0; j < n; j++) a[j] = cscale(a[j], da[j]);
            to_k1d(a, n);
            for (int j = 0; j < n; j++) a[j] = cscale(a[j], t.propG[j]);
        }
        double s = 0;
        for (int j = 0; j < n; j++) s += cnorm(a[j]);
        s *= p->dz * n;
        if (ntot) ntot[it] = s;
        N0 = s;
    }
    orc_dft1d(a, n, -1);
    memcpy(a0, a, sizeof(orc_c) * n);
    if (gs_out) memcpy(gs_out, a0, sizeof(orc_c) * n);
    if (rep) {
        rep->N0 = N0;
        rep->nv = n;
        rep->gs_residual = orc1d_residual(p, a0, &rep->gs_mu);
    }

    double V = (p->Va + p->Vb) / 2.0, r2 = sqrt(2.0);
    for (int i = 0; i < 2 * ntW; i++) { m[i] = 0; var[i] = 0; }
    if (vis) for (int i = 0; i < ntW; i++) vis[i] = 0;
    if (avn) for (int j = 0; j < n; j++) avn[j] = 0;
    double nd = p->noise_dev / 0.5;
    for (int e = 0; e < p->ne; e++) {
        for (int j = 0; j < n; j++) {   /* :281-297 */
            double z[4];
            orc_noise_cell(p->seed, (uint32_t)e, (uint32_t)j, 0u, z);
            a[j] = cmk(a0[j].re + p->Va * nd * z[0], a0[j].im + p->Va * nd * z[1]);
            b[j] = cmk(p->Vb * nd * z[2], p->Vb * nd * z[3]);
        }
        to_k1d(a, n);
        to_k1d(b, n);
        for (int nn = -ntW; nn < ntW; nn++) {
            if (nn > -ntW) {
                if (nn == 0) {   /* pi/2 pulse (:308-318) */
                    for (int j = 0; j < n; j++) {
                        orc_c bo = b[j];
                        b[j] = cscale(csub(b[j], cmul(cmk(0, 1), a[j])), 1.0 / r2);
                        a[j] = cscale(csub(a[j], cmul(cmk(0, 1), bo)), 1.0 / r2);
                    }
                }
                for (int j = 0; j < n; j++) { a[j] = cmul(a[j], t.propW[j]); b[j] = cmul(b[j], t.propW[j]); }
                orc_dft1d(a, n, -1);
                orc_dft1d(b, n, -1);
                for (int j = 0; j < n; j++) {   /* single-shot nonlinear step (:331-337) */
                    double na = cnorm(a[j]), nb = cnorm(b[j]);
                    double dA = -p->g11 * na - p->g12 * nb + t.pa[j];
                    double dB = -p->g22 * nb - p->g12 * na + t.pb[j];
                    a[j] = cmul(a[j], cexp_(cmk(0, -p->dt * dA)));
                    b[j] = cmul(b[j], cexp_(cmk(0, -p->dt * dB)));
                }
                to_k1d(a, n);
                to_k1d(b, n);
                for (int j = 0; j < n; j++) { a[j] = cmul(a[j], t.propW[j]); b[j] = cmul(b[j], t.propW[j]); }
            }
            if (nn >= 0) {   /* second pulse U^dagger and measurement in k-space (:358-391) */
                double sa = 0, sb = 0, cre = 0, cim = 0, na = 0, nb = 0;
                for (int j = 0; j < n; j++) {
                    orc_c b1 = cscale(cadd(b[j], cmul(cmk(0, 1), a[j])), 1.0 / r2);
                    orc_c aa = cscale(cadd(a[j], cmul(cmk(0, 1), b[j])), 1.0 / r2);
                    sa += cnorm(aa);
                    sb += cnorm(b1);
                    cre += a[j].re * b[j].re + a[j].im * b[j].im;
                    cim += a[j].re * b[j].im - a[j].im * b[j].re;
                    na += cnorm(a[j]);
                    nb += cnorm(b[j]);
                }
                double N1 = n * (sa * p->dz - 0.5 * V);
                double N2 = n * (sb * p->dz - 0.5 * V);
                m[nn] += N1; m[nn + ntW] += N2;
                var[nn] += N1 * N1; var[nn + ntW] += N2 * N2;
                if (vis) vis[nn] += 2.0 * sqrt(cre * cre + cim * cim) / (na + nb);
            }
        }
        if (avn) {   /* :396-400 */
            memcpy(a1, a, sizeof(orc_c) * n);
            orc_dft1d(a1, n, -1);
            for (int j = 0; j < n; j++) avn[j] += cnorm(a1[j]) - 0.5 * V / p->dz;
        }
    }
    /* :405-413, as written: no absolute value before the square root */
    for (int i = 0; i < 2 * ntW; i++) {
        m[i] = m[i] / p->ne;
        var[i] = var[i] / p->ne - m[i] * m[i];
        var[i] = sqrt(var[i] / (p->ne - 1));
    }
    if (vis) for (int i = 0; i < ntW; i++) vis[i] /= p->ne;
    if (avn) for (int j = 0; j < n; j++) avn[j] /= p->ne;

    free(a); free(b); free(a1); free(a0); free(da);
    tables1d_free(&t);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Loss-induced Wigner noise during the evolution                       */

/* Four standard normal deviates for the loss noise of step `step`: Philox
 * counter {cell, traj, step, 1 + group}, key = seed, Box-Muller on 24-bit
 * uniforms u = ((x >> 8) + 1) 2^-24. */
void orc_loss_noise_cell(uint64_t seed, uint32_t traj, uint32_t cell, uint32_t step, uint32_t group, double out[4])
{
    uint32_t ctr[4] = {cell, traj, step, 1u + group};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    for (int p = 0; p < 2; p++) {
        double u1 = (double)((r[2 * p] >> 8) + 1u) / 16777216.0;
        double u2 = (double)((r[2 * p + 1] >> 8) + 1u) / 16777216.0;
        double rad = sqrt(-2.0 * log(u1));
        out[2 * p] = rad * cos(2.0 * M_PI * u2);
        out[2 * p + 1] = rad * sin(2.0 * M_PI * u2);
    }
}

/* The Ito increment of the loss noise for one cell, from the state a0, b0 at
 * the start of the step.  The truncated Wigner equation of the report
 * (docs/wigner_simulation.report/report.tex:624-656; poster
 * docs/icap2010.poster/poster.tex:245-290), at the order of the drift the
 * kernels integrate (-(l111 n1^2 + l12 n2)/2 and -(l12 n1 + l22 n2)/2),
 * diffuses with
 *   D11 = 3/2 l111 n1^2 + 1/2 l12 n2,  D22 = 1/2 l12 n1 + l22 n2,
 *   D12 = 1/2 l12 a0 conj(b0),
 * i.e. <d psi_i d psi_j*> = D_ij dt / dV per cell.  The increment is B dW with
 * B the lower Cholesky factor of D and dW_k = (z_2k + i z_2k+1) sqrt(dt/(2 dV)). */
static void loss_increment(double l111, double l12, double l22, orc_c a0, orc_c b0, const double z[4], double w,
                           orc_c *da, orc_c *db)
{
    double n1 = cnorm(a0), n2 = cnorm(b0);
    double D11 = 1.5 * l111 * n1 * n1 + 0.5 * l12 * n2;
    double D22 = 0.5 * l12 * n1 + l22 * n2;
    orc_c w1 = cmk(z[0] * w, z[1] * w), w2 = cmk(z[2] * w, z[3] * w);
    double b11 = sqrt(D11), b22;
    orc_c b21 = cmk(0.0, 0.0);
    if (D11 > 0) {
        /* conj(D12) / b11 = l12 conj(a0) b0 / (2 b11) */
        orc_c cab = cmk(a0.re * b0.re + a0.im * b0.im, a0.re * b0.im - a0.im * b0.re);
        b21 = cscale(cab, 0.5 * l12 / b11);
        double r = D22 - 0.25 * l12 * l12 * n1 * n2 / D11;
        b22 = sqrt(r > 0 ? r : 0.0);
    } else {
        b22 = sqrt(D22);
    }
    *da = cscale(w1, b11);
    *db = cadd(cmul(b21, w1), cscale(w2, b22));
}

void orc_nonlinear_rt_noise(const orc_calc *c, orc_c *aa, orc_c *bb, int ne, double dt, uint64_t seed,
                            uint32_t step, uint32_t group)
{
    long n = (long)c->cells * ne;
    double w = sqrt(dt / (2.0 * c->dx * c->dy * c->dz));
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        double V = orc_potential(c, (int)(i % c->cells));
        orc_c a0 = aa[i], b0 = bb[i], a = a0, b = b0;
        orc_c fa = cmk(0, 0), fb = cmk(0, 0), da, db;
        /* the deterministic step of orc_nonlinear_rt */
        for (int it = 0; it < c->itmax; it++) {
            double na = cnorm(a), nb = cnorm(b);
            orc_c pa = cmk(-(c->l111 * na * na + c->l12 * nb) / 2.0, V + c->g11 * na + c->g12 * nb);
            orc_c pb = cmk(-(c->l22 * nb + c->l12 * na) / 2.0, V + c->g22 * nb + c->g12 * na - c->detuning);
            fa = cexp_(cscale(pa, dt / 2.0));
            fb = cexp_(cscale(pb, dt / 2.0));
            a = cmul(a0, fa);
            b = cmul(b0, fb);
        }
        /* plus the noise increment evaluated at the start of the step */
        double z[4];
        orc_loss_noise_cell(seed, (uint32_t)(i / c->cells), (uint32_t)(i % c->cells), step, group, z);
        loss_increment(c->l111, c->l12, c->l22, a0, b0, z, w, &da, &db);
        aa[i] = cadd(cmul(a, fa), da);
        bb[i] = cadd(cmul(b, fb), db);
    }
}

/* propagate with the loss noise: steps step0, step0 + 1, ... */
void orc_propagate_noise(const orc_calc *c, orc_c *a, orc_c *b, int ne, double dt, int n_steps, uint64_t seed,
                         uint32_t step0, uint32_t group)
{
    double *k2 = malloc(sizeof(double) * c->cells);
    orc_ktable(c, k2);
    for (int s = 0; s < n_steps; s++) {
        orc_kinetic_rt(c, k2, a, b, ne, dt);
        orc_to_xspace(c, a, ne);
        orc_to_xspace(c, b, ne);
        orc_nonlinear_rt_noise(c, a, b, ne, dt, seed, step0 + (uint32_t)s, group);
        orc_to_kspace(c, a, ne);
        orc_to_kspace(c, b, ne);
        orc_kinetic_rt(c, k2, a, b, ne, dt);
    }
    free(k2);
}

/* calculateAverages (wigner.cu:305-337, the debug report; its smallReduce of
 * quirk Q6 fixed): on x-space states, out = {N, norm, var Re a, var Im a,
 * var Re b, var Im b} with N = (sum|a|^2 + sum|b|^2) dV / ne,
 * norm = N_model / ((sum|a|^2 + sum|b|^2) / ne), and each variance the
 * ensemble variance (1/ne) sum_j (x_j - mean)^2 of one quadrature per cell
 * (kernelAverage, wigner_kernel.cu:383-392), averaged over cells, times norm. */
void orc_averages(const orc_calc *c, const orc_c *a, const orc_c *b, int ne, double out[6])
{
    double sa = 0, sb = 0, v[4] = {0, 0, 0, 0};
    for (int i = 0; i < c->cells; i++) {
        double m[4] = {0, 0, 0, 0};
        for (int e = 0; e < ne; e++) {
            long idx = (long)e * c->cells + i;
            m[0] += a[idx].re; m[1] += a[idx].im; m[2] += b[idx].re; m[3] += b[idx].im;
            sa += cnorm(a[idx]);
            sb += cnorm(b[idx]);
        }
        for (int k = 0; k < 4; k++) m[k] /= ne;
        double d[4] = {0, 0, 0, 0};
        for (int e = 0; e < ne; e++) {
            long idx = (long)e * c->cells + i;
            double x[4] = {a[idx].re, a[idx].im, b[idx].re, b[idx].im};
            for (int k = 0; k < 4; k++) d[k] += (x[k] - m[k]) * (x[k] - m[k]);
        }
        for (int k = 0; k < 4; k++) v[k] += d[k];
    }
    double dV = c->dx * c->dy * c->dz;
    double norm = (double)c->N / ((sa + sb) / ne);
    out[0] = (sa + sb) * dV / ne;
    out[1] = norm;
    for (int k = 0; k < 4; k++) out[2 + k] = v[k] / ((double)c->cells * ne) * norm;
}

/* A resonant pulse of area theta and phase phi:
 * U = [[cos(theta/2), -i e^{-i phi} sin(theta/2)], [-i e^{i phi} sin(theta/2), cos(theta/2)]]
 * (U(pi/2, 0), U(pi, 0): wigner_kernel.cu:320-341; U(pi/2, -alpha): halfPiRotate, :344-360). */
void orc_rotation(orc_c *a, orc_c *b, long n, double theta, double phi)
{
    double C = cos(0.5 * theta), S = sin(0.5 * theta);
    orc_c u01 = cmk(-sin(phi) * S, -cos(phi) * S), u10 = cmk(sin(phi) * S, -cos(phi) * S);
    for (long i = 0; i < n; i++) {
        orc_c a0 = a[i], b0 = b[i];
        a[i] = cadd(cscale(a0, C), cmul(u01, b0));
        b[i] = cadd(cmul(u10, a0), cscale(b0, C));
    }
}
