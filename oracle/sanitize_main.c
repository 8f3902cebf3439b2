/* Exercises the oracle's entry points on small lattices for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitizers.py):
 * any out-of-bounds access, use after free, leak or undefined operation aborts
 * the run with a report.  Test infrastructure, like the oracle itself. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bec_oracle.h"

static orc_model model(int nx, int nz, int ne)
{
    orc_model m;
    memset(&m, 0, sizeof(m));
    m.a11 = 100.4; m.a12 = 97.66; m.a22 = 95.0; m.m = 1.443160648e-25;
    m.fx = m.fy = 97.6; m.fz = 11.96; m.N = 30000; m.detuning = -41.0;
    m.gamma111 = 5.4e-30; m.gamma12 = 0.78e-13; m.gamma22 = 1.5e-14;
    m.Va = m.Vb = 1.0; m.nvx = m.nvy = nx; m.nvz = nz; m.itmax = 1;
    m.dt_steady = 1e-5; m.t_equilib = 0.0; m.dt_evo = 4e-5; m.ne = ne;
    return m;
}

int main(void)
{
    const int nx = 8, nz = 16, ne = 3;
    orc_model m = model(nx, nz, ne);
    orc_calc c;
    orc_derive(&m, &c);
    const int cells = c.cells;
    orc_c *gs = calloc(cells, sizeof(orc_c)), *a = calloc((size_t)cells * ne, sizeof(orc_c));
    orc_c *b = calloc((size_t)cells * ne, sizeof(orc_c)), *t = calloc(cells, sizeof(orc_c));
    double *k2 = calloc(cells, sizeof(double));
    orc_gs_report rep;
    orc_ground_state(&c, NULL, ORC_GS_RULE_REFERENCE, 1e-6, 500, 1, gs, &rep);
    double mu = 0;
    orc_residual(&c, gs, &mu);
    orc_dft3d_bruteforce(gs, t, nz, nx, nx, -1);
    orc_fft3d(t, nz, nx, nx, 1, +1);
    orc_ktable(&c, k2);
    orc_init_evolution(&c, gs, 0x5EED, 1, a, b);
    orc_propagate(&c, a, b, ne, c.dt_evo, 3);
    orc_propagate_noise(&c, a, b, ne, c.dt_evo, 2, 0x1055, 0, 0);
    orc_to_xspace(&c, a, ne);
    orc_to_xspace(&c, b, ne);
    double sums[4 * 3], out6[6], *da = calloc(cells, sizeof(double)), *db = calloc(cells, sizeof(double));
    orc_traj_sums(&c, a, b, ne, sums);
    orc_cell_density_sums(&c, a, b, ne, 0.3, da, db);
    orc_averages(&c, a, b, ne, out6);
    double vis = orc_visibility_grid(&c, a, b, ne);
    orc_rotation(a, b, (long)cells * ne, 1.1, 0.4);
    /* two-component ground state */
    orc_c *ga = calloc(cells, sizeof(orc_c)), *gb = calloc(cells, sizeof(orc_c));
    orc_gs2_report rep2;
    orc_ground_state2(&c, 0.5, NULL, NULL, 1e-6, 200, c.mu, 1.0, 0.85, ga, gb, &rep2);
    /* the 2D program, both ground-state forms, and the 1D program */
    orc2d_params p2;
    memset(&p2, 0, sizeof(p2));
    p2.nvx = 16; p2.nvy = 8; p2.xmax = 10; p2.ymax = 10; p2.g11 = 6.18; p2.g12 = 6.03; p2.g22 = 5.88;
    p2.E = 2.5; p2.mu = 26.8; p2.Va = 0.2; p2.Vb = 0.5; p2.dtG = 1.0 / 31; p2.ntG = 8; p2.itmax = 3;
    p2.dtW = 0.01; p2.ntW = 4; p2.n_meas = 3; p2.ne = 2; p2.n0 = 1; p2.seed = 7; p2.noise_dev = 0.5;
    p2.gs_rule = 0; p2.gs_tol = 1e-6; p2.gs_max_iter = 100; p2.n_atoms = 1000; p2.gs_tau = 0.5; p2.gs_beta = 0.6;
    double m2[6], v2[6], vis2[3], avn[16 * 8];
    orc_c gs2[16 * 8];
    orc2d_report r2;
    orc2d_run(&p2, m2, v2, vis2, gs2, avn, &r2);
    p2.gs_rule = 2;
    orc2d_run(&p2, m2, v2, vis2, gs2, avn, &r2);
    orc1d_params p1;
    memset(&p1, 0, sizeof(p1));
    p1.zmax = 10; p1.dz = 1; p1.g11 = 6.18; p1.g12 = 6.03; p1.g22 = 5.88; p1.E = 2.5; p1.mu = 26.8;
    p1.Va = 0.5; p1.Vb = 1; p1.dt = 0.01; p1.ntG = 10; p1.itmax = 3; p1.ntW = 4; p1.ne = 2; p1.n0 = 1;
    p1.seed = 3; p1.noise_dev = 0.25;
    const int nv = orc1d_nv(&p1);
    double *m1 = calloc(2 * p1.ntW, sizeof(double)), *v1 = calloc(2 * p1.ntW, sizeof(double));
    double *vis1 = calloc(p1.ntW, sizeof(double)), *avn1 = calloc(nv, sizeof(double));
    double *ntot1 = calloc(p1.ntG, sizeof(double));
    orc_c *gs1 = calloc(nv, sizeof(orc_c));
    orc1d_report r1;
    orc1d_run(&p1, m1, v1, vis1, gs1, avn1, ntot1, &r1);
    printf("sanitized run: gs iterations %d, mu %.6f, visibility %.6f, 2D N0 %.3f, 1D nv %d\n", rep.iterations, mu,
           vis, r2.N0, nv);
    free(gs); free(a); free(b); free(t); free(k2); free(da); free(db); free(ga); free(gb);
    free(m1); free(v1); free(vis1); free(avn1); free(ntot1); free(gs1);
    return 0;
}
