/*
 * oracle/sanitize_check.c -- drives every entry point of the FP64 oracle on
 * small cases, for the address and undefined-behaviour sanitizer build of
 * tests/test_sanitizers.py.
 *
 * TEST INFRASTRUCTURE: built by that test only, never part of the product.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "quill_oracle.h"

static int run_case(int solver, int pusher, int qed)
{
    const double PI = 3.141592653589793;
    const int nx = 24, ny = 14, nz = 12;
    oq_params p = {0};
    p.nx = nx; p.ny = ny; p.nz = nz;
    p.dx = 0.05 * 2 * PI; p.dy = 1.1 * p.dx; p.dz = p.dx;
    p.dt = (solver == OQ_FIVE_POINT ? 0.35 : 0.5) * p.dx;
    p.solver = solver; p.pusher = pusher; p.qed = qed;
    p.e_s = qed ? 2000.0 : 1e10;
    p.seed = 5;
    oq_state *s = oq_create(&p);
    /* plasma from the loader, plus positrons, photons and ions */
    if (oq_load_uniform(s, 2, 0.05, 0.3, 9, 1, nx) <= 0) return 1;
    int64_t m = 60;
    double *a = (double *)calloc((size_t)m * 7, sizeof(double));
    int32_t *id = (int32_t *)malloc((size_t)m * sizeof(int32_t));
    for (int64_t n = 0; n < m; ++n) {
        a[n] = 3 + (n % 17); a[m + n] = 3 + (n % 8) + 0.3; a[2 * m + n] = 3 + (n % 6) + 0.6;
        a[3 * m + n] = 30.0 * ((n % 3) - 1); a[4 * m + n] = 0.2; a[5 * m + n] = -0.1;
        a[6 * m + n] = 0.01;
        id[n] = 100000 + (int32_t)n;
    }
    oq_add_particles(s, m, a, a + m, a + 2 * m, a + 3 * m, a + 4 * m, a + 5 * m, a + 6 * m, 1.0, id);
    oq_add_particles(s, m, a, a + m, a + 2 * m, a + 3 * m, a + 4 * m, a + 5 * m, a + 6 * m, 0.0, id);
    oq_add_particles(s, m, a, a + m, a + 2 * m, a + 3 * m, a + 4 * m, a + 5 * m, a + 6 * m, 1.0 / 1836.0, id);
    /* fields: a laser initial condition and a source on each side */
    oq_init_laser_cos(s, 50.0, 20.0, 0.6 * PI, 0.8 * PI, 0.8 * PI, 0.5 * nx * p.dx, 1, -1.0, 0.2, 0.1, 0.0, 1, 0.1,
                      0, 0, 0);
    double L[16] = {1.0, 0.5, 0.4 * PI, 0.5 * PI, 0.5 * PI, -1.0, 1, -1.0, 0, 0, 0, 1, 0, 0, 0, 0};
    if (oq_add_laser_source(s, L, 0, 3)) return 2;
    L[5] = nx * p.dx + 1.0;
    L[7] = 1.0;
    if (oq_add_laser_source(s, L, 1, nx - 4)) return 3;
    oq_init_boundaries(s);
    oq_diag d[4];
    oq_step(s, 4, d);
    double ppd[4] = {0.3, 0.3, 0.3, 0.3}, icmr[1] = {1.0 / 1836.0};
    oq_pmerging(s, 1, ppd, 1, icmr);
    oq_pmerging(s, 2, ppd, 1, icmr);
    oq_moving_window(s);
    int32_t cells[6] = {nx - 3, 4, 5, nx - 3, 5, 5};
    double dens[2] = {0.05, 0.07};
    int npic[3] = {2, 1, 1};
    oq_fill_cells(s, -1.0, 2, cells, dens, npic, 0.1, 7);
    oq_step(s, 2, d);
    oq_compute_rho(s);
    double w = oq_max_w(s, 0, nx);
    int64_t np = oq_count(s);
    double *rec = (double *)malloc((size_t)(np ? np : 1) * 10 * sizeof(double));
    int32_t *ids = (int32_t *)malloc((size_t)(np ? np : 1) * sizeof(int32_t));
    oq_get_particles(s, rec, ids);
    printf("solver %d pusher %d qed %d: %lld particles, max w %.3g, energy %.3g\n", solver, pusher, qed,
           (long long)np, w, d[1].energy_f);
    free(rec); free(ids); free(a); free(id);
    oq_destroy(s);
    return isfinite(w) ? 0 : 4;
}

int main(void)
{
    int rc = 0;
    for (int solver = 0; solver < 3; ++solver) rc |= run_case(solver, solver == 0 ? OQ_BORIS : OQ_VAY, 0);
    rc |= run_case(OQ_FDTD, OQ_VAY, 1);
    /* the loader and a field evaluation of the laser source on their own */
    const int64_t n = oq_synth_plasma(10, 6, 6, 2, 0.1, 0.2, 3, 1, 10, NULL, NULL);
    double *rec8 = (double *)malloc((size_t)n * 8 * sizeof(double));
    int32_t *id = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    rc |= oq_synth_plasma(10, 6, 6, 2, 0.1, 0.2, 3, 1, 10, rec8, id) != 2 * 7 * 3 * 3;
    free(rec8);
    free(id);
    return rc;
}
