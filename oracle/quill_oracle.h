/*
 * oracle/quill_oracle.h -- FP64 CPU restatement of QUILL's per-step PIC loop
 * on one subdomain (n_sr = 1).
 *
 * TEST INFRASTRUCTURE.  This is the checker the GPU path is compared with; it
 * is never linked into, loaded by or called from the product (quill_amd/).
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * use it.
 *
 * Layout follows the reference containers (include/containers.h:17-52):
 * every field is an AoS array double[nx][ny][nz][3], k fastest.  Positions
 * are absolute, in cell units from the subdomain origin (src/main.cpp:2573,
 * src/advance.cpp:281-285).
 */
#ifndef QUILL_ORACLE_H
#define QUILL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OQ_FDTD = 0, OQ_NDFX = 1, OQ_FIVE_POINT = 2 };   /* containers.h:82-84 */
enum { OQ_VAY = 0, OQ_BORIS = 1 };                      /* containers.h:86-88 */
enum { OQ_E = 0, OQ_B = 1, OQ_J = 2, OQ_BE = 3 };

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, dt;        /* c/omega units (config value * 2 pi) */
    int solver;                   /* OQ_FDTD / OQ_NDFX / OQ_FIVE_POINT */
    int pusher;                   /* OQ_VAY / OQ_BORIS */
    int qed;                      /* photon emission + Breit-Wheeler + vacuum pairs */
    double e_s;                   /* Sauter field, mc omega / e units */
    double external_bz;           /* added to the gathered Bz (advance.cpp:136) */
    double q_vac;                 /* charge of vacuum-born macro-particles */
    uint64_t seed;                /* QED random stream */
    double norm_n;                /* compute_N multiplier (fout.cpp:82-84) */
    double norm_ef;               /* compute_energy field multiplier (fout.cpp:132) */
    double norm_ep;               /* compute_energy particle multiplier (fout.cpp:133-137) */
} oq_params;

typedef struct {
    double N_e, N_p, N_ph;                 /* fout.cpp:41-85 */
    int64_t n_e, n_p, n_ph, n_ion;         /* macro-particle counts */
    double energy_f, energy_e, energy_p, energy_ph, energy_ion;  /* fout.cpp:87-138 */
} oq_diag;

typedef struct oq_state oq_state;

oq_state *oq_create(const oq_params *p);
void oq_destroy(oq_state *s);
/* pointer to one field array double[nx][ny][nz][3] (OQ_E, OQ_B, OQ_J, OQ_BE) */
double *oq_field(oq_state *s, int which);
/* particles: absolute positions (cell units) and u = p/mc; g is computed as in
 * fill_cell_by_particles (init.cpp:461-464); place() semantics drop those
 * outside is_inside (plist.cpp:42-48).  Returns number kept. */
int64_t oq_add_particles(oq_state *s, int64_t n, const double *x, const double *y, const double *z,
                         const double *ux, const double *uy, const double *uz, const double *q,
                         double cmr, const int32_t *id);
int64_t oq_count(const oq_state *s);
/* copy out: x,y,z,ux,uy,uz,g,q,cmr,chi (10 doubles per particle, row-major) and ids */
void oq_get_particles(const oq_state *s, double *rec10, int32_t *id);

void oq_init_boundaries(oq_state *s);     /* solver->init_boundaries(); interpolate_be() */
void oq_interpolate_be(oq_state *s);      /* advance.cpp:20-83 */
void oq_maxwell(oq_state *s);             /* maxwell.cpp:10-22 */
void oq_padvance(oq_state *s);            /* advance.cpp:90-387 */
void oq_birth_from_vacuum(oq_state *s);   /* mc.cpp:142-196 */
/* pmerging (mc.cpp:202-462): mode 1 "ti" deletes every particle with
 * probability ppd[0]; mode 2 "nl" deletes a particle of species s with
 * probability ppd[s] / (1 + (2 |q| / <|q|>_s)^4) and divides the survivors'
 * charge by one minus that probability; both then rescale each species so
 * that its total charge is unchanged.  Species: 0 electrons, 1 positrons,
 * 2 photons, 3 + n ions of charge-to-mass ratio ion_cmr[n]. */
void oq_pmerging(oq_state *s, int mode, const double *ppd, int n_ion, const double *ion_cmr);
/* moving_window (advance.cpp:389-454): erase plane i = 0, move E, B, Be and
 * the particles one cell toward -x, zero plane nx - 1 */
void oq_moving_window(oq_state *s);
/* get_max_w (low-level.cpp:516-532): max of E^2 + B^2 over planes [ilo, ihi) */
double oq_max_w(const oq_state *s, int ilo, int ihi);
/* fill_cell_by_particles (init.cpp:402-478) with T = 0 on ncells cells (i j k
 * triples): npic[0] x npic[1] x npic[2] particles each at 0.5 / b + cell +
 * index / b, u = (ux0, 0, 0), q = +-density / (npic product) (negative for
 * cmr < 0), ids first_id + m * npt + particle index.  Returns particles kept. */
int64_t oq_fill_cells(oq_state *s, double cmr, int64_t ncells, const int32_t *cells, const double *density,
                      const int npic[3], double ux0, int32_t first_id);
void oq_diagnostics(oq_state *s, oq_diag *d);   /* fout.cpp:41-138 */
void oq_compute_rho(oq_state *s);         /* fout.cpp:7-39 (rho stored in J slots) */
/* one full step in the order of src/main.cpp:2319-2332; diag may be NULL */
void oq_step(oq_state *s, int nsteps, oq_diag *diag);

/* laser initial condition, f_init_cos with sscos = 0 (init.cpp:6-210) */
void oq_init_laser_cos(oq_state *s, double a0y, double a0z, double xsigma, double ysigma,
                       double zsigma, double x0, int b_sign, double x1, double phase,
                       double y0, double z0, int append, double phi,
                       double xtarget, double ytarget, double ztarget);

/* A laser entering through an x plane (the reference's lasers are initial
 * conditions only): the pulse of oq_init_laser_cos's parameters L[16]
 * (a0y a0z xsigma ysigma zsigma x0 b_sign x1 phase y0 z0 append phi xt yt
 * zt), carried along its direction at c from its position at the next
 * oq_maxwell call, is fed in as a total-field / scattered-field source:
 * side 0 holds the total field at i >= plane (2 <= plane <= nx - 6), side 1
 * at i < plane (6 <= plane <= nx - 2).  At most 4 sources; returns 0, or -1
 * on bad arguments. */
int oq_add_laser_source(oq_state *s, const double L[16], int side, int plane);
/* the same pulse after travelling `shift` (c/omega units) along its
 * direction, on the grid of s: staggered = 1 samples each component where it
 * lives on the staggered grid (E_a half a cell along a, B_a half a cell
 * across a), 0 at the nodes as f_init_cos does.  E, B: [nx][ny][nz][3]. */
void oq_laser_field(oq_state *s, const double L[16], double shift, int staggered, double *E, double *B);

/* uniform thermal electron plasma from the shared Philox loader (philox.h);
 * cells with i in [ilo, ihi) (and is_inside); returns particles added */
int64_t oq_load_uniform(oq_state *s, int ppc, double density_ncr, double sigma_u,
                        uint64_t seed, int ilo, int ihi);

/* the same loader writing to caller arrays: rec8 holds x,y,z,ux,uy,uz,q,cmr
 * per particle (row-major), id the global particle id.  Pass rec8 = NULL to
 * count only.  Returns the number of particles. */
int64_t oq_synth_plasma(int nx, int ny, int nz, int ppc, double density_ncr, double sigma_u,
                        uint64_t seed, int ilo, int ihi, double *rec8, int32_t *id);

/* Philox4x32-10 of philox.h on one counter block (ctr[4], key[2]) -> out[4],
 * exposed so the generator can be checked against the published
 * known-answer vectors */
void oq_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* single-particle kernels, exposed for closed-form tests */
void oq_push(int pusher, double cmr, double dt, const double e[3], const double b[3],
             double u[3], double *g);
double oq_chi(double e_s, const double e[3], const double b[3], const double u[3], double g);
double oq_rate_w(double e_s, double g, double *r, double chi);       /* mc.cpp:40-74 */
double oq_rate_tilde_w(double e_s, double g, double r, double chi);  /* mc.cpp:76-97 */
double oq_vacuum_rate(double e_s, const double e[3], const double b[3]); /* mc.cpp:99-116 */

#ifdef __cplusplus
}
#endif
#endif
