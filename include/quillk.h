/*
 * quillk.h -- C ABI of the MI355X-native per-step PIC loop (libquillk.so).
 *
 * One context holds one subdomain (the reference's spatial_region at
 * n_sr = 1, include/main.h:17-170): its fields, its particle species and the
 * solver / pusher / QED options.  All device memory is owned by the context;
 * inputs are copied in and outputs copied out, nothing is aliased across the
 * boundary.  Every entry point returns QK_OK (0) or an error code and never
 * aborts the process; qk_last_error() gives the text of the last failure on
 * the calling thread.
 *
 * Threading: a context is used by one host thread at a time (the caller
 * serialises calls on a context); distinct contexts are independent.
 *
 * Units and layouts follow the reference: lengths and times in c/omega,
 * fields in m c omega / e, momenta u = p / m c, positions in cell units from
 * the subdomain origin (src/main.cpp:2573-2580, src/advance.cpp:281-285).
 * Host field arrays are AoS double[nx][ny][nz][3], k fastest, the layout of
 * field3d<celle> (include/containers.h:17-52).
 */
#ifndef QUILLK_H
#define QUILLK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QK_OK 0
#define QK_ERR_INVALID 1    /* bad argument or configuration */
#define QK_ERR_CAPACITY 2   /* particle buffer too small */
#define QK_ERR_HIP 3        /* HIP runtime failure */
#define QK_ERR_STATE 4      /* call out of order */
#define QK_ERR_NOT_BUILT 5  /* feature not compiled in */

/* maxwell_solver_enum (include/containers.h:82-84) */
#define QK_SOLVER_FDTD 0
#define QK_SOLVER_NDFX 1
#define QK_SOLVER_FIVE_POINT 2
/* pusher_enum (include/containers.h:86-88) */
#define QK_PUSHER_VAY 0
#define QK_PUSHER_BORIS 1
/* arithmetic of particles and fields on the GPU */
#define QK_FP32 0
#define QK_FP64 1
/* current deposition strategy */
#define QK_DEPOSIT_TILED 0    /* per-workgroup LDS accumulation, flushed with atomics */
#define QK_DEPOSIT_GLOBAL 1   /* every contribution is a global atomic (plain method) */

/* species slots; cmr = -1, +1, 0, ion_cmr[n] (src/plist.cpp:31, init.cpp:413) */
#define QK_ELECTRON 0
#define QK_POSITRON 1
#define QK_PHOTON 2
#define QK_ION0 3
#define QK_MAX_ION 4
#define QK_MAX_SPECIES (QK_ION0 + QK_MAX_ION)

typedef struct {
    int32_t nx, ny, nz;           /* cells; nx <= 2048, ny, nz <= 1024 */
    double dx, dy, dz, dt;        /* c/omega units (config value x 2 pi) */
    int32_t solver;               /* QK_SOLVER_* */
    int32_t pusher;               /* QK_PUSHER_* */
    int32_t qed;                  /* photon emission, Breit-Wheeler, vacuum pairs */
    int32_t precision;            /* QK_FP32 or QK_FP64 */
    int32_t deposit;              /* QK_DEPOSIT_* */
    int32_t sort_every;           /* steps between cell-tile sorts; 0 = never */
    int32_t track_ids;            /* keep a per-particle id (tests); costs 4 B/particle in sorts */
    int32_t device;               /* HIP device ordinal */
    double e_s;                   /* Sauter field (lambda / lambda_C), src/main.cpp:2276 */
    double external_bz;           /* added to the gathered Bz, src/advance.cpp:136 */
    double q_vac;                 /* vacuum-pair macro charge, src/main.cpp:2324 */
    uint64_t seed;                /* QED random streams */
    int64_t capacity;             /* particle slots per species (0: 2 x first load) */
    int32_t n_ion;                /* ion populations */
    double ion_cmr[QK_MAX_ION];   /* their charge-to-mass ratios */
    double norm_n;                /* compute_N multiplier, src/main.cpp:2326 (0 -> 1) */
    double norm_ef;               /* compute_energy field multiplier, src/main.cpp:2327 (0 -> 1) */
    double norm_ep;               /* compute_energy particle multiplier (0 -> 1) */
    int32_t merging;              /* pmerging after a step that crosses the limit: 0 off, 1 "ti", 2 "nl" */
    double crnp;                  /* the limit: crpc x cells (src/main.cpp:1701), macro-particles per species */
    double sort_fallback;         /* also sort a species when more than this fraction of it was outside its
                                     tile's window two steps before (0: sort_every only) */
} qk_params;

/* per-step diagnostics, compute_N / compute_energy (src/fout.cpp:41-138) */
typedef struct {
    double N_e, N_p, N_ph;
    int64_t n_e, n_p, n_ph, n_ion;
    double energy_f, energy_e, energy_p, energy_ph, energy_ion;
    double energy_ions[QK_MAX_ION];   /* per ion population (ienergy, fout.cpp:120-137); energy_ion is their sum */
    int64_t n_nonfinite;              /* particles dropped this step because their momentum or displacement was
                                         not finite (corrupt input); the reference would carry the NaN on */
} qk_diag;

typedef struct qk_ctx qk_ctx;

/* spatial_region::init + create_arrays (src/low-level.cpp:61-150) */
int qk_create(const qk_params *params, qk_ctx **out);
/* ~spatial_region (src/low-level.cpp:152-184) */
int qk_destroy(qk_ctx *ctx);
/* text of the last error on this thread ("" if none) */
const char *qk_last_error(void);

/* write E and B: the public ce / cb members (include/main.h:68-69) */
int qk_set_fields(qk_ctx *ctx, const double *E, const double *B);
/* read E, B, J (current of the last step) and Be (cbe); any pointer may be NULL */
int qk_get_fields(qk_ctx *ctx, double *E, double *B, double *J, double *Be);

/* append particles of one species (new_particle + place, src/low-level.cpp:186,
 * src/plist.cpp:22-40): absolute positions in cells, u = p/mc, macro charge q.
 * Particles outside is_inside (src/plist.cpp:42-48) are dropped.  ids may be
 * NULL.  *kept (may be NULL) receives the number stored. */
int qk_add_particles(qk_ctx *ctx, int species, int64_t n, const double *x, const double *y,
                     const double *z, const double *ux, const double *uy, const double *uz,
                     const double *q, const int32_t *ids, int64_t *kept);
/* uniform thermal plasma generated on the device: ppc particles per is_inside
 * cell with ilo <= i < ihi, positions uniform in the cell and each momentum
 * component normal(0, sigma_u), charge -density/ppc (electron sign for
 * cmr < 0); counter-based Philox keyed by (seed, cell*ppc + slot) */
int qk_load_uniform_plasma(qk_ctx *ctx, int species, int ppc, double density_ncr, double sigma_u,
                           uint64_t seed, int ilo, int ihi, int64_t *added);
/* laser initial condition f_init_cos, cos envelope (src/init.cpp:6-210); L holds
 * a0y a0z xsigma ysigma zsigma x0 b_sign x1 phase y0 z0 append phi xt yt zt */
int qk_init_laser_cos(qk_ctx *ctx, const double L[16]);
/* f_init_boundaries + interpolate_be (src/main.cpp:2288-2289) */
int qk_init_boundaries(qk_ctx *ctx);

/* nsteps of the loop body of src/main.cpp:2319-2332: birth_from_vacuum,
 * padvance, compute_N, compute_energy, fadvance.  diag (may be NULL) receives
 * one record per step. */
int qk_step(qk_ctx *ctx, int nsteps, qk_diag *diag);
/* wait for all queued work of the context */
int qk_sync(qk_ctx *ctx);

/* live particles of one species */
int qk_particle_count(qk_ctx *ctx, int species, int64_t *n);
/* copy out live particles: rec10 receives n x (x y z ux uy uz g q cmr chi);
 * ids may be NULL.  cap is the row capacity of rec10. */
int qk_get_particles(qk_ctx *ctx, int species, int64_t cap, double *rec10, int32_t *ids, int64_t *n_out);
/* compute_rho (src/fout.cpp:7-39): node charge of electrons, positrons and
 * photons in the three slots of rho (AoS [nx][ny][nz][3]) */
int qk_compute_rho(qk_ctx *ctx, double *rho);
/* qk_compute_rho without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The node charges
 * accumulate in a fixed device array of the library (48 MiB on each GPU it
 * runs on, held until the process ends), so grids of more than 2097152 cells
 * are refused with QK_ERR_CAPACITY.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (contexts used from several threads) race. */
int qk_compute_rho_noalloc(qk_ctx *ctx, double *rho);
/* A laser entering through an x plane.  No counterpart in the reference,
 * whose lasers are initial conditions only (f_init_cos, src/init.cpp:6-210,
 * called from src/main.cpp:1150-1227): here the same pulse (L as in
 * qk_init_laser_cos, the cos envelope), carried along its direction at c
 * from where it stands at the next step, is fed in as a total-field /
 * scattered-field source.  side 0: the total field is held at planes
 * i >= plane (2 <= plane <= nx - 6), the pulse enters through x-min; side 1:
 * at i < plane (6 <= plane <= nx - 2), through x-max.  At most 4 sources per
 * side, sharing one plane.  Each update that reads across the plane is
 * corrected by the solver's own update applied to the incident field
 * (sampled on the staggered grid) on the other side. */
int qk_add_laser_source(qk_ctx *ctx, const double L[16], int side, int plane);


/* spatial_region::pmerging (src/mc.cpp:202-462): thin every species once;
 * mode 1 "ti" deletes each particle with probability ppd[0], mode 2 "nl" a
 * particle of species s with probability ppd[s] / (1 + (2|q|/<|q|>_s)^4),
 * and both rescale each species' charges to keep its total.  ppd has
 * QK_ION0 + n_ion entries (electrons, positrons, photons, ions).  qk_step
 * calls it by itself when params.merging is set (the trigger of
 * evaluate_merging_condition, src/main.cpp:1691-1748). */
int qk_pmerging(qk_ctx *ctx, int mode, const double *ppd);

/* spatial_region::moving_window (src/advance.cpp:389-454): erase the
 * particles of the plane i = 0, move E, B, Be and every particle one cell
 * toward -x (x -= 1), zero the plane nx - 1.  J is left as it is.  The caller
 * decides when (check_moving_window, src/main.cpp:1953-1978) and refills the
 * front (add_moving_window_particles, :1750-1835), as main() does. */
int qk_moving_window(qk_ctx *ctx);
/* spatial_region::get_max_w (src/low-level.cpp:512-532): the largest
 * E^2 + B^2 over the planes ilo <= i < ihi (the "auto" window trigger) */
int qk_max_w(qk_ctx *ctx, int ilo, int ihi, double *max_w);
/* fill_cell_by_particles (src/init.cpp:402-478) with T = 0 on a list of cells:
 * cells holds ncells (i, j, k) triples, density their n; each cell gets
 * npic[0] x npic[1] x npic[2] particles at 0.5/b + cell + index/b per axis,
 * u = (ux0, 0, 0), charge density / (npic product) (negative for cmr < 0).
 * Cells outside is_inside are skipped (place() deletes their particles).
 * With track_ids, the particles of the m-th listed cell get the ids
 * first_id + m x npt + (ip npic[1] + jp) npic[2] + kp.  *added (may be NULL)
 * receives the number stored. */
int qk_fill_cells(qk_ctx *ctx, int species, int64_t ncells, const int32_t *cells, const double *density,
                  const int32_t npic[3], double ux0, int32_t first_id, int64_t *added);

/* Adjoint of the deterministic step (QED off).  qk_tape_begin records the
 * trajectory that the following qk_step calls make: a snapshot of the whole
 * state now and every `every` steps (memory: one state per snapshot plus
 * `every` states while the adjoint runs).  It needs qed = 0, merging = 0,
 * sort_every = 0, sort_fallback = 0 (particles keep their slots) and no
 * photons.  With QK_DEPOSIT_TILED it sorts the particles into tiles once,
 * first, which permutes the slots (read the new order with qk_get_particles
 * and ids).  No counterpart in the reference: it differentiates the loop
 * body of src/main.cpp:2319-2332 (padvance and fadvance). */
int qk_tape_begin(qk_ctx *ctx, int every);
/* Tangent-linear run along the recorded steps (no counterpart in the
 * reference): right after qk_tape_begin, seed a perturbation of the state
 * there -- dE, dB (AoS [nx][ny][nz][3], perturbing the fields as given to
 * qk_set_fields before qk_init_boundaries) and per species dpart[s] (n x
 * (x y z ux uy uz q), slots of the tape); any pointer may be NULL (zero).
 * Each following qk_step carries it forward in FP64 along the same branches
 * (face crossings) as the forward step; qk_tangent_get reads the tangent of
 * the current state.  With qk_adjoint on the same tape,
 * <lam, tangent> = <gradient, seed> for any lam and seed: the dot-product
 * check of the adjoint. */
int qk_tangent_begin(qk_ctx *ctx, const double *dE, const double *dB, const double *const *dpart);
int qk_tangent_get(qk_ctx *ctx, double *dE, double *dB, double *const *dpart);

/* Given the adjoints of the state after the recorded steps -- lam_E, lam_B
 * (AoS [nx][ny][nz][3], may be NULL for zero) and per species slot
 * (QK_MAX_SPECIES pointers, each NULL or n x 7 doubles: x y z ux uy uz q of
 * the n slots the species had after qk_tape_begin, in slot order -- the
 * order they were added with QK_DEPOSIT_GLOBAL; zero for particles deleted
 * on the way) -- return the adjoints of
 * the state at qk_tape_begin in grad_E, grad_B (before init_boundaries: zero
 * on the planes it clears) and grad_part (same layout).  Positions are in
 * cells.  The state after the call is the state before it. */
int qk_adjoint(qk_ctx *ctx, const double *lam_E, const double *lam_B, const double *const *lam_part,
               double *grad_E, double *grad_B, double *const *grad_part);

/* timing of the last qk_step call, measured with HIP events on the context's
 * stream: total milliseconds and the part spent in the particle kernels */
int qk_last_timing(qk_ctx *ctx, double *ms_total, double *ms_particles, double *ms_fields);
/* fraction of the particle pushes of the last qk_step call whose current went
 * through global atomics because the particle had left its tile's LDS window
 * since the last sort (0 for QK_DEPOSIT_GLOBAL, where every push does) */
int qk_last_fallback(qk_ctx *ctx, double *fraction);
/* the context's HIP stream (hipStream_t) */
int qk_get_stream(qk_ctx *ctx, void **stream);

#ifdef __cplusplus
}
#endif
#endif
