/*
 * oracle/golden_format.h -- binary layout of the case files exchanged with
 * the compiled reference harness (oracle/ref_harness.cpp).
 *
 * TEST INFRASTRUCTURE.  tests/golden/golden_io.py mirrors this layout with a
 * numpy structured dtype; keep the two in step.
 *
 * Input file:  header | E[ncell*3] | B[ncell*3] | particles n x 8 doubles
 *              (x, y, z, ux, uy, uz, q, cmr) | ids n x int32
 *              | (mode 3 only) push table
 * Output file: header | E | B | J | Be | int64 n_out | n_out x 10 doubles
 *              (x, y, z, ux, uy, uz, g, q, cmr, chi) | n_out x int32 ids
 *              | nsteps x 12 doubles of diagnostics
 * All fields are AoS double[nx][ny][nz][3], k fastest (include/containers.h:17-52).
 */
#ifndef QUILL_GOLDEN_FORMAT_H
#define QUILL_GOLDEN_FORMAT_H

#include <stdint.h>

#define QG_MAGIC 0x4C4C5551   /* "QULL" little endian */

enum {
    QG_MODE_RUN = 0,      /* init_boundaries + interpolate_be, then nsteps full steps */
    QG_MODE_LASER = 1,    /* f_init_cos only, dump E and B */
    QG_MODE_RATES = 2,    /* chi / w / tilde_w / mathcal_W on the rows of a table */
    QG_MODE_PUSH = 3,     /* pushers on the rows of a table */
    QG_MODE_TIME = 4,     /* timing: synthetic plasma + laser, warm-up + timed steps */
    QG_MODE_FOUT = 5,     /* output writers: fields + particles in, the reference's plane files out */
    QG_MODE_MERGE = 6,    /* pmerging once: particles in (ppd in laser[], ti/nl in extra[0]), particles out */
    QG_MODE_MW = 7,       /* RUN with a moving window; extra: mode (1 on, 2 auto), mwspeed, mwtolerance,
                             column density, npic x y z, t_add_mw; per step nmw and the auto ratio follow the diagnostics */
    QG_MODE_FILM = 8      /* one film() call: arguments in laser[], append flag and gradwidth_y in extra[0..1] */
};

typedef struct {
    int32_t magic;
    int32_t mode;
    int32_t nx, ny, nz;
    int32_t solver, pusher, qed;
    int32_t nsteps;
    int32_t n_particles;     /* RUN: particles in the input; RATES/PUSH: table rows */
    int32_t seed;
    int32_t n_ion;           /* ion populations (cmr values in ion_cmr) */
    double dx, dy, dz, dt;
    double e_s, external_bz, q_vac;
    double norm_n, norm_ef, norm_ep;
    double laser[16];        /* a0y a0z xsigma ysigma zsigma x0 b_sign x1 phase y0 z0 append phi xt yt zt */
    double ion_cmr[4];
    double extra[8];         /* TIME: ppc, density, sigma_u, warmup, ilo, ihi; FOUT: i_yz, binary */
} qg_header;

/* diagnostics row: N_e N_p N_ph n_e n_p n_ph energy_f energy_e energy_p energy_ph energy_ion n_ion */
#define QG_DIAG_WIDTH 12

#endif
