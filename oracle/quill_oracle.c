/*
 * oracle/quill_oracle.c -- FP64 CPU restatement of QUILL's per-step PIC loop
 * (one subdomain, n_sr = 1).
 *
 * TEST INFRASTRUCTURE ONLY.  The GPU product never links, loads or calls this
 * code; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it
 * as the checker.
 *
 * Each function cites the reference file:line whose algorithm it restates.
 * Pinning (what makes this restatement trustworthy), by function:
 *   push (Boris, Vay)        golden vectors from the compiled reference
 *                            (tests/golden) + closed-form orbits
 *   gather / interpolate_be  golden vectors + exact reproduction of linear fields
 *   jdeposition              golden vectors + charge conservation (div E - rho
 *                            invariant) + brute-force path integral
 *   solvers + boundaries     golden vectors (all three solvers) + vacuum energy
 *   laser IC (f_init_cos)    golden vectors
 *   diagnostics              golden vectors (energies, counts)
 *   QED rates                closed forms restated in tests + golden vectors of
 *                            the rate functions from the compiled reference
 *   QED event sequences      statistical only (the reference's generator is
 *                            seeded from the wall clock, mc.cpp:12-23)
 *   moving window, max_w,    golden runs of the compiled reference with the
 *   fill_cells               window on (fixed and automatic shifts)
 *   pmerging                 statistics of 16 merges by the compiled reference
 *   compute_rho              the Gauss-law invariant div E - rho + golden planes
 *   uniform plasma loader    no reference counterpart (the builder's Philox
 *                            loader); the GPU loader must match it exactly
 */
#include "quill_oracle.h"
#include "philox.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OQ_PI 3.141592653589793   /* include/physical_constants.h:3 */

typedef struct {
    double x, y, z, ux, uy, uz, g, q, cmr, chi;
    int32_t id;
    int32_t alive;
} oq_particle;

/* the pulse of f_init_cos (sscos = 0): its frame and its fields at a point */
typedef struct {
    double a0y, a0z, xs, ys, zs;          /* envelope half lengths (init.cpp:34-37) */
    double x0, y0, z0, xt, yt, zt, phase;
    double r0[3], ey[3], ez[3];           /* propagation frame (init.cpp:55-93) */
    double y12, z12;
    int b_sign;
} oq_pulse;

#define OQ_MAX_SOURCES 4

struct oq_state {
    oq_params p;
    int64_t ncell;
    double *f[4];              /* E, B, J, Be : [nx][ny][nz][3] */
    oq_particle *part;
    int64_t np, cap;
    uint64_t rng_ctr;          /* sequential Philox stream for QED draws */
    /* laser sources at an x plane (oq_add_laser_source) */
    int nsrc;
    oq_pulse src[OQ_MAX_SOURCES];
    int src_side[OQ_MAX_SOURCES], src_plane[OQ_MAX_SOURCES];
    int64_t src_t0[OQ_MAX_SOURCES];
    int src_shift0[OQ_MAX_SOURCES];
    int64_t field_steps;       /* oq_maxwell calls so far */
    int window_shifts;         /* oq_moving_window calls so far */
    oq_state *tmp;             /* same grid, the sources' correction passes */
};

/* ------------------------------------------------------------------ */
/* helpers                                                             */

static inline int64_t IDX(const oq_state *s, int i, int j, int k)
{
    return (((int64_t)i * s->p.ny + j) * s->p.nz + k) * 3;
}
#define FE(i, j, k, c) (s->f[OQ_E][IDX(s, i, j, k) + (c)])
#define FB(i, j, k, c) (s->f[OQ_B][IDX(s, i, j, k) + (c)])
#define FJ(i, j, k, c) (s->f[OQ_J][IDX(s, i, j, k) + (c)])
#define FBE(i, j, k, c) (s->f[OQ_BE][IDX(s, i, j, k) + (c)])

static double oq_rand(oq_state *s)
{
    /* uniform in [0,1), 53 bits, from a sequential Philox stream */
    uint64_t c = s->rng_ctr++;
    oq_u32x4 ctr = {{(uint32_t)c, (uint32_t)(c >> 32), 0x0ED5u, 0x0ACEu}};
    oq_u32x4 r = oq_philox4x32_10(ctr, (uint32_t)s->p.seed, (uint32_t)(s->p.seed >> 32));
    uint64_t m = ((uint64_t)r.v[0] << 21) ^ (uint64_t)(r.v[1] >> 11);
    return (double)(m & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
}

/* is_inside (plist.cpp:42-48): 0 < i < n-2 along each axis */
static inline int is_inside(const oq_state *s, int i, int j, int k)
{
    return i > 0 && i < s->p.nx - 2 && j > 0 && j < s->p.ny - 2 && k > 0 && k < s->p.nz - 2;
}

static oq_particle *new_particle(oq_state *s)
{
    if (s->np == s->cap) {
        s->cap = s->cap ? s->cap * 2 : 1024;
        s->part = (oq_particle *)realloc(s->part, (size_t)s->cap * sizeof(oq_particle));
    }
    oq_particle *p = &s->part[s->np++];
    memset(p, 0, sizeof(*p));
    p->g = 1;
    p->cmr = -1;   /* particle() defaults, containers.cpp:39-56 */
    p->alive = 1;
    return p;
}

/* ------------------------------------------------------------------ */
/* lifecycle                                                           */

oq_state *oq_create(const oq_params *p)
{
    oq_state *s = (oq_state *)calloc(1, sizeof(oq_state));
    s->p = *p;
    if (s->p.norm_n == 0) s->p.norm_n = 1;
    if (s->p.norm_ef == 0) s->p.norm_ef = 1;
    if (s->p.norm_ep == 0) s->p.norm_ep = 1;
    s->ncell = (int64_t)p->nx * p->ny * p->nz;
    for (int a = 0; a < 4; ++a) s->f[a] = (double *)calloc((size_t)s->ncell * 3, sizeof(double));
    return s;
}

void oq_destroy(oq_state *s)
{
    if (!s) return;
    oq_destroy(s->tmp);
    for (int a = 0; a < 4; ++a) free(s->f[a]);
    free(s->part);
    free(s);
}

double *oq_field(oq_state *s, int which) { return s->f[which]; }
int64_t oq_count(const oq_state *s) { return s->np; }

int64_t oq_add_particles(oq_state *s, int64_t n, const double *x, const double *y, const double *z,
                         const double *ux, const double *uy, const double *uz, const double *q,
                         double cmr, const int32_t *id)
{
    int64_t kept = 0;
    for (int64_t m = 0; m < n; ++m) {
        if (!is_inside(s, (int)x[m], (int)y[m], (int)z[m])) continue;   /* place(), plist.cpp:22-31 */
        oq_particle *p = new_particle(s);
        p->x = x[m]; p->y = y[m]; p->z = z[m];
        p->ux = ux[m]; p->uy = uy[m]; p->uz = uz[m];
        p->q = q[m];
        p->cmr = cmr;
        double u2 = p->ux * p->ux + p->uy * p->uy + p->uz * p->uz;
        p->g = (cmr != 0) ? sqrt(1 + u2) : sqrt(u2);                    /* init.cpp:461-464 */
        p->chi = 0;
        p->id = id ? id[m] : (int32_t)m;
        ++kept;
    }
    return kept;
}

void oq_get_particles(const oq_state *s, double *rec, int32_t *id)
{
    int64_t m = 0;
    for (int64_t n = 0; n < s->np; ++n) {
        const oq_particle *p = &s->part[n];
        if (!p->alive) continue;
        double *r = rec + 10 * m;
        r[0] = p->x; r[1] = p->y; r[2] = p->z;
        r[3] = p->ux; r[4] = p->uy; r[5] = p->uz;
        r[6] = p->g; r[7] = p->q; r[8] = p->cmr; r[9] = p->chi;
        if (id) id[m] = p->id;
        ++m;
    }
}

static void compact(oq_state *s)
{
    int64_t m = 0;
    for (int64_t n = 0; n < s->np; ++n)
        if (s->part[n].alive) s->part[m++] = s->part[n];
    s->np = m;
}

/* moving_window (advance.cpp:389-454): the particles of the plane i = 0
 * are erased, E, B and Be move one plane toward -x (F[i] = F[i + 1]), the
 * plane nx - 1 is zeroed and every particle's x drops by one.  J stays. */
void oq_moving_window(oq_state *s)
{
    const int nx = s->p.nx;
    const int64_t plane = (int64_t)s->p.ny * s->p.nz * 3;
    for (int64_t n = 0; n < s->np; ++n)
        if (s->part[n].alive && (int)s->part[n].x == 0) s->part[n].alive = 0;
    compact(s);
    const int moved[3] = {OQ_E, OQ_B, OQ_BE};
    for (int a = 0; a < 3; ++a) {
        double *F = s->f[moved[a]];
        memmove(F, F + plane, (size_t)((int64_t)(nx - 1) * plane) * sizeof(double));
        memset(F + (int64_t)(nx - 1) * plane, 0, (size_t)plane * sizeof(double));
    }
    for (int64_t n = 0; n < s->np; ++n) s->part[n].x -= 1;
    ++s->window_shifts;
}

/* get_max_w (low-level.cpp:516-532) */
double oq_max_w(const oq_state *s, int ilo, int ihi)
{
    double m = 0;
    for (int i = ilo; i < ihi; ++i)
        for (int j = 0; j < s->p.ny; ++j)
            for (int k = 0; k < s->p.nz; ++k) {
                const double *e = &s->f[OQ_E][IDX(s, i, j, k)], *b = &s->f[OQ_B][IDX(s, i, j, k)];
                double w = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
                if (w > m) m = w;
            }
    return m;
}

/* fill_cell_by_particles (init.cpp:402-478) with T = 0 and dsplmt = 0 on
 * listed cells; cell m's particles get ids first_id + m * npt + their index
 * in (ip, jp, kp) order; place() drops the cells outside is_inside */
int64_t oq_fill_cells(oq_state *s, double cmr, int64_t ncells, const int32_t *cells, const double *density,
                      const int npic[3], double ux0, int32_t first_id)
{
    const int bi = npic[0], bj = npic[1], bk = npic[2], npt = bi * bj * bk;
    int64_t kept = 0;
    for (int64_t m = 0; m < ncells; ++m) {
        const int ci = cells[3 * m], cj = cells[3 * m + 1], ck = cells[3 * m + 2];
        if (!is_inside(s, ci, cj, ck)) continue;
        double q0 = 1 / (double)npt * density[m];
        if (cmr < 0) q0 = -q0;
        for (int ip = 0; ip < bi; ++ip)
            for (int jp = 0; jp < bj; ++jp)
                for (int kp = 0; kp < bk; ++kp) {
                    oq_particle *p = new_particle(s);
                    p->cmr = cmr;
                    p->q = q0;
                    p->x = (0.5 / bi + (double)ci) + (double)ip / (double)bi;
                    p->y = (0.5 / bj + (double)cj) + (double)jp / (double)bj;
                    p->z = (0.5 / bk + (double)ck) + (double)kp / (double)bk;
                    p->ux = ux0;
                    p->uy = 0;
                    p->uz = 0;
                    p->g = cmr != 0 ? sqrt(1 + ux0 * ux0) : sqrt(ux0 * ux0);
                    p->chi = 0;
                    p->id = first_id + (int32_t)(m * npt) + (ip * bj + jp) * bk + kp;
                    ++kept;
                }
    }
    return kept;
}

/* ------------------------------------------------------------------ */
/* pushers (pusher.cpp)                                               */

/* Vay 2008 (pusher.cpp:5-42): uses the stored gamma of the previous half step */
static void push_vay(double cmr, double dt, const double ef[3], const double bf[3],
                     double u[3], double *gp)
{
    double h = dt * cmr;
    double e[3] = {ef[0] * h, ef[1] * h, ef[2] * h};
    double hb = 0.5 * h;
    double b[3] = {bf[0] * hb, bf[1] * hb, bf[2] * hb};
    double g = *gp;
    double w[3] = {u[0] + e[0] + (u[1] * b[2] - u[2] * b[1]) / g,
                   u[1] + e[1] + (u[2] * b[0] - u[0] * b[2]) / g,
                   u[2] + e[2] + (u[0] * b[1] - u[1] * b[0]) / g};
    double b2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    double bw = b[0] * w[0] + b[1] * w[1] + b[2] * w[2];
    double s = 1 + w[0] * w[0] + w[1] * w[1] + w[2] * w[2] - b2;
    g = sqrt(0.5 * (s + sqrt(s * s + 4 * (b2 + bw * bw))));
    double den = 1 / (1 + b2 / (g * g));
    double bwg = bw / (g * g);
    u[0] = (w[0] + (w[1] * b[2] - w[2] * b[1]) / g + b[0] * bwg) * den;
    u[1] = (w[1] + (w[2] * b[0] - w[0] * b[2]) / g + b[1] * bwg) * den;
    u[2] = (w[2] + (w[0] * b[1] - w[1] * b[0]) / g + b[2] * bwg) * den;
    *gp = g;
}

/* Boris (pusher.cpp:44-97) */
static void push_boris(double cmr, double dt, const double ef[3], const double bf[3],
                       double u[3], double *gp)
{
    double h = 0.5 * dt * cmr;
    double e[3] = {ef[0] * h, ef[1] * h, ef[2] * h};
    double um[3] = {u[0] + e[0], u[1] + e[1], u[2] + e[2]};
    double g = sqrt(1 + um[0] * um[0] + um[1] * um[1] + um[2] * um[2]);
    double hg = h / g;
    double t[3] = {bf[0] * hg, bf[1] * hg, bf[2] * hg};
    double w[3] = {um[0] + (um[1] * t[2] - um[2] * t[1]),
                   um[1] + (um[2] * t[0] - um[0] * t[2]),
                   um[2] + (um[0] * t[1] - um[1] * t[0])};
    double f = 2 / (1 + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    um[0] += (w[1] * t[2] - w[2] * t[1]) * f;
    um[1] += (w[2] * t[0] - w[0] * t[2]) * f;
    um[2] += (w[0] * t[1] - w[1] * t[0]) * f;
    u[0] = um[0] + e[0];
    u[1] = um[1] + e[1];
    u[2] = um[2] + e[2];
    *gp = sqrt(1 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
}

void oq_push(int pusher, double cmr, double dt, const double e[3], const double b[3],
             double u[3], double *g)
{
    if (pusher == OQ_BORIS) push_boris(cmr, dt, e, b, u, g);
    else push_vay(cmr, dt, e, b, u, g);
}

static void push_particle(const oq_state *s, oq_particle *p, const double e[3], const double b[3], double dt)
{
    double u[3] = {p->ux, p->uy, p->uz};
    oq_push(s->p.pusher, p->cmr, dt, e, b, u, &p->g);
    p->ux = u[0]; p->uy = u[1]; p->uz = u[2];
}

/* ------------------------------------------------------------------ */
/* gather (deposition.cpp:431-471)                                     */
/* Each component is bilinear in the two directions transverse to it   */
/* and taken at the lower node along its own direction.                */

static void gather(const oq_state *s, const double *F, double x, double y, double z, double out[3])
{
    int i = (int)floor(x), j = (int)floor(y), k = (int)floor(z);
    double xa = x - i, xb = 1 - xa, ya = y - j, yb = 1 - ya, za = z - k, zb = 1 - za;
    const double *c000 = F + IDX(s, i, j, k);
    const double *c100 = F + IDX(s, i + 1, j, k);
    const double *c010 = F + IDX(s, i, j + 1, k);
    const double *c001 = F + IDX(s, i, j, k + 1);
    const double *c110 = F + IDX(s, i + 1, j + 1, k);
    const double *c101 = F + IDX(s, i + 1, j, k + 1);
    const double *c011 = F + IDX(s, i, j + 1, k + 1);
    out[0] = c000[0] * yb * zb + c010[0] * ya * zb + c001[0] * yb * za + c011[0] * ya * za;
    out[1] = c000[1] * xb * zb + c100[1] * xa * zb + c001[1] * xb * za + c101[1] * xa * za;
    out[2] = c000[2] * xb * yb + c100[2] * xa * yb + c010[2] * xb * ya + c110[2] * xa * ya;
}

/* ------------------------------------------------------------------ */
/* current deposition (deposition.cpp:13-356)                          */

/* one straight segment d inside cell w, from position p (simple_jdep,
 * deposition.cpp:13-39): linear-shape charge-conserving weights */
static void deposit_segment(oq_state *s, double q, const double p[3], const double d[3], const int w[3])
{
    double xa = p[0] - w[0], xb = 1 - xa;
    double ya = p[1] - w[1], yb = 1 - ya;
    double za = p[2] - w[2], zb = 1 - za;
    int i = w[0], j = w[1], k = w[2];
    double qx = d[0] * q, qy = d[1] * q, qz = d[2] * q;
    /* jx on the four x-edges i, j..j+1, k..k+1 */
    FJ(i, j + 1, k + 1, 0) += qx * (ya * za + 0.5 * (d[2] * ya + d[1] * za) + d[1] * d[2] / 3);
    FJ(i, j, k + 1, 0)     += qx * (yb * za + 0.5 * (d[2] * yb - d[1] * za) - d[1] * d[2] / 3);
    FJ(i, j + 1, k, 0)     += qx * (ya * zb + 0.5 * (d[1] * zb - d[2] * ya) - d[1] * d[2] / 3);
    FJ(i, j, k, 0)         += qx * (yb * zb - 0.5 * (d[2] * yb + d[1] * zb) + d[1] * d[2] / 3);
    /* jy on the y-edges */
    FJ(i + 1, j, k + 1, 1) += qy * (xa * za + 0.5 * (d[0] * za + d[2] * xa) + d[0] * d[2] / 3);
    FJ(i, j, k + 1, 1)     += qy * (xb * za + 0.5 * (d[2] * xb - d[0] * za) - d[0] * d[2] / 3);
    FJ(i + 1, j, k, 1)     += qy * (xa * zb + 0.5 * (d[0] * zb - d[2] * xa) - d[0] * d[2] / 3);
    FJ(i, j, k, 1)         += qy * (xb * zb - 0.5 * (d[0] * zb + d[2] * xb) + d[0] * d[2] / 3);
    /* jz on the z-edges */
    FJ(i + 1, j + 1, k, 2) += qz * (xa * ya + 0.5 * (d[0] * ya + d[1] * xa) + d[0] * d[1] / 3);
    FJ(i, j + 1, k, 2)     += qz * (xb * ya + 0.5 * (d[1] * xb - d[0] * ya) - d[0] * d[1] / 3);
    FJ(i + 1, j, k, 2)     += qz * (xa * yb + 0.5 * (d[0] * yb - d[1] * xa) - d[0] * d[1] / 3);
    FJ(i, j, k, 2)         += qz * (xb * yb - 0.5 * (d[0] * yb + d[1] * xb) + d[0] * d[1] / 3);
}

/* regulate (low-level.cpp:316-366): the three values in descending order */
static void sort3_desc(double a, double b, double c, double v[3])
{
    if (a > b) {
        if (b > c) { v[0] = a; v[1] = b; v[2] = c; }
        else if (a > c) { v[0] = a; v[1] = c; v[2] = b; }
        else { v[0] = c; v[1] = a; v[2] = b; }
    } else {
        if (c > b) { v[0] = c; v[1] = b; v[2] = a; }
        else if (a > c) { v[0] = b; v[1] = a; v[2] = c; }
        else { v[0] = b; v[1] = c; v[2] = a; }
    }
}

/*
 * jdeposition (deposition.cpp:41-356).  The straight path d from p is split
 * at every crossed cell face (at most one per axis).  With one or two
 * crossings the segments run through the start cell, then the cell with the
 * first-crossed axis index switched, then the end cell; ties between two
 * crossing fractions go to the axis the reference tests first (y before z,
 * x before z, y before x).  With three crossings the reference sorts the
 * fractions (regulate) and takes each segment's cell as the floor of its
 * midpoint.  The particle position is advanced segment by segment.
 */
static void jdeposition(oq_state *s, oq_particle *p, const double d[3])
{
    double pos[3] = {p->x, p->y, p->z};
    int c1[3] = {(int)floor(pos[0]), (int)floor(pos[1]), (int)floor(pos[2])};
    int c2[3] = {(int)floor(pos[0] + d[0]), (int)floor(pos[1] + d[1]), (int)floor(pos[2] + d[2])};
    int cross[3] = {c1[0] != c2[0], c1[1] != c2[1], c1[2] != c2[2]};
    int ncross = cross[0] + cross[1] + cross[2];
    double seg[3];
    int w[3] = {c1[0], c1[1], c1[2]};

#define ADVANCE(sg) do { pos[0] += (sg)[0]; pos[1] += (sg)[1]; pos[2] += (sg)[2]; } while (0)

    if (ncross == 0) {
        deposit_segment(s, p->q, pos, d, w);
        ADVANCE(d);
    } else if (ncross == 1) {
        int a = cross[0] ? 0 : (cross[1] ? 1 : 2);
        double face = (double)(c1[a] > c2[a] ? c1[a] : c2[a]);
        double t = face - pos[a];
        for (int m = 0; m < 3; ++m) seg[m] = (m == a) ? t : d[m] * t / d[a];
        deposit_segment(s, p->q, pos, seg, w);
        ADVANCE(seg);
        for (int m = 0; m < 3; ++m) seg[m] = d[m] - seg[m];
        w[a] = c2[a];
        deposit_segment(s, p->q, pos, seg, w);
        ADVANCE(seg);
    } else if (ncross == 2) {
        /* axes (first, second) in the reference's tie order */
        int a0, a1;
        if (!cross[0]) { a0 = 1; a1 = 2; }        /* y & z: y wins ties */
        else if (!cross[1]) { a0 = 0; a1 = 2; }   /* x & z: x wins ties */
        else { a0 = 1; a1 = 0; }                  /* x & y: y wins ties */
        double t0 = ((double)(c1[a0] > c2[a0] ? c1[a0] : c2[a0]) - pos[a0]) / d[a0];
        double t1 = ((double)(c1[a1] > c2[a1] ? c1[a1] : c2[a1]) - pos[a1]) / d[a1];
        int first = a0, second = a1;
        double tf = t0, ts = t1;
        if (t0 > t1) { first = a1; second = a0; tf = t1; ts = t0; }
        for (int m = 0; m < 3; ++m) seg[m] = d[m] * tf;
        deposit_segment(s, p->q, pos, seg, w);
        ADVANCE(seg);
        for (int m = 0; m < 3; ++m) seg[m] = d[m] * (ts - tf);
        w[first] = c2[first];
        deposit_segment(s, p->q, pos, seg, w);
        ADVANCE(seg);
        for (int m = 0; m < 3; ++m) seg[m] = d[m] * (1 - ts);
        w[second] = c2[second];
        deposit_segment(s, p->q, pos, seg, w);
        ADVANCE(seg);
    } else {
        double t[3];
        for (int m = 0; m < 3; ++m)
            t[m] = ((double)(c1[m] > c2[m] ? c1[m] : c2[m]) - pos[m]) / d[m];
        double v[3];
        sort3_desc(t[0], t[1], t[2], v);
        double frac[4] = {v[2], v[1] - v[2], v[0] - v[1], 1 - v[0]};
        for (int n = 0; n < 4; ++n) {
            for (int m = 0; m < 3; ++m) seg[m] = d[m] * frac[n];
            if (n > 0)
                for (int m = 0; m < 3; ++m) w[m] = (int)floor(pos[m] + 0.5 * seg[m]);
            deposit_segment(s, p->q, pos, seg, w);
            ADVANCE(seg);
        }
    }
#undef ADVANCE
    p->x = pos[0]; p->y = pos[1]; p->z = pos[2];
}

/* ------------------------------------------------------------------ */
/* fields: B interpolation to E positions (advance.cpp:20-83)          */
/* Each component of Be is the mean of B over a 2x2x2 stencil; on a    */
/* boundary plane the half of the stencil that leaves the grid is      */
/* dropped (4-point mean).  Cells on two or more boundary planes keep  */
/* their value.                                                        */

void oq_interpolate_be(oq_state *s)
{
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    /* stencil offsets per component: {x offset, y offset, z offset}, each 0 or +-1 */
    static const int off[3][3] = {{+1, -1, -1}, {-1, +1, -1}, {-1, -1, +1}};
    const int n[3] = {nx, ny, nz};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int c[3] = {i, j, k};
                int nb = (i == 0 || i == nx - 1) + (j == 0 || j == ny - 1) + (k == 0 || k == nz - 1);
                if (nb >= 2) continue;
                for (int comp = 0; comp < 3; ++comp) {
                    int lo[3], hi[3], cnt = 1;
                    for (int a = 0; a < 3; ++a) {
                        int o = c[a] + off[comp][a];
                        lo[a] = c[a] < o ? c[a] : o;
                        hi[a] = c[a] < o ? o : c[a];
                        if (o < 0 || o > n[a] - 1) { lo[a] = hi[a] = c[a]; }
                        else cnt *= 2;
                    }
                    (void)cnt;
                    double sum = 0;
                    int count = 0;
                    for (int a = lo[0]; a <= hi[0]; ++a)
                        for (int b = lo[1]; b <= hi[1]; ++b)
                            for (int d = lo[2]; d <= hi[2]; ++d) { sum += FB(a, b, d, comp); ++count; }
                    FBE(i, j, k, comp) = (count == 8 ? 0.125 : 0.25) * sum;
                }
            }
}

/* ------------------------------------------------------------------ */
/* Maxwell solvers (maxwell.cpp)                                       */
/*                                                                     */
/* The reference writes each solver as an interior sweep plus plane    */
/* sweeps.  Their union is restated here as per-component index sets;  */
/* updates of B read only E (and of E only B, J), so sweep order does  */
/* not matter.                                                         */

static inline int in(int v, int lo, int hi) { return v >= lo && v <= hi; }

/* zero-padded read used by the five-point planes at n-1 (maxwell.cpp:52-84) */
static inline double rdz(const oq_state *s, const double *F, int i, int j, int k, int c)
{
    if (i < 0 || j < 0 || k < 0 || i >= s->p.nx || j >= s->p.ny || k >= s->p.nz) return 0;
    return F[IDX(s, i, j, k) + c];
}

typedef struct { double ax, ay, az, bx, by, bz; } coef6;

static coef6 solver_coefs(const oq_state *s)
{
    coef6 c = {0, 0, 0, 1, 1, 1};
    const double dx = s->p.dx, dy = s->p.dy, dz = s->p.dz, dt = s->p.dt;
    if (s->p.solver == OQ_NDFX) {
        /* maxwell.cpp:238-249 */
        double kappa = 1.25 * (sqrt(3.0) - 1) * 0.5 / sqrt(3.0);
        c.bx = 1 - kappa;
        c.ax = 0.5 * (1 - c.bx);
        c.by = 1 - kappa * dx * dx / (dy * dy);
        c.bz = 1 - kappa * dx * dx / (dz * dz);
        c.ay = 0.5 * (1 - c.by);
        c.az = 0.5 * (1 - c.bz);
    } else if (s->p.solver == OQ_FIVE_POINT) {
        /* maxwell.cpp:24-33 */
        c.ax = 0.25 * (1 - dx / dt * sin(0.5 * OQ_PI * fmin(1.0, dt / dx)));
        c.ay = 0.25 * (1 - dy / dt * sin(0.5 * OQ_PI * fmin(1.0, dt / dy)));
        c.az = 0.25 * (1 - dz / dt * sin(0.5 * OQ_PI * fmin(1.0, dt / dz)));
        c.bx = 1 - 2 * c.ax;
        c.by = 1 - 2 * c.ay;
        c.bz = 1 - 2 * c.az;
    }
    return c;
}

/* unit offsets */
static const int AX[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

/*
 * Difference of field component comp of F along axis a at cell (i,j,k).
 * fwd = 1: forward difference F[+1]-F[0] (B update); fwd = 0: backward
 * F[0]-F[-1] (E update).
 *  FDTD       plain difference (maxwell.cpp:445-522)
 *  NDFX       difference smoothed along the component's own axis with that
 *             axis' (b, a) (maxwell.cpp:251-328)
 *  five-point b_a (F[+1]-F[0]) + a_a (F[+2]-F[+1]+F[0]-F[-1]) and its
 *             backward mirror, with its own-axis coefficients
 *             (maxwell.cpp:35-125)
 */
static double diff(const oq_state *s, const coef6 *c, const double *F, int comp, int a,
                   int i, int j, int k, int fwd)
{
    const int *u = AX[a];
    /* base pair of nodes: (lo, hi) along a */
    int li = fwd ? i : i - u[0], lj = fwd ? j : j - u[1], lk = fwd ? k : k - u[2];
    int hi_i = li + u[0], hj = lj + u[1], hk = lk + u[2];
    if (s->p.solver == OQ_FDTD) {
        return F[IDX(s, hi_i, hj, hk) + comp] - F[IDX(s, li, lj, lk) + comp];
    } else if (s->p.solver == OQ_NDFX) {
        const int *v = AX[comp];
        double bc = comp == 0 ? c->bx : (comp == 1 ? c->by : c->bz);
        double ac = comp == 0 ? c->ax : (comp == 1 ? c->ay : c->az);
        double d0 = F[IDX(s, hi_i, hj, hk) + comp] - F[IDX(s, li, lj, lk) + comp];
        double dp = F[IDX(s, hi_i + v[0], hj + v[1], hk + v[2]) + comp] - F[IDX(s, li + v[0], lj + v[1], lk + v[2]) + comp];
        double dm = F[IDX(s, hi_i - v[0], hj - v[1], hk - v[2]) + comp] - F[IDX(s, li - v[0], lj - v[1], lk - v[2]) + comp];
        return bc * d0 + ac * (dp + dm);
    } else {
        double ba = a == 0 ? c->bx : (a == 1 ? c->by : c->bz);
        double aa = a == 0 ? c->ax : (a == 1 ? c->ay : c->az);
        /* nodes lo-1, lo, hi, hi+1 along a */
        double fm = rdz(s, F, li - u[0], lj - u[1], lk - u[2], comp);
        double f0 = rdz(s, F, li, lj, lk, comp);
        double f1 = rdz(s, F, hi_i, hj, hk, comp);
        double f2 = rdz(s, F, hi_i + u[0], hj + u[1], hk + u[2], comp);
        return ba * (f1 - f0) + aa * (f2 - f1 + f0 - fm);
    }
}

/* which cells get component comp of B updated (maxwell.cpp:445-484, 251-290, 35-85) */
static int b_updates(const oq_state *s, int comp, int i, int j, int k)
{
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    int c[3] = {i, j, k}, n[3] = {nx, ny, nz};
    if (s->p.solver == OQ_FIVE_POINT) {
        int top = 0;
        for (int a = 0; a < 3; ++a) {
            if (c[a] == n[a] - 1) ++top;
            else if (!in(c[a], 1, n[a] - 3)) return 0;
        }
        return top <= 1;
    }
    int lo = s->p.solver == OQ_FDTD ? 0 : 1;   /* lowest plane updated */
    /* along comp: interior only; the two transverse axes may sit on plane lo,
     * but not both at once */
    int nlow = 0;
    for (int a = 0; a < 3; ++a) {
        if (a == comp) {
            if (!in(c[a], lo + 1, n[a] - 2)) return 0;
        } else {
            if (!in(c[a], lo, n[a] - 2)) return 0;
            if (c[a] == lo) ++nlow;
        }
    }
    return nlow <= 1;
}

/* which cells get component comp of E updated (maxwell.cpp:486-522, 292-328, 87-125) */
static int e_updates(const oq_state *s, int comp, int i, int j, int k)
{
    int c[3] = {i, j, k}, n[3] = {s->p.nx, s->p.ny, s->p.nz};
    int lo = s->p.solver == OQ_FDTD ? 0 : 1;
    int top = s->p.solver == OQ_FIVE_POINT ? 3 : 2;   /* highest index n - top */
    for (int a = 0; a < 3; ++a) {
        int first = (a == comp) ? lo : lo + 1;
        if (!in(c[a], first, n[a] - top)) return 0;
    }
    return 1;
}

/* the B half step on planes [ilo, ihi) (every update reads the grid's own
 * update sets, so a plane range only skips work) */
static void advance_b_range(oq_state *s, const coef6 *cf, int ilo, int ihi)
{
    const int ny = s->p.ny, nz = s->p.nz;
    const double dtdx = s->p.dt / s->p.dx, dtdy = s->p.dt / s->p.dy, dtdz = s->p.dt / s->p.dz;
    const double *E = s->f[OQ_E];
    for (int i = ilo; i < ihi; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (b_updates(s, 0, i, j, k))
                    FB(i, j, k, 0) += 0.5 * (-dtdy * diff(s, cf, E, 2, 1, i, j, k, 1) + dtdz * diff(s, cf, E, 1, 2, i, j, k, 1));
                if (b_updates(s, 1, i, j, k))
                    FB(i, j, k, 1) += 0.5 * (dtdx * diff(s, cf, E, 2, 0, i, j, k, 1) - dtdz * diff(s, cf, E, 0, 2, i, j, k, 1));
                if (b_updates(s, 2, i, j, k))
                    FB(i, j, k, 2) += 0.5 * (dtdy * diff(s, cf, E, 0, 1, i, j, k, 1) - dtdx * diff(s, cf, E, 1, 0, i, j, k, 1));
            }
}

static void advance_e_range(oq_state *s, const coef6 *cf, int ilo, int ihi)
{
    const int ny = s->p.ny, nz = s->p.nz;
    const double dtdx = s->p.dt / s->p.dx, dtdy = s->p.dt / s->p.dy, dtdz = s->p.dt / s->p.dz;
    const double *B = s->f[OQ_B];
    for (int i = ilo; i < ihi; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (e_updates(s, 0, i, j, k))
                    FE(i, j, k, 0) += dtdy * diff(s, cf, B, 2, 1, i, j, k, 0) - dtdz * diff(s, cf, B, 1, 2, i, j, k, 0) - FJ(i, j, k, 0) * s->p.dx;
                if (e_updates(s, 1, i, j, k))
                    FE(i, j, k, 1) += -dtdx * diff(s, cf, B, 2, 0, i, j, k, 0) + dtdz * diff(s, cf, B, 0, 2, i, j, k, 0) - FJ(i, j, k, 1) * s->p.dy;
                if (e_updates(s, 2, i, j, k))
                    FE(i, j, k, 2) += dtdx * diff(s, cf, B, 1, 0, i, j, k, 0) - dtdy * diff(s, cf, B, 0, 1, i, j, k, 0) - FJ(i, j, k, 2) * s->p.dz;
            }
}

static void advance_b_half(oq_state *s, const coef6 *cf) { advance_b_range(s, cf, 0, s->p.nx); }
static void advance_e(oq_state *s, const coef6 *cf) { advance_e_range(s, cf, 0, s->p.nx); }

static int pulse_frame(const oq_state *s, const double L[16], oq_pulse *P);
static int pulse_eval(const oq_pulse *P, double px, double py, double pz, double shift, double ev[3], double bv[3]);

/* whether plane i lies on the total-field side of source n */
static inline int total_side(const oq_state *s, int n, int i)
{
    return s->src_side[n] == 0 ? i >= s->src_plane[n] : i < s->src_plane[n];
}

/*
 * Total-field / scattered-field laser source.  On the source's total-field
 * side of its plane the grid holds the total field, on the other side the
 * field scattered off the incident pulse.  An update that reads across the
 * plane would mix the two, so every update is corrected by the same update
 * applied to the incident field on the far side of the plane: with chi = 1 on
 * the total side, node n of component u gets
 *     sum_m c_m (chi(n) - chi(n + m)) u_inc(n + m),
 * i.e. + L(u_inc on the scattered side) on total nodes and
 * - L(u_inc on the total side) on scattered nodes, L being the solver's own
 * update (all three solvers, boundary planes included).  The incident field
 * is the pulse of f_init_cos carried at c, sampled where each component
 * lives on the staggered grid: E_a at a half cell along a, B_a at half cells
 * across a.  b_field: correct the B half step (input E_inc at time t), else
 * the E step (input B_inc at t).
 */
static void inject(oq_state *s, const coef6 *cf, int b_field, double t)
{
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const double d[3] = {s->p.dx, s->p.dy, s->p.dz};
    oq_state *T = s->tmp;
    const int in_f = b_field ? OQ_E : OQ_B, out_f = b_field ? OQ_B : OQ_E;
    const size_t plane_len = (size_t)ny * nz * 3;
    for (int n = 0; n < s->nsrc; ++n) {
        const double shift = t - s->src_t0[n] * s->p.dt;
        /* plane i lies at (i + window shifts since the source was added) dx
         * in the frame the pulse was given in */
        const int ish = s->window_shifts - s->src_shift0[n];
        /* the updates reach two planes: the incident field on the planes
         * within 4 of the source plane feeds every correction, which is
         * nonzero within 3 of it; the planes cleared and updated cover what
         * those corrections read */
        const int ilo = s->src_plane[n] - 6 > 0 ? s->src_plane[n] - 6 : 0;
        const int ihi = s->src_plane[n] + 7 < nx ? s->src_plane[n] + 7 : nx;
        for (int pass = 0; pass < 2; ++pass) {
            /* pass 0: incident field on the scattered side, read at total
             * nodes; pass 1: on the total side, read at scattered nodes */
            memset(T->f[in_f] + ilo * plane_len, 0, (ihi - ilo) * plane_len * sizeof(double));
            memset(T->f[out_f] + ilo * plane_len, 0, (ihi - ilo) * plane_len * sizeof(double));
            for (int i = ilo; i < ihi; ++i) {
                if (abs(i - s->src_plane[n]) > 4 || total_side(s, n, i) != pass) continue;
                for (int j = 0; j < ny; ++j)
                    for (int k = 0; k < nz; ++k)
                        for (int c = 0; c < 3; ++c) {
                            /* staggered position of component c */
                            double h[3];
                            for (int a = 0; a < 3; ++a) h[a] = (b_field ? a == c : a != c) ? 0.5 : 0.0;
                            double ev[3], bv[3];
                            if (pulse_eval(&s->src[n], (i + ish + h[0]) * d[0], (j + h[1]) * d[1], (k + h[2]) * d[2],
                                           shift, ev, bv))
                                T->f[in_f][IDX(T, i, j, k) + c] = b_field ? ev[c] : bv[c];
                        }
            }
            if (b_field) advance_b_range(T, cf, ilo, ihi);
            else advance_e_range(T, cf, ilo, ihi);
            for (int i = ilo; i < ihi; ++i) {
                if (abs(i - s->src_plane[n]) > 3 || total_side(s, n, i) == pass) continue;
                const double sg = pass == 0 ? 1.0 : -1.0;
                for (int j = 0; j < ny; ++j)
                    for (int k = 0; k < nz; ++k)
                        for (int c = 0; c < 3; ++c)
                            s->f[out_f][IDX(s, i, j, k) + c] += sg * T->f[out_f][IDX(T, i, j, k) + c];
            }
        }
    }
}

void oq_maxwell(oq_state *s)
{
    coef6 cf = solver_coefs(s);
    const double t = s->field_steps * s->p.dt;
    advance_b_half(s, &cf);   /* maxwell.cpp:10-22 */
    if (s->nsrc) inject(s, &cf, 1, t);
    advance_e(s, &cf);
    if (s->nsrc) inject(s, &cf, 0, t + 0.5 * s->p.dt);
    advance_b_half(s, &cf);
    if (s->nsrc) inject(s, &cf, 1, t + s->p.dt);
    ++s->field_steps;
}

void oq_laser_field(oq_state *s, const double L[16], double shift, int staggered, double *E, double *B)
{
    oq_pulse P;
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const double d[3] = {s->p.dx, s->p.dy, s->p.dz};
    memset(E, 0, (size_t)s->ncell * 3 * sizeof(double));
    memset(B, 0, (size_t)s->ncell * 3 * sizeof(double));
    if (pulse_frame(s, L, &P)) return;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                for (int c = 0; c < 3; ++c)
                    for (int f = 0; f < 2; ++f) {
                        double h[3], ev[3], bv[3];
                        for (int a = 0; a < 3; ++a) h[a] = staggered ? ((f == 0 ? a == c : a != c) ? 0.5 : 0.0) : 0.0;
                        if (pulse_eval(&P, (i + h[0]) * d[0], (j + h[1]) * d[1], (k + h[2]) * d[2], shift, ev, bv))
                            (f == 0 ? E : B)[IDX(s, i, j, k) + c] = f == 0 ? ev[c] : bv[c];
                    }
}

int oq_add_laser_source(oq_state *s, const double L[16], int side, int plane)
{
    if (s->nsrc == OQ_MAX_SOURCES) return -1;
    if (side == 0 ? !(plane >= 2 && plane <= s->p.nx - 6) : !(side == 1 && plane >= 6 && plane <= s->p.nx - 2))
        return -1;
    oq_pulse P;
    if (pulse_frame(s, L, &P)) return -1;
    if (!s->tmp) s->tmp = oq_create(&s->p);
    s->src[s->nsrc] = P;
    s->src_side[s->nsrc] = side;
    s->src_plane[s->nsrc] = plane;
    s->src_t0[s->nsrc] = s->field_steps;
    s->src_shift0[s->nsrc] = s->window_shifts;
    ++s->nsrc;
    return 0;
}

/*
 * init_boundaries: zero field components on boundary planes, once
 * (maxwell.cpp:532-597 FDTD, 338-438 NDFX, 135-236 five-point).  On every
 * plane the two transverse indices run over [0, n-2].
 *   FDTD:  low planes: the two E components tangential to the plane and the
 *          normal B component; high planes (n-1): all six.
 *   NDFX:  low (0) and high (n-1) planes: all six; plane 1: tangential E and
 *          normal B.
 *   five-point: planes 0 and n-1: all six; plane 1: tangential E; plane n-2:
 *          tangential B.
 */
static void zero_plane(oq_state *s, int a, int pos, int emask, int bmask)
{
    int n[3] = {s->p.nx, s->p.ny, s->p.nz};
    int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
    for (int u = 0; u < n[b1] - 1; ++u)
        for (int v = 0; v < n[b2] - 1; ++v) {
            int c[3];
            c[a] = pos; c[b1] = u; c[b2] = v;
            for (int comp = 0; comp < 3; ++comp) {
                if (emask & (1 << comp)) FE(c[0], c[1], c[2], comp) = 0;
                if (bmask & (1 << comp)) FB(c[0], c[1], c[2], comp) = 0;
            }
        }
}

void oq_init_boundaries(oq_state *s)
{
    int n[3] = {s->p.nx, s->p.ny, s->p.nz};
    for (int a = 0; a < 3; ++a) {
        int normal = 1 << a, tang = 7 & ~normal;
        if (s->p.solver == OQ_FDTD) {
            zero_plane(s, a, 0, tang, normal);
            zero_plane(s, a, n[a] - 1, 7, 7);
        } else if (s->p.solver == OQ_NDFX) {
            zero_plane(s, a, 0, 7, 7);
            zero_plane(s, a, 1, tang, normal);
            zero_plane(s, a, n[a] - 1, 7, 7);
        } else {
            zero_plane(s, a, 0, 7, 7);
            zero_plane(s, a, n[a] - 1, 7, 7);
            zero_plane(s, a, 1, tang, 0);
            zero_plane(s, a, n[a] - 2, 0, tang);
        }
    }
    oq_interpolate_be(s);   /* main.cpp:2288-2289 */
}

/* ------------------------------------------------------------------ */
/* QED (mc.cpp:26-139)                                                 */

double oq_chi(double e_s, const double e[3], const double b[3], const double u[3], double g)
{
    double a = g * e[0] + u[1] * b[2] - u[2] * b[1];
    double s2 = a * a;
    a = g * e[1] - u[0] * b[2] + u[2] * b[0];
    s2 += a * a;
    a = g * e[2] + u[0] * b[1] - u[1] * b[0];
    s2 += a * a;
    a = u[0] * e[0] + u[1] * e[1] + u[2] * e[2];
    s2 -= a * a;
    return sqrt(s2) / e_s;
}

/* photon emission spectrum dW/dr (mc.cpp:40-74); for chi <= 0.133 the
 * fraction r is rescaled in place by r_max = 1/(1+2/(15 chi)) */
double oq_rate_w(double e_s, double g, double *rp, double chi)
{
    double r = *rp;
    if ((1 - r) * g < 1 || r == 0 || chi == 0) return 0;
    double scale = 1;
    if (chi <= 0.133) {
        double r_max = 1 / (1 + 2 / (3 * 5 * chi));
        r = r_max * r;
        *rp = r;
        scale = r_max;
    }
    double d = r / ((1 - r) * chi);
    double a = pow(d, 1.0 / 3);
    double kap = a * a;
    double t1 = 1 / sqrt(sqrt((kap + 0.80049) * (kap + 0.80049) * (kap + 0.80049)));
    double t2 = (2 * (1 - r) + r * r) * chi * a / r * sqrt(sqrt(kap + 0.70861 * (1 - 0.65 * kap / (1 + kap * kap))));
    return -scale * 0.0020591 * e_s / (g * g) * exp(-2 * d / 3) * (t1 - t2);
}

/* Breit-Wheeler pair spectrum (mc.cpp:76-97) */
double oq_rate_tilde_w(double e_s, double g, double r, double chi)
{
    if (chi == 0 || r * g < 1 || (1 - r) * g < 1) return 0;
    double d = 1 / (r * (1 - r) * chi);
    double a = pow(d, 1.0 / 3);
    double kap = a * a;
    double t1 = 1 / sqrt(sqrt((kap + 0.80049) * (kap + 0.80049) * (kap + 0.80049)));
    double t2 = (2 * r * (1 - r) - 1) * chi * a * sqrt(sqrt(kap + 0.70861 * (1 - 0.65 * kap / (1 + kap * kap))));
    return 0.0020591 * e_s / (g * g) * exp(-2 * d / 3) * (t1 - t2);
}

/* Schwinger pair rate per unit volume and time (mc.cpp:99-116) */
double oq_vacuum_rate(double e_s, const double e[3], const double b[3])
{
    double f = 0.5 * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2] - b[0] * b[0] - b[1] * b[1] - b[2] * b[2]);
    double gg = e[0] * b[0] + e[1] * b[1] + e[2] * b[2];
    double a = sqrt(f * f + gg * gg);
    double xi = sqrt(a + f), th = sqrt(a - f);
    if (xi == 0) return 0;
    if (th != 0) return e_s * e_s / (4 * OQ_PI) * xi * th / tanh(OQ_PI * th / xi) * exp(-OQ_PI * e_s / xi);
    return e_s * e_s / (4 * OQ_PI) * xi * xi / OQ_PI * exp(-OQ_PI * e_s / xi);
}

/* bear_particle (mc.cpp:118-139) */
static oq_particle *bear(oq_state *s, double cmr, const double pos[3], const double dir[3],
                         double g, double chi, double q)
{
    oq_particle *p = new_particle(s);
    p->cmr = cmr;
    p->q = q;
    p->g = g;
    p->chi = chi;
    p->x = pos[0]; p->y = pos[1]; p->z = pos[2];
    double d2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
    double a = (cmr == 0) ? g / sqrt(d2) : sqrt((g * g - 1) / d2);
    p->ux = a * dir[0]; p->uy = a * dir[1]; p->uz = a * dir[2];
    p->id = 0;
    return p;
}

/* place() after a move (plist.cpp:7-40): keep if inside, else delete */
static void place(oq_state *s, oq_particle *p)
{
    if (!is_inside(s, (int)p->x, (int)p->y, (int)p->z)) p->alive = 0;
}

static void displacement(const oq_state *s, const oq_particle *p, double d[3])
{
    /* get_displacement + /d_alpha (containers.cpp:58-67, advance.cpp:281-284) */
    double a = s->p.dt / p->g;
    d[0] = p->ux * a / s->p.dx;
    d[1] = p->uy * a / s->p.dy;
    d[2] = p->uz * a / s->p.dz;
}

/* ------------------------------------------------------------------ */
/* the particle phase (advance.cpp:90-387)                             */

void oq_padvance(oq_state *s)
{
    const double dt = s->p.dt, dx = s->p.dx, dy = s->p.dy, dz = s->p.dz;
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    /* p_boundary (plist.cpp:73-102): erase particles in layers 0, n-2, n-1 */
    for (int64_t n = 0; n < s->np; ++n) {
        oq_particle *p = &s->part[n];
        int i = (int)p->x, j = (int)p->y, k = (int)p->z;
        if (i == 0 || i == nx - 2 || i == nx - 1 || j == 0 || j == ny - 2 || j == ny - 1 ||
            k == 0 || k == nz - 2 || k == nz - 1)
            p->alive = 0;
    }
    memset(s->f[OQ_J], 0, (size_t)s->ncell * 3 * sizeof(double));   /* advance.cpp:112-123 */

    const int64_t n0 = s->np;   /* particles born in this pass are not visited */
    for (int64_t n = 0; n < n0; ++n) {
        oq_particle *p = &s->part[n];
        if (!p->alive) continue;
        double e[3], b[3], d[3];
        gather(s, s->f[OQ_E], p->x, p->y, p->z, e);
        gather(s, s->f[OQ_BE], p->x, p->y, p->z, b);
        b[2] += s->p.external_bz;

        if (p->cmr == 0) {   /* photon, advance.cpp:155-252 */
            int pair = 0;
            double u_int[3] = {p->ux, p->uy, p->uz}, g_prev = p->g, r = 0;
            if (s->p.qed) {
                double chi_prev = p->chi;
                p->chi = oq_chi(s->p.e_s, e, b, u_int, p->g);
                double chibar = 0.5 * (p->chi + chi_prev);
                r = oq_rand(s);
                pair = !(oq_rand(s) >= oq_rate_tilde_w(s->p.e_s, g_prev, r, chibar) * g_prev * dt);
            }
            if (!pair) {
                displacement(s, p, d);
                p->x += d[0]; p->y += d[1]; p->z += d[2];
                place(s, p);
            } else {
                /* the photon may move in memory when the array grows: copy it */
                oq_particle ph = *p;
                double pos[3] = {ph.x, ph.y, ph.z};
                for (int sp = 0; sp < 2; ++sp) {
                    double cmr = sp == 0 ? -1 : 1;
                    double gb = sp == 0 ? r * g_prev : (1 - r) * ph.g;
                    double chib = sp == 0 ? r * ph.chi : (1 - r) * ph.chi;
                    double qb = sp == 0 ? -ph.q : ph.q;
                    double dir[3] = {ph.ux, ph.uy, ph.uz};
                    oq_particle *bn = bear(s, cmr, pos, dir, gb, chib, qb);
                    double up[3] = {bn->ux, bn->uy, bn->uz}, gp = bn->g;
                    push_particle(s, bn, e, b, dt);
                    double ui[3] = {0.5 * (bn->ux + up[0]), 0.5 * (bn->uy + up[1]), 0.5 * (bn->uz + up[2])};
                    bn->chi = oq_chi(s->p.e_s, e, b, ui, 0.5 * (bn->g + gp));
                    displacement(s, bn, d);
                    jdeposition(s, bn, d);
                    place(s, bn);
                }
                s->part[n].alive = 0;   /* photon destroyed */
            }
        } else if (fabs(p->cmr) == 1) {   /* electron / positron, advance.cpp:253-359 */
            int emit = 0;
            double u_prev[3] = {p->ux, p->uy, p->uz}, g_prev = p->g, r = 0;
            if (s->p.qed) {
                double chi_prev = p->chi;
                push_particle(s, p, e, b, dt);
                double ui[3] = {0.5 * (p->ux + u_prev[0]), 0.5 * (p->uy + u_prev[1]), 0.5 * (p->uz + u_prev[2])};
                p->chi = oq_chi(s->p.e_s, e, b, ui, 0.5 * (p->g + g_prev));
                double chibar = 0.5 * (p->chi + chi_prev);
                r = oq_rand(s);
                emit = !(oq_rand(s) >= oq_rate_w(s->p.e_s, g_prev, &r, chibar) * g_prev * dt);
            } else {
                push_particle(s, p, e, b, dt);
            }
            if (!emit) {
                displacement(s, p, d);
                jdeposition(s, p, d);
                place(s, p);
            } else {
                /* photon emission (advance.cpp:291-357) */
                double a = 0.5 * dt / g_prev;
                double pos[3] = {p->x - a * u_prev[0] / dx, p->y - a * u_prev[1] / dy, p->z - a * u_prev[2] / dz};
                double chi_now = p->chi;
                double q_abs = fabs(p->q);
                oq_particle *ph = bear(s, 0, pos, u_prev, r * g_prev, r * chi_now, q_abs);
                double a3 = 3 * a * g_prev / ph->g;
                ph->x += a3 * ph->ux / dx;
                ph->y += a3 * ph->uy / dy;
                ph->z += a3 * ph->uz / dz;
                place(s, ph);
                p = &s->part[n];   /* the array may have moved */
                p->chi = (1 - r) * p->chi;
                p->g = (1 - r) * g_prev;
                double up2 = u_prev[0] * u_prev[0] + u_prev[1] * u_prev[1] + u_prev[2] * u_prev[2];
                double sc = sqrt((p->g * p->g - 1) / up2);
                p->ux = u_prev[0] * sc; p->uy = u_prev[1] * sc; p->uz = u_prev[2] * sc;
                a = 0.5 * dt / g_prev;
                p->x -= a * u_prev[0] / dx; p->y -= a * u_prev[1] / dy; p->z -= a * u_prev[2] / dz;
                a = a * g_prev / p->g;
                p->x += a * p->ux / dx; p->y += a * p->uy / dy; p->z += a * p->uz / dz;
                if (is_inside(s, (int)p->x, (int)p->y, (int)p->z)) {
                    gather(s, s->f[OQ_E], p->x, p->y, p->z, e);
                    gather(s, s->f[OQ_BE], p->x, p->y, p->z, b);
                    b[2] += s->p.external_bz;
                    push_particle(s, p, e, b, dt);
                    displacement(s, p, d);
                    jdeposition(s, p, d);
                    place(s, p);
                } else {
                    p->alive = 0;
                }
            }
        } else {   /* ion, advance.cpp:360-371 */
            push_particle(s, p, e, b, dt);
            displacement(s, p, d);
            jdeposition(s, p, d);
            place(s, p);
        }
    }
    compact(s);
}

/* pmerging (mc.cpp:202-462) */
static int merge_species(double cmr, int n_ion, const double *ion_cmr)
{
    if (cmr == -1) return 0;
    if (cmr == 1) return 1;
    if (cmr == 0) return 2;
    for (int n = 0; n < n_ion; ++n)
        if (cmr == ion_cmr[n]) return 3 + n;
    return -1;
}

/* _ppd (mc.cpp:198-200) */
static double merge_weight(double q, double q0)
{
    return 1 / (1 + pow(2 * q / q0, 4));
}

void oq_pmerging(oq_state *s, int mode, const double *ppd, int n_ion, const double *ion_cmr)
{
    enum { NS = 3 + 4 };
    double tot[NS] = {0}, del[NS] = {0}, cnt[NS] = {0};
    const int ns = 3 + n_ion;
    for (int64_t n = 0; n < s->np; ++n) {
        oq_particle *p = &s->part[n];
        const int sp = p->alive ? merge_species(p->cmr, n_ion, ion_cmr) : -1;
        if (sp < 0) continue;
        tot[sp] += fabs(p->q);
        cnt[sp] += 1;
    }
    if (mode == 1) {
        for (int64_t n = 0; n < s->np; ++n) {
            oq_particle *p = &s->part[n];
            const int sp = p->alive ? merge_species(p->cmr, n_ion, ion_cmr) : -1;
            if (sp < 0) continue;
            if (oq_rand(s) < ppd[0]) {
                del[sp] += fabs(p->q);
                p->alive = 0;
            }
        }
        for (int sp = 0; sp < ns; ++sp)
            if (tot[sp] - del[sp] != 0) tot[sp] = tot[sp] / (tot[sp] - del[sp]);
    } else {
        double av[NS], kept[NS] = {0};
        for (int sp = 0; sp < ns; ++sp) av[sp] = tot[sp] / cnt[sp];
        for (int64_t n = 0; n < s->np; ++n) {
            oq_particle *p = &s->part[n];
            const int sp = p->alive ? merge_species(p->cmr, n_ion, ion_cmr) : -1;
            if (sp < 0) continue;
            if (oq_rand(s) < ppd[sp] * merge_weight(fabs(p->q), av[sp])) p->alive = 0;
        }
        for (int64_t n = 0; n < s->np; ++n) {
            oq_particle *p = &s->part[n];
            const int sp = p->alive ? merge_species(p->cmr, n_ion, ion_cmr) : -1;
            if (sp < 0) continue;
            p->q = p->q / (1 - ppd[sp] * merge_weight(fabs(p->q), av[sp]));
            kept[sp] += fabs(p->q);
        }
        for (int sp = 0; sp < ns; ++sp) tot[sp] = kept[sp] != 0 ? tot[sp] / kept[sp] : 0;
    }
    for (int64_t n = 0; n < s->np; ++n) {
        oq_particle *p = &s->part[n];
        const int sp = p->alive ? merge_species(p->cmr, n_ion, ion_cmr) : -1;
        if (sp >= 0) p->q *= tot[sp];
    }
    compact(s);
}

/* birth_from_vacuum (mc.cpp:142-196) */
void oq_birth_from_vacuum(oq_state *s)
{
    if (!s->p.qed) return;
    const double dt = s->p.dt;
    const double vol = s->p.dx * s->p.dy * s->p.dz;
    double dir[3] = {1, 0, 0};
    for (int i = 2; i < s->p.nx - 2; ++i)
        for (int j = 2; j < s->p.ny - 2; ++j)
            for (int k = 2; k < s->p.nz - 2; ++k) {
                double pos[3] = {i + 0.5, j + 0.5, k + 0.5}, e[3], b[3];
                gather(s, s->f[OQ_E], pos[0], pos[1], pos[2], e);
                gather(s, s->f[OQ_BE], pos[0], pos[1], pos[2], b);
                if (oq_rand(s) < oq_vacuum_rate(s->p.e_s, e, b) * dt * vol) {
                    oq_particle *p1 = bear(s, -1, pos, dir, 1, 0, -s->p.q_vac);
                    push_particle(s, p1, e, b, -dt / 2);
                    oq_particle *p2 = bear(s, 1, pos, dir, 1, 0, s->p.q_vac);
                    push_particle(s, p2, e, b, -dt / 2);
                }
            }
}

/* ------------------------------------------------------------------ */
/* diagnostics (fout.cpp:41-138), n_sr = 1: the whole grid             */

void oq_diagnostics(oq_state *s, oq_diag *d)
{
    memset(d, 0, sizeof(*d));
    const double *E = s->f[OQ_E], *B = s->f[OQ_B];
    double ef = 0;
    for (int64_t c = 0; c < s->ncell * 3; ++c) ef += E[c] * E[c] + B[c] * B[c];
    for (int64_t n = 0; n < s->np; ++n) {
        const oq_particle *p = &s->part[n];
        if (!p->alive) continue;
        if (p->cmr == -1) { d->N_e -= p->q; d->n_e++; d->energy_e -= (p->g - 1) * p->q; }
        else if (p->cmr == 1) { d->N_p += p->q; d->n_p++; d->energy_p += (p->g - 1) * p->q; }
        else if (p->cmr == 0) { d->N_ph += p->q; d->n_ph++; d->energy_ph += p->g * p->q; }
        else { d->n_ion++; d->energy_ion += (p->g - 1) * p->q / p->cmr; }
    }
    d->N_e *= s->p.norm_n; d->N_p *= s->p.norm_n; d->N_ph *= s->p.norm_n;
    d->energy_f = ef * s->p.norm_ef;
    d->energy_e *= s->p.norm_ep; d->energy_p *= s->p.norm_ep;
    d->energy_ph *= s->p.norm_ep; d->energy_ion *= s->p.norm_ep;
}

/* compute_rho (fout.cpp:7-39) with rhodeposition (deposition.cpp:358-429):
 * node charge with trilinear weights; electrons into the jx slot, positrons
 * into jy, photons into jz (ions are not stored here) */
void oq_compute_rho(oq_state *s)
{
    memset(s->f[OQ_J], 0, (size_t)s->ncell * 3 * sizeof(double));
    for (int64_t n = 0; n < s->np; ++n) {
        const oq_particle *p = &s->part[n];
        if (!p->alive) continue;
        int slot = p->cmr == -1 ? 0 : (p->cmr == 1 ? 1 : (p->cmr == 0 ? 2 : -1));
        if (slot < 0) continue;
        int i = (int)floor(p->x), j = (int)floor(p->y), k = (int)floor(p->z);
        if (!is_inside(s, i, j, k)) continue;
        double xa = p->x - i, xb = 1 - xa, ya = p->y - j, yb = 1 - ya, za = p->z - k, zb = 1 - za;
        FJ(i, j, k, slot) += p->q * xb * yb * zb;
        FJ(i + 1, j, k, slot) += p->q * xa * yb * zb;
        FJ(i, j + 1, k, slot) += p->q * xb * ya * zb;
        FJ(i, j, k + 1, slot) += p->q * xb * yb * za;
        FJ(i + 1, j + 1, k, slot) += p->q * xa * ya * zb;
        FJ(i + 1, j, k + 1, slot) += p->q * xa * yb * za;
        FJ(i, j + 1, k + 1, slot) += p->q * xb * ya * za;
        FJ(i + 1, j + 1, k + 1, slot) += p->q * xa * ya * za;
    }
}

/* ------------------------------------------------------------------ */
/* one step, src/main.cpp:2319-2332 at n_sr = 1                        */

void oq_step(oq_state *s, int nsteps, oq_diag *diag)
{
    for (int l = 0; l < nsteps; ++l) {
        oq_birth_from_vacuum(s);
        oq_padvance(s);
        if (diag) oq_diagnostics(s, &diag[l]);
        oq_maxwell(s);
        oq_interpolate_be(s);   /* fadvance, advance.cpp:13-18 */
    }
}

/* ------------------------------------------------------------------ */
/* laser IC: f_init_cos, sscos = 0 (init.cpp:6-210)                    */


static int pulse_frame(const oq_state *s, const double L[16], oq_pulse *P)
{
    /* L: a0y a0z xsigma ysigma zsigma x0 b_sign x1 phase y0 z0 append phi xt yt zt */
    P->a0y = L[0];
    P->a0z = L[1];
    /* half lengths of the cos^2 envelope with the same \int a^2 as a Gaussian (init.cpp:34-37) */
    P->xs = L[2] * 2 * sqrt(2 * OQ_PI) / 3;
    P->ys = L[3] * 2 * sqrt(2 * OQ_PI) / 3;
    P->zs = L[4] * 2 * sqrt(2 * OQ_PI) / 3;
    P->x0 = L[5];
    P->b_sign = L[6] != 0;
    double x1 = L[7];
    P->phase = L[8];
    P->y0 = L[9];
    P->z0 = L[10];
    double phi = L[12];
    P->xt = L[13];
    P->yt = L[14];
    P->zt = L[15];
    P->y12 = 0.5 * s->p.ny * s->p.dy;
    P->z12 = 0.5 * s->p.nz * s->p.dz;
    /* propagation frame: r0 points from the pulse centre to the target (init.cpp:55-93) */
    double rn = sqrt(x1 * x1 + P->y0 * P->y0 + P->z0 * P->z0);
    if (!(rn > 0)) return -1;
    double *r0 = P->r0, *ey_ = P->ey, *ez_ = P->ez;
    r0[0] = -x1 / rn; r0[1] = -P->y0 / rn; r0[2] = -P->z0 / rn;
    double t = sqrt(r0[0] * r0[0] + r0[1] * r0[1]);
    if (t != 0) { ey_[0] = -r0[1] / t; ey_[1] = r0[0] / t; ey_[2] = 0; }
    else { ey_[0] = 0; ey_[1] = 1; ey_[2] = 0; }
    ez_[0] = r0[1] * ey_[2] - r0[2] * ey_[1];
    ez_[1] = r0[2] * ey_[0] - r0[0] * ey_[2];
    ez_[2] = r0[0] * ey_[1] - r0[1] * ey_[0];
    if (phi != 0) {
        double c = cos(phi), sn = sin(phi), y[3], z[3];
        for (int m = 0; m < 3; ++m) { y[m] = ey_[m]; z[m] = ez_[m]; }
        for (int m = 0; m < 3; ++m) { ey_[m] = y[m] * c + z[m] * sn; ez_[m] = z[m] * c - y[m] * sn; }
    }
    return 0;
}

/* Fields of the pulse at the point (px, py, pz) (c/omega units from the
 * subdomain origin) after it has travelled a distance `shift` along its
 * direction of propagation: +r0 when b_sign, -r0 otherwise (E x B).  shift = 0
 * is f_init_cos itself.  Returns 0 (and leaves ev, bv untouched) outside the
 * envelope. */
static int pulse_eval(const oq_pulse *P, double px, double py, double pz, double shift, double ev[3], double bv[3])
{
    double X = px - P->x0 - P->xt, Y = py - P->y12 - P->y0 - P->yt, Z = pz - P->z12 - P->z0 - P->zt;
    double x = P->r0[0] * X + P->r0[1] * Y + P->r0[2] * Z;
    double y = P->ey[0] * X + P->ey[1] * Y + P->ey[2] * Z;
    double z = P->ez[0] * X + P->ez[1] * Y + P->ez[2] * Z;
    if (shift != 0) x -= P->b_sign ? shift : -shift;
    const double xs = P->xs, ys = P->ys, zs = P->zs, a0y = P->a0y, a0z = P->a0z;
    double xi = x + P->phase;
    if (!(x < xs && x > -xs && y < ys && y > -ys && z < zs && z > -zs)) return 0;
    double cx = cos(OQ_PI * x / 2 / xs), cy = cos(OQ_PI * y / 2 / ys), cz = cos(OQ_PI * z / 2 / zs);
    double sx = sin(OQ_PI * x / 2 / xs), sy = sin(OQ_PI * y / 2 / ys), sz = sin(OQ_PI * z / 2 / zs);
    double tr = cy * cy * cz * cz;
    /* transverse fields with the longitudinal-envelope correction
     * and the divergence-free longitudinal components */
    double fy = a0y * tr * (cos(xi) * cx * cx - OQ_PI / xs * sin(xi) * cx * sx);
    double fz = a0z * tr * (sin(xi) * cx * cx + OQ_PI / xs * cos(xi) * cx * sx);
    double fx = a0y * sin(xi) * cx * cx * cz * cz * OQ_PI / ys * cy * sy
              - a0z * cos(xi) * cx * cx * cy * cy * OQ_PI / zs * cz * sz;
    double gx = a0z * cos(xi) * cx * cx * OQ_PI / ys * cz * cz * cy * sy
              + a0y * sin(xi) * cx * cx * cy * cy * OQ_PI / zs * cz * sz;
    double gy = -fz, gz = fy;
    if (!P->b_sign) { gx = -gx; gy = -gy; gz = -gz; }
    for (int m = 0; m < 3; ++m) {
        ev[m] = fx * P->r0[m] + fy * P->ey[m] + fz * P->ez[m];
        bv[m] = gx * P->r0[m] + gy * P->ey[m] + gz * P->ez[m];
    }
    return 1;
}

void oq_init_laser_cos(oq_state *s, double a0y, double a0z, double xsigma, double ysigma,
                       double zsigma, double x0, int b_sign, double x1, double phase,
                       double y0, double z0, int append, double phi,
                       double xtarget, double ytarget, double ztarget)
{
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    const double dx = s->p.dx, dy = s->p.dy, dz = s->p.dz;
    const double L[16] = {a0y, a0z, xsigma, ysigma, zsigma, x0, b_sign, x1, phase, y0, z0, append, phi,
                          xtarget, ytarget, ztarget};
    oq_pulse P;
    if (pulse_frame(s, L, &P)) return;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                double *E = &FE(i, j, k, 0), *B = &FB(i, j, k, 0);
                double ev[3], bv[3];
                if (pulse_eval(&P, i * dx, j * dy, k * dz, 0.0, ev, bv)) {
                    for (int m = 0; m < 3; ++m) {
                        if (append) { E[m] += ev[m]; B[m] += bv[m]; }
                        else { E[m] = ev[m]; B[m] = bv[m]; }
                    }
                } else if (!append) {
                    for (int m = 0; m < 3; ++m) { E[m] = 0; B[m] = 0; }
                }
            }
}

/* ------------------------------------------------------------------ */
/* synthetic uniform plasma from the shared loader (philox.h)          */

void oq_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    oq_u32x4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    oq_u32x4 r = oq_philox4x32_10(c, key[0], key[1]);
    for (int m = 0; m < 4; ++m) out[m] = r.v[m];
}

int64_t oq_synth_plasma(int nx, int ny, int nz, int ppc, double density_ncr, double sigma_u,
                        uint64_t seed, int ilo, int ihi, double *rec8, int32_t *id)
{
    /* electrons in the is_inside cells (plist.cpp:42-48) with ilo <= i < ihi;
     * q = -n/ppc as fill_cell_by_particles gives (init.cpp:413-414) */
    if (ilo < 1) ilo = 1;
    if (ihi > nx - 2) ihi = nx - 2;
    double q = -density_ncr / ppc;
    int64_t m = 0;
    for (int i = ilo; i < ihi; ++i)
        for (int j = 1; j < ny - 2; ++j)
            for (int k = 1; k < nz - 2; ++k) {
                uint64_t cell = ((uint64_t)i * ny + j) * nz + k;
                for (int s = 0; s < ppc; ++s, ++m) {
                    if (!rec8) continue;
                    double off[3], u[3];
                    uint64_t gid = cell * (uint64_t)ppc + s;
                    oq_synth_particle(seed, gid, sigma_u, off, u);
                    double *r = rec8 + 8 * m;
                    r[0] = i + off[0]; r[1] = j + off[1]; r[2] = k + off[2];
                    r[3] = u[0]; r[4] = u[1]; r[5] = u[2];
                    r[6] = q; r[7] = -1;
                    if (id) id[m] = (int32_t)gid;
                }
            }
    return m;
}

int64_t oq_load_uniform(oq_state *s, int ppc, double density_ncr, double sigma_u,
                        uint64_t seed, int ilo, int ihi)
{
    int64_t n = oq_synth_plasma(s->p.nx, s->p.ny, s->p.nz, ppc, density_ncr, sigma_u, seed, ilo, ihi, NULL, NULL);
    double *rec = (double *)malloc((size_t)(n ? n : 1) * 8 * sizeof(double));
    int32_t *id = (int32_t *)malloc((size_t)(n ? n : 1) * sizeof(int32_t));
    oq_synth_plasma(s->p.nx, s->p.ny, s->p.nz, ppc, density_ncr, sigma_u, seed, ilo, ihi, rec, id);
    for (int64_t m = 0; m < n; ++m) {
        const double *r = rec + 8 * m;
        oq_particle *p = new_particle(s);
        p->x = r[0]; p->y = r[1]; p->z = r[2];
        p->ux = r[3]; p->uy = r[4]; p->uz = r[5];
        p->g = sqrt(1 + r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
        p->q = r[6];
        p->cmr = -1;
        p->chi = 0;
        p->id = id[m];
    }
    free(rec);
    free(id);
    return n;
}
