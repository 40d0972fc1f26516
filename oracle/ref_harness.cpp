/*
 * oracle/ref_harness.cpp -- single-rank driver for the compiled reference.
 *
 * TEST INFRASTRUCTURE.  oracle/Makefile links this file with the reference's
 * own translation units (src/{advance,containers,deposition,fout,init,
 * low-level,maxwell,mc,parameters,plist,pusher}.cpp compiled where they lie
 * under the reference tree).  It replaces src/main.cpp, which needs MPI: it
 * defines the seven globals the hot-path units declare extern
 * (src/advance.cpp:9-11, src/low-level.cpp:12-14, src/plist.cpp:4-5) and
 * drives spatial_region (include/main.h:17-170) the way main() does for
 * n_sr = 1 (src/main.cpp:2276-2332).  The binary goes to oracle/_ref/ and is
 * never committed.
 *
 * Usage: quill_ref <input case file> <output file>
 * Case-file layout: oracle/golden_format.h.
 */
#include <chrono>
#include <fstream>
#include <string>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "main.h"
#include "pusher.h"
#include "golden_format.h"
#include "philox.h"

/* globals owned by src/main.cpp in the reference build */
bool qed_enabled = false;
bool dump_photons = false;
bool catching_enabled = false;
double lambda = 1e-4;
int mpi_rank = 0;
int n_sr = 1;
int nm = 4;

static void die(const char *m)
{
    std::fprintf(stderr, "quill_ref: %s\n", m);
    std::exit(2);
}

static void rd(FILE *f, void *p, size_t n)
{
    if (n && std::fread(p, 1, n, f) != n) die("short read");
}
static void wr(FILE *f, const void *p, size_t n)
{
    if (n && std::fwrite(p, 1, n, f) != n) die("short write");
}

static maxwell_solver_enum solver_of(int s)
{
    return s == 0 ? maxwell_solver_enum::FDTD : (s == 1 ? maxwell_solver_enum::NDFX : maxwell_solver_enum::FIVE_POINT);
}
static pusher_enum pusher_of(int p) { return p == 1 ? pusher_enum::BORIS : pusher_enum::VAY; }

/* create a spatial_region as main() does for rank 0 of 1 (src/main.cpp:2275-2277) */
static spatial_region *make_region(const qg_header &h, std::vector<double> &icmr)
{
    spatial_region *sr = new spatial_region;
    icmr.assign(h.ion_cmr, h.ion_cmr + (h.n_ion > 0 ? h.n_ion : 0));
    sr->init(0, h.dx, h.dy, h.dz, h.dt, h.e_s, 1, 1, 1, h.n_ion, icmr.empty() ? nullptr : icmr.data(),
             "/tmp", solver_of(h.solver), pusher_of(h.pusher));
    sr->create_arrays(h.nx, h.ny, h.nz, h.seed);
    return sr;
}

static void set_start(spatial_region *sr, const qg_header &h)
{
    /* padvance walks each list from pl.start (src/advance.cpp:131) */
    for (int i = 0; i < h.nx; ++i)
        for (int j = 0; j < h.ny; ++j)
            for (int k = 0; k < h.nz; ++k) sr->cp[i][j][k].pl.start = sr->cp[i][j][k].pl.head;
}

static void add_particle(spatial_region *sr, const double *r, int id)
{
    particle *p = sr->new_particle();
    p->x = r[0]; p->y = r[1]; p->z = r[2];
    p->ux = r[3]; p->uy = r[4]; p->uz = r[5];
    p->q = r[6];
    p->cmr = r[7];
    double u2 = p->ux * p->ux + p->uy * p->uy + p->uz * p->uz;
    p->g = (p->cmr != 0) ? std::sqrt(1 + u2) : std::sqrt(u2);
    p->chi = 0;
    p->trn = id;
    p->next = nullptr;
    p->previous = nullptr;
    sr->place(*p);
}

static void put_fields(spatial_region *sr, const qg_header &h, const double *E, const double *B)
{
    size_t c = 0;
    for (int i = 0; i < h.nx; ++i)
        for (int j = 0; j < h.ny; ++j)
            for (int k = 0; k < h.nz; ++k, c += 3) {
                sr->ce[i][j][k].ex = E[c]; sr->ce[i][j][k].ey = E[c + 1]; sr->ce[i][j][k].ez = E[c + 2];
                sr->cb[i][j][k].bx = B[c]; sr->cb[i][j][k].by = B[c + 1]; sr->cb[i][j][k].bz = B[c + 2];
            }
}

static void dump_fields(spatial_region *sr, const qg_header &h, FILE *f, bool all)
{
    size_t n = (size_t)h.nx * h.ny * h.nz * 3;
    std::vector<double> a(n), b(n), j(n), be(n);
    size_t c = 0;
    for (int i = 0; i < h.nx; ++i)
        for (int jj = 0; jj < h.ny; ++jj)
            for (int k = 0; k < h.nz; ++k, c += 3) {
                a[c] = sr->ce[i][jj][k].ex; a[c + 1] = sr->ce[i][jj][k].ey; a[c + 2] = sr->ce[i][jj][k].ez;
                b[c] = sr->cb[i][jj][k].bx; b[c + 1] = sr->cb[i][jj][k].by; b[c + 2] = sr->cb[i][jj][k].bz;
                j[c] = sr->cj[i][jj][k].jx; j[c + 1] = sr->cj[i][jj][k].jy; j[c + 2] = sr->cj[i][jj][k].jz;
                be[c] = sr->cbe[i][jj][k].bex; be[c + 1] = sr->cbe[i][jj][k].bey; be[c + 2] = sr->cbe[i][jj][k].bez;
            }
    wr(f, a.data(), n * 8);
    wr(f, b.data(), n * 8);
    if (all) {
        wr(f, j.data(), n * 8);
        wr(f, be.data(), n * 8);
    }
}

static void dump_particles(spatial_region *sr, const qg_header &h, FILE *f)
{
    std::vector<double> rec;
    std::vector<int32_t> ids;
    for (int i = 0; i < h.nx; ++i)
        for (int j = 0; j < h.ny; ++j)
            for (int k = 0; k < h.nz; ++k)
                for (particle *p = sr->cp[i][j][k].pl.head; p; p = p->next) {
                    double r[10] = {p->x, p->y, p->z, p->ux, p->uy, p->uz, p->g, p->q, p->cmr, p->chi};
                    rec.insert(rec.end(), r, r + 10);
                    ids.push_back(p->trn);
                }
    int64_t n = (int64_t)ids.size();
    wr(f, &n, 8);
    wr(f, rec.data(), rec.size() * 8);
    wr(f, ids.data(), ids.size() * 4);
}

/* one step in the order of src/main.cpp:2319-2332 (n_sr = 1, nm*(i!=0) = 0) */
static void one_step(spatial_region *sr, const qg_header &h, double *diag)
{
    sr->birth_from_vacuum(h.q_vac);
    sr->padvance(h.external_bz);
    sr->compute_N(0, 0, h.norm_n);
    sr->compute_energy(0, 0, h.norm_ef, h.norm_ep);
    if (diag) {
        double ei = 0;
        for (int n = 0; n < h.n_ion; ++n) ei += sr->ienergy[n];
        double nion = 0;
        for (int n = 0; n < h.n_ion; ++n) nion += sr->N_qp_i[n];
        double d[QG_DIAG_WIDTH] = {sr->N_e, sr->N_p, sr->N_ph, (double)sr->N_qp_e, (double)sr->N_qp_p,
                                   (double)sr->N_qp_g, sr->energy_f, sr->energy_e, sr->energy_p,
                                   sr->energy_ph, ei, nion};
        std::memcpy(diag, d, sizeof d);
    }
    sr->fadvance();
}

static void laser(spatial_region *sr, const qg_header &h)
{
    const double *L = h.laser;
    sr->f_init_cos(L[0], L[1], L[2], L[3], L[4], L[5], 0, L[6] != 0, L[7], L[8], L[9], L[10], L[11] != 0,
                   L[12], L[13], L[14], L[15]);
}

int main(int argc, char **argv)
{
    if (argc != 3) die("usage: quill_ref <in> <out>");
    FILE *fi = std::fopen(argv[1], "rb");
    if (!fi) die("cannot open input");
    qg_header h;
    rd(fi, &h, sizeof h);
    if (h.magic != QG_MAGIC) die("bad magic");
    qed_enabled = h.qed != 0;
    const size_t nf = (size_t)h.nx * h.ny * h.nz * 3;

    FILE *fo = std::fopen(argv[2], "wb");
    if (!fo) die("cannot open output");
    wr(fo, &h, sizeof h);

    if (h.mode == QG_MODE_PUSH) {
        /* rows: cmr dt ex ey ez bx by bz ux uy uz g -> boris (ux uy uz g) vay (ux uy uz g) */
        std::vector<double> t((size_t)h.n_particles * 12);
        rd(fi, t.data(), t.size() * 8);
        for (int r = 0; r < h.n_particles; ++r) {
            const double *a = &t[(size_t)r * 12];
            double out[8];
            for (int which = 0; which < 2; ++which) {
                particle p;
                p.cmr = a[0];
                p.ux = a[8]; p.uy = a[9]; p.uz = a[10]; p.g = a[11];
                vector3d e, b;
                e.x = a[2]; e.y = a[3]; e.z = a[4];
                b.x = a[5]; b.y = a[6]; b.z = a[7];
                if (which == 0) push_boris(p, e, b, a[1]);
                else push_vay(p, e, b, a[1]);
                out[4 * which + 0] = p.ux; out[4 * which + 1] = p.uy;
                out[4 * which + 2] = p.uz; out[4 * which + 3] = p.g;
            }
            wr(fo, out, sizeof out);
        }
    } else if (h.mode == QG_MODE_RATES) {
        /* rows: g r chi ex ey ez bx by bz ux uy uz ->
         *       chi(e,b,u,g) w r_after_w tilde_w mathcal_W */
        std::vector<double> icmr;
        spatial_region *sr = make_region(h, icmr);
        std::vector<double> t((size_t)h.n_particles * 12);
        rd(fi, t.data(), t.size() * 8);
        for (int r = 0; r < h.n_particles; ++r) {
            const double *a = &t[(size_t)r * 12];
            vector3d e, b, u;
            e.x = a[3]; e.y = a[4]; e.z = a[5];
            b.x = a[6]; b.y = a[7]; b.z = a[8];
            u.x = a[9]; u.y = a[10]; u.z = a[11];
            double g = a[0], rr = a[1], ch = a[2];
            double out[5];
            out[0] = sr->chi(e, b, u, g);
            out[1] = sr->w(g, rr, ch);
            out[2] = rr;
            double r2 = a[1];
            out[3] = sr->tilde_w(g, r2, ch);
            out[4] = sr->mathcal_W(e, b);
            wr(fo, out, sizeof out);
        }
        delete sr;
    } else if (h.mode == QG_MODE_LASER) {
        std::vector<double> icmr;
        spatial_region *sr = make_region(h, icmr);
        laser(sr, h);
        dump_fields(sr, h, fo, false);
        delete sr;
    } else if (h.mode == QG_MODE_RUN) {
        std::vector<double> icmr;
        spatial_region *sr = make_region(h, icmr);
        std::vector<double> E(nf), B(nf);
        rd(fi, E.data(), nf * 8);
        rd(fi, B.data(), nf * 8);
        put_fields(sr, h, E.data(), B.data());
        std::vector<double> pr((size_t)h.n_particles * 8);
        std::vector<int32_t> ids((size_t)h.n_particles);
        rd(fi, pr.data(), pr.size() * 8);
        rd(fi, ids.data(), ids.size() * 4);
        for (int n = 0; n < h.n_particles; ++n) add_particle(sr, &pr[(size_t)n * 8], ids[n]);
        set_start(sr, h);
        sr->f_init_boundaries();   /* src/main.cpp:2288-2289 */
        sr->interpolate_be();
        std::vector<double> diag((size_t)h.nsteps * QG_DIAG_WIDTH);
        for (int l = 0; l < h.nsteps; ++l) one_step(sr, h, &diag[(size_t)l * QG_DIAG_WIDTH]);
        dump_fields(sr, h, fo, true);
        dump_particles(sr, h, fo);
        wr(fo, diag.data(), diag.size() * 8);
        delete sr;
    } else if (h.mode == QG_MODE_TIME) {
        /* synthetic uniform plasma (oracle/philox.h) + the reference's own laser IC,
         * then warm-up and timed steps of the full reference step */
        std::vector<double> icmr;
        spatial_region *sr = make_region(h, icmr);
        int ppc = (int)h.extra[0];
        double dens = h.extra[1], sig = h.extra[2];
        int warm = (int)h.extra[3], ilo = (int)h.extra[4], ihi = (int)h.extra[5];
        if (ilo < 1) ilo = 1;
        if (ihi > h.nx - 2) ihi = h.nx - 2;
        int64_t np = 0;
        for (int i = ilo; i < ihi; ++i)
            for (int j = 1; j < h.ny - 2; ++j)
                for (int k = 1; k < h.nz - 2; ++k) {
                    uint64_t cell = ((uint64_t)i * h.ny + j) * h.nz + k;
                    for (int s = 0; s < ppc; ++s) {
                        double off[3], u[3];
                        uint64_t gid = cell * (uint64_t)ppc + s;
                        oq_synth_particle((uint64_t)(uint32_t)h.seed, gid, sig, off, u);
                        double r[8] = {i + off[0], j + off[1], k + off[2], u[0], u[1], u[2], -dens / ppc, -1};
                        add_particle(sr, r, 0);
                        ++np;
                    }
                }
        set_start(sr, h);
        laser(sr, h);
        sr->f_init_boundaries();
        sr->interpolate_be();
        int64_t live = np;   /* electrons alive before the next step (compute_N counts them after each) */
        for (int l = 0; l < warm; ++l) {
            one_step(sr, h, nullptr);
            live = sr->N_qp_e;
        }
        /* particles pushed by the timed steps: those alive before each */
        int64_t pushed = 0;
        auto t0 = std::chrono::steady_clock::now();
        for (int l = 0; l < h.nsteps; ++l) {
            pushed += live;
            one_step(sr, h, nullptr);
            live = sr->N_qp_e;
        }
        auto t1 = std::chrono::steady_clock::now();
        double sec = std::chrono::duration<double>(t1 - t0).count();
        std::printf("{\"seconds\": %.6f, \"steps\": %d, \"particles\": %lld, \"pushed\": %lld, \"cells\": %lld}\n",
                    sec, h.nsteps, (long long)np, (long long)pushed, (long long)h.nx * h.ny * h.nz);
        double res[2] = {sec, (double)np};
        wr(fo, res, sizeof res);
        delete sr;
    } else if (h.mode == QG_MODE_FOUT) {
        /* the plane files of write_fields / write_density (src/main.cpp:369-476,
         * 745-995) for n_sr = 1, written by the reference's own spatial_region
         * writers next to the output file (<out>.ex ...); the output file gets
         * E, B, rho (cj after compute_rho) and Be */
        std::vector<double> icmr;
        spatial_region *sr = make_region(h, icmr);
        std::vector<double> E(nf), B(nf);
        rd(fi, E.data(), nf * 8);
        rd(fi, B.data(), nf * 8);
        put_fields(sr, h, E.data(), B.data());
        std::vector<double> pr((size_t)h.n_particles * 8);
        std::vector<int32_t> ids((size_t)h.n_particles);
        rd(fi, pr.data(), pr.size() * 8);
        rd(fi, ids.data(), ids.size() * 4);
        for (int n = 0; n < h.n_particles; ++n) add_particle(sr, &pr[(size_t)n * 8], ids[n]);
        set_start(sr, h);   /* compute_rho walks the lists from pl.start (src/fout.cpp:29) */
        sr->interpolate_be();
        const int iyz = (int)h.extra[0];
        const std::ios_base::openmode mode =
            h.extra[1] != 0 ? (std::ios_base::out | std::ios_base::binary) : std::ios_base::out;
        const std::string base = argv[2];
        typedef void (spatial_region::*plane_fn)(std::ofstream *, int, int, std::ios_base::openmode);
        typedef void (spatial_region::*yz_fn)(std::ofstream *, int, std::ios_base::openmode);
        const char *names[6] = {"ex", "ey", "ez", "bx", "by", "bz"};
        const plane_fn planes[6] = {&spatial_region::fout_ex, &spatial_region::fout_ey, &spatial_region::fout_ez,
                                    &spatial_region::fout_bx, &spatial_region::fout_by, &spatial_region::fout_bz};
        const yz_fn yz[6] = {&spatial_region::fout_ex_yzplane, &spatial_region::fout_ey_yzplane,
                             &spatial_region::fout_ez_yzplane, &spatial_region::fout_bx_yzplane,
                             &spatial_region::fout_by_yzplane, &spatial_region::fout_bz_yzplane};
        for (int c = 0; c < 6; ++c) {
            const std::string name = base + "." + names[c];
            {
                std::ofstream f(name.c_str(), mode);
                (sr->*planes[c])(&f, 0, h.nx, mode);
            }
            std::ofstream f(name.c_str(), mode | std::ios_base::app);
            (sr->*yz[c])(&f, iyz, mode | std::ios_base::app);
        }
        for (int inv = 0; inv < 2; ++inv) {
            const std::string name = base + (inv ? ".inv" : ".w");
            {
                std::ofstream f(name.c_str(), std::ios_base::out);
                if (inv) sr->fout_inv(&f, 0, h.nx, true);
                else sr->fout_w(&f, 0, h.nx, true);
            }
            std::ofstream f(name.c_str(), std::ios_base::out | std::ios_base::app);
            if (inv) sr->fout_inv_yzplane(&f, iyz);
            else sr->fout_w_yzplane(&f, iyz);
        }
        sr->compute_rho();
        {
            std::ofstream fx((base + ".rho").c_str(), mode), fy((base + ".rho_p").c_str(), mode),
                fz((base + ".rho_ph").c_str(), mode);
            sr->fout_rho(&fx, &fy, &fz, 0, h.nx, mode);
        }
        {
            const std::ios_base::openmode am = mode | std::ios_base::app;
            std::ofstream fx((base + ".rho").c_str(), am), fy((base + ".rho_p").c_str(), am),
                fz((base + ".rho_ph").c_str(), am);
            sr->fout_rho_yzplane(&fx, &fy, &fz, iyz, am);
        }
        dump_fields(sr, h, fo, true);
        delete sr;
    } else if (h.mode == QG_MODE_MERGE) {
        /* one call of pmerging (src/mc.cpp:202-462) on the given particles with
         * ppd = laser[0 .. 3 + n_ion) and the "ti" (extra[0] = 1) or "nl" (2)
         * algorithm, then the surviving particles */
        std::vector<double> icmr;
        spatial_region *sr = make_region(h, icmr);
        std::vector<double> pr((size_t)h.n_particles * 8);
        std::vector<int32_t> ids((size_t)h.n_particles);
        rd(fi, pr.data(), pr.size() * 8);
        rd(fi, ids.data(), ids.size() * 4);
        for (int n = 0; n < h.n_particles; ++n) add_particle(sr, &pr[(size_t)n * 8], ids[n]);
        set_start(sr, h);
        std::vector<double> ppd(h.laser, h.laser + 3 + (h.n_ion > 0 ? h.n_ion : 0));
        sr->pmerging(ppd.data(), h.extra[0] == 1 ? "ti" : "nl");
        dump_particles(sr, h, fo);
        delete sr;
    } else if (h.mode == QG_MODE_MW) {
        /* RUN with a moving window, in the order of main()'s loop
         * (src/main.cpp:2333-2336, 2396-2398): after each step
         * check_moving_window (:1953-1978, "on" with mwspeed or "auto" with
         * mwtolerance), moving_window(), then while (l + 1) dt < t_add_mw a
         * column of fill_cell_by_particles at i = nx - 3 (:1750-1786, uniform
         * density, no profiles).  New particles get ids base + f ny nz npt +
         * (j nz + k) npt + creation index for the f-th fill.  After the
         * diagnostics: per step nmw and the auto-mode ratio. */
        std::vector<double> icmr;
        spatial_region *sr = make_region(h, icmr);
        std::vector<double> E(nf), B(nf);
        rd(fi, E.data(), nf * 8);
        rd(fi, B.data(), nf * 8);
        put_fields(sr, h, E.data(), B.data());
        std::vector<double> pr((size_t)h.n_particles * 8);
        std::vector<int32_t> ids((size_t)h.n_particles);
        rd(fi, pr.data(), pr.size() * 8);
        rd(fi, ids.data(), ids.size() * 4);
        for (int n = 0; n < h.n_particles; ++n) add_particle(sr, &pr[(size_t)n * 8], ids[n]);
        set_start(sr, h);
        sr->f_init_boundaries();
        sr->interpolate_be();
        const int mw = (int)h.extra[0];
        const double mwspeed = h.extra[1], tol = h.extra[2], dens = h.extra[3], t_add = h.extra[7];
        int_vector3d npic;
        npic.i = (int)h.extra[4];
        npic.j = (int)h.extra[5];
        npic.k = (int)h.extra[6];
        const int npt = npic.i * npic.j * npic.k;
        const int id_base = 1 << 24;
        int nmw = 1, fills = 0;
        std::vector<double> diag((size_t)h.nsteps * QG_DIAG_WIDTH), shifts(2 * (size_t)h.nsteps);
        for (int l = 0; l < h.nsteps; ++l) {
            one_step(sr, h, &diag[(size_t)l * QG_DIAG_WIDTH]);
            bool shift = false;
            double ratio = 0;
            if (mw == 1) {
                shift = (l + 1) * h.dt * mwspeed > nmw * h.dx;
            } else if (mw == 2) {
                double max_w = sr->get_max_w();
                ratio = max_w > 0 ? sr->get_max_w(h.nx - nm - 2, h.nx - nm) / max_w : 0.0;
                shift = ratio > tol;
            }
            if (shift) {
                sr->moving_window();
                ++nmw;
            }
            shifts[2 * (size_t)l] = nmw;
            shifts[2 * (size_t)l + 1] = ratio;
            if (shift && (l + 1) * h.dt < t_add && dens > 0) {
                int_vector3d a;
                a.i = h.nx - 3;
                for (int j = 0; j < h.ny; ++j)
                    for (int k = 0; k < h.nz; ++k) {
                        a.j = j;
                        a.k = k;
                        sr->fill_cell_by_particles(-1, a, npic, dens);
                        if (!sr->is_inside(a.i, j, k)) continue;
                        /* place() put each new particle at the head of the list */
                        particle *p = sr->cp[a.i][j][k].pl.head;
                        for (int e = 0; e < npt && p; ++e, p = p->next)
                            p->trn = id_base + (fills * h.ny * h.nz + j * h.nz + k) * npt + (npt - 1 - e);
                    }
                ++fills;
            }
        }
        dump_fields(sr, h, fo, true);
        dump_particles(sr, h, fo);
        wr(fo, diag.data(), diag.size() * 8);
        wr(fo, shifts.data(), shifts.size() * 8);
        delete sr;
    } else if (h.mode == QG_MODE_FILM) {
        /* one call of film() (src/init.cpp:537-646) on an empty region with
         * the arguments laser[0..15] = x0 x1 ne_y0 ne_y1 ions cmr gradwidth y0
         * y1 z0 z1 T vx xnpic ynpic znpic, extra[0] = append_for_moving_window,
         * extra[1] = gradwidth_y (is_profiled off), then the particles */
        std::vector<double> icmr;
        spatial_region *sr = make_region(h, icmr);
        const double *L = h.laser;
        sr->film(L[0], L[1], L[2], L[3], (int)L[4], L[5], L[6], L[7], L[8], L[9], L[10], L[11], L[12], false,
                 (int)L[13], (int)L[14], (int)L[15], h.extra[0] != 0, h.extra[1]);
        dump_particles(sr, h, fo);
        delete sr;
    } else {
        die("unknown mode");
    }
    std::fclose(fo);
    std::fclose(fi);
    return 0;
}
