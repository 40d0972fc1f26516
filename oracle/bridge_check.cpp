/*
 * oracle/bridge_check.cpp -- end-to-end check of the reference-side binding
 * (quill_amd/bridge/quill_gpu.cpp).
 *
 * TEST INFRASTRUCTURE.  Linked by oracle/Makefile with the reference's own
 * hot-path translation units and libquillk.so.  It stands in for the part of
 * src/main.cpp the binding reads: it defines main()'s globals (the ones the
 * reference units and the binding declare extern), sets a region up with the
 * reference's own calls (init, create_arrays, f_init_cos, new_particle +
 * place, f_init_boundaries, interpolate_be; src/main.cpp:2275-2289), hands
 * it to the GPU through quill_gpu_create, and runs the loop body of
 * src/main.cpp:2319-2332 on both: the reference's own step on the CPU and
 * quill_gpu_step in FP64 on the GPU.  Prints one JSON line with the largest
 * relative field and diagnostic differences.
 *
 * Usage: bridge_check <solver 0|1|2> <pusher 0|1> <steps>
 * Exit: 0 agreement to 1e-10, 1 disagreement, 3 no GPU context (the error
 * text on stderr).
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "main.h"
#include "philox.h"
#include "quill_gpu.h"

/* main()'s globals (src/main.cpp:20-122) that the reference units
 * (src/advance.cpp:9-11, src/low-level.cpp:12-14, src/plist.cpp:4-5) and the
 * binding read */
double dx, dy, dz, dt;
int ny_global, nz_global;
double lambda = 1e-4;
double external_bz = 0.1;
int n_ion_populations = 1;
double *icmr = nullptr;
maxwell_solver_enum solver;
pusher_enum pusher;
int mpi_rank = 0;
int n_sr = 1;
int nm = 4;
vector<int> nx_sr;
bool qed_enabled = false;
bool dump_photons = false;
bool catching_enabled = false;

static spatial_region *make_region(int nx)
{
    spatial_region *sr = new spatial_region;
    sr->init(mpi_rank, dx, dy, dz, dt, lambda / 2.4263086e-10, 1, 1, 1, n_ion_populations, icmr, "/tmp", solver,
             pusher);
    sr->create_arrays(nx, ny_global, nz_global, 7);
    return sr;
}

int main(int argc, char **argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: bridge_check <solver 0|1|2> <pusher 0|1> <steps>\n");
        return 2;
    }
    const int s = std::atoi(argv[1]), pu = std::atoi(argv[2]), nsteps = std::atoi(argv[3]);
    solver = s == 0 ? maxwell_solver_enum::FDTD : (s == 1 ? maxwell_solver_enum::NDFX : maxwell_solver_enum::FIVE_POINT);
    pusher = pu == 1 ? pusher_enum::BORIS : pusher_enum::VAY;
    const double PI = 3.141592653589793;
    const int nx = 28;
    nx_sr.assign(1, nx);
    ny_global = 20;
    nz_global = 18;
    dx = 0.05 * 2 * PI;
    dy = 1.1 * dx;
    dz = dx;
    dt = (s == 2 ? 0.35 : 0.5) * dx;
    static double ion_cmr[1] = {1.0 / 1836.1526721};
    icmr = ion_cmr;

    spatial_region *sr = make_region(nx);
    /* a tilted, elliptically polarised pulse (f_init_cos, src/init.cpp:6-210) */
    sr->f_init_cos(0.8, 0.3, 0.6 * PI, 0.8 * PI, 0.8 * PI, 0.5 * nx * dx, 0, true, -1.0, 0.2, 0.1, 0.0, true, 0.1,
                   0, 0, 0);
    /* electrons, positrons and ions from the shared Philox loader */
    const double cmrs[3] = {-1.0, 1.0, ion_cmr[0]};
    int id = 0;
    for (int i = 1; i < nx - 2; ++i)
        for (int j = 1; j < ny_global - 2; ++j)
            for (int k = 1; k < nz_global - 2; ++k)
                for (int m = 0; m < 3; ++m) {
                    double off[3], u[3];
                    const uint64_t cell = ((uint64_t)i * ny_global + j) * nz_global + k;
                    oq_synth_particle(11, cell * 3 + m, 0.3, off, u);
                    particle *p = sr->new_particle();
                    p->x = i + off[0]; p->y = j + off[1]; p->z = k + off[2];
                    p->ux = u[0]; p->uy = u[1]; p->uz = u[2];
                    p->q = m == 0 ? -0.02 : 0.02;
                    p->cmr = cmrs[m];
                    p->g = std::sqrt(1 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                    p->chi = 0;
                    p->trn = id++;
                    p->next = p->previous = nullptr;
                    sr->place(*p);
                }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny_global; ++j)
            for (int k = 0; k < nz_global; ++k) sr->cp[i][j][k].pl.start = sr->cp[i][j][k].pl.head;
    sr->f_init_boundaries();
    sr->interpolate_be();

    quill_gpu_options opt;
    opt.precision = QK_FP64;
    opt.sort_every = 5;
    qk_ctx *ctx = quill_gpu_create(*sr, opt);
    if (!ctx) {
        std::fprintf(stderr, "bridge_check: no GPU context: %s\n", qk_last_error());
        return 3;
    }
    /* the GPU's diagnostics and fields land in a second region */
    spatial_region *sg = make_region(nx);
    const double norm_n = dx * dy * dz * 1.11485e13 * lambda / (8 * PI * PI * PI);
    const double norm_ef = 0.5 * dx * dy * dz * 3.691e4 * lambda / 1e7;
    const double norm_ep = 8.2e-14 * dx * dy * dz * 1.11485e13 * lambda / (8 * PI * PI * PI);
    double diag_err = 0;
    int rc = 0;
    for (int l = 0; l < nsteps && rc == 0; ++l) {
        sr->padvance(external_bz);
        sr->compute_N(0, 0, norm_n);
        sr->compute_energy(0, 0, norm_ef, norm_ep);
        sr->fadvance();
        rc = quill_gpu_step(ctx, *sg);
        const double a[6] = {sr->N_e, sr->N_p, sr->energy_f, sr->energy_e, sr->energy_p, sr->ienergy[0]};
        const double b[6] = {sg->N_e, sg->N_p, sg->energy_f, sg->energy_e, sg->energy_p, sg->ienergy[0]};
        for (int m = 0; m < 6; ++m)
            if (a[m] != 0) diag_err = std::fmax(diag_err, std::fabs(a[m] - b[m]) / std::fabs(a[m]));
        if (sr->N_qp_e != sg->N_qp_e || sr->N_qp_p != sg->N_qp_p) diag_err = 1;
    }
    if (rc == 0) rc = quill_gpu_fields_to_region(ctx, *sg);
    if (rc != 0) {
        std::fprintf(stderr, "bridge_check: %s\n", qk_last_error());
        return 1;
    }
    double err[4] = {0, 0, 0, 0}, scale[4] = {0, 0, 0, 0};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny_global; ++j)
            for (int k = 0; k < nz_global; ++k) {
                const double r[4][3] = {{sr->ce[i][j][k].ex, sr->ce[i][j][k].ey, sr->ce[i][j][k].ez},
                                        {sr->cb[i][j][k].bx, sr->cb[i][j][k].by, sr->cb[i][j][k].bz},
                                        {sr->cj[i][j][k].jx, sr->cj[i][j][k].jy, sr->cj[i][j][k].jz},
                                        {sr->cbe[i][j][k].bex, sr->cbe[i][j][k].bey, sr->cbe[i][j][k].bez}};
                const double g[4][3] = {{sg->ce[i][j][k].ex, sg->ce[i][j][k].ey, sg->ce[i][j][k].ez},
                                        {sg->cb[i][j][k].bx, sg->cb[i][j][k].by, sg->cb[i][j][k].bz},
                                        {sg->cj[i][j][k].jx, sg->cj[i][j][k].jy, sg->cj[i][j][k].jz},
                                        {sg->cbe[i][j][k].bex, sg->cbe[i][j][k].bey, sg->cbe[i][j][k].bez}};
                for (int f = 0; f < 4; ++f)
                    for (int c = 0; c < 3; ++c) {
                        err[f] = std::fmax(err[f], std::fabs(r[f][c] - g[f][c]));
                        scale[f] = std::fmax(scale[f], std::fabs(r[f][c]));
                    }
            }
    double worst = diag_err;
    for (int f = 0; f < 4; ++f) {
        err[f] /= scale[f] > 0 ? scale[f] : 1;
        worst = std::fmax(worst, err[f]);
    }
    std::printf("{\"E\": %.3e, \"B\": %.3e, \"J\": %.3e, \"Be\": %.3e, \"diagnostics\": %.3e, \"steps\": %d, "
                "\"electrons\": %d}\n", err[0], err[1], err[2], err[3], diag_err, nsteps, sr->N_qp_e);
    qk_destroy(ctx);
    delete sg;
    delete sr;
    return worst < 1e-10 ? 0 : 1;
}
