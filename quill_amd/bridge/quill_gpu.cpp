// quill_gpu.cpp -- reference-side binding of libquillk.so (see quill_gpu.h).
//
// Reads main()'s globals (src/main.cpp:20-122): the grid steps dx, dy, dz,
// dt, the solver and pusher choices, lambda, external_bz, the ion
// populations, and the subdomain sizes nx_sr[mpi_rank], ny_global,
// nz_global; the same quantities main() passes to spatial_region::init and
// create_arrays (src/main.cpp:2276-2277) and to the step (:2324-2327).
#include "quill_gpu.h"

#include <cmath>
#include <iostream>
#include <string>
#include <vector>

extern double dx, dy, dz, dt;                 // src/main.cpp:21
extern int ny_global, nz_global;              // src/main.cpp:23
extern double lambda;                         // src/main.cpp:28
extern double external_bz;                    // src/main.cpp:55
extern int n_ion_populations;                 // src/main.cpp:57
extern double *icmr;                          // src/main.cpp:58
extern maxwell_solver_enum solver;            // src/main.cpp:96
extern pusher_enum pusher;                    // src/main.cpp:97
extern int mpi_rank;                          // src/main.cpp:114
extern vector<int> nx_sr;                     // src/main.cpp:115
#ifdef QUILL_NOQED
extern const bool qed_enabled;
#else
extern bool qed_enabled;                      // src/main.cpp:84-88
#endif

namespace {

const double PI = 3.141592653589793;

// species slot of a charge-to-mass ratio (src/plist.cpp:31, init.cpp:413)
int species_of(double cmr)
{
    if (cmr == -1) return QK_ELECTRON;
    if (cmr == 1) return QK_POSITRON;
    if (cmr == 0) return QK_PHOTON;
    for (int n = 0; n < n_ion_populations && n < QK_MAX_ION; ++n)
        if (cmr == icmr[n]) return QK_ION0 + n;
    return -1;
}

void copy_fields_out(spatial_region &sr, int nx, int ny, int nz, std::vector<double> &E, std::vector<double> &B)
{
    E.resize(3 * (size_t)nx * ny * nz);
    B.resize(E.size());
    size_t o = 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k, o += 3) {
                E[o] = sr.ce[i][j][k].ex; E[o + 1] = sr.ce[i][j][k].ey; E[o + 2] = sr.ce[i][j][k].ez;
                B[o] = sr.cb[i][j][k].bx; B[o + 1] = sr.cb[i][j][k].by; B[o + 2] = sr.cb[i][j][k].bz;
            }
}

}  // namespace

qk_ctx *quill_gpu_create(spatial_region &sr, const quill_gpu_options &opt)
{
    const int nx = nx_sr[mpi_rank], ny = ny_global, nz = nz_global;
    if (sr.get_nx() != nx || sr.get_ny() != ny || sr.get_nz() != nz) {
        std::cerr << "quill_gpu_create: region is not " << nx << " x " << ny << " x " << nz << '\n';
        return nullptr;
    }
    if (n_ion_populations > QK_MAX_ION) {
        std::cerr << "quill_gpu_create: at most " << QK_MAX_ION << " ion populations\n";
        return nullptr;
    }
    qk_params p = {};
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.dx = dx;
    p.dy = dy;
    p.dz = dz;
    p.dt = dt;
    p.solver = solver == maxwell_solver_enum::FDTD ? QK_SOLVER_FDTD
               : solver == maxwell_solver_enum::NDFX ? QK_SOLVER_NDFX : QK_SOLVER_FIVE_POINT;
    p.pusher = pusher == pusher_enum::BORIS ? QK_PUSHER_BORIS : QK_PUSHER_VAY;
    p.qed = qed_enabled ? 1 : 0;
    p.precision = opt.precision;
    p.deposit = opt.deposit;
    p.sort_every = opt.sort_every;
    p.sort_fallback = opt.sort_fallback;
    p.device = opt.device;
    p.e_s = lambda / 2.4263086e-10;                                  // src/main.cpp:2276
    p.external_bz = external_bz;                                      // src/main.cpp:2325
    p.q_vac = 8 * PI * PI / (dx * dy * dz) * 2.818e-13 / lambda;      // src/main.cpp:2324
    p.seed = (uint64_t)(mpi_rank + 1);
    p.n_ion = n_ion_populations;
    for (int n = 0; n < n_ion_populations; ++n) p.ion_cmr[n] = icmr[n];
    // compute_N / compute_energy multipliers (src/main.cpp:2326-2327)
    p.norm_n = dx * dy * dz * 1.11485e13 * lambda / (8 * PI * PI * PI);
    p.norm_ef = 0.5 * dx * dy * dz * 3.691e4 * lambda / 1e7;
    p.norm_ep = 8.2e-14 * dx * dy * dz * 1.11485e13 * lambda / (8 * PI * PI * PI);
    qk_ctx *ctx = nullptr;
    if (qk_create(&p, &ctx) != QK_OK) return nullptr;

    std::vector<double> E, B;
    copy_fields_out(sr, nx, ny, nz, E, B);
    if (qk_set_fields(ctx, E.data(), B.data()) != QK_OK) {
        qk_destroy(ctx);
        return nullptr;
    }
    // every particle of the cell lists (cp[i][j][k].pl, include/main.h:72;
    // linked through next, src/plist.cpp:22-40), one SoA per species
    struct Soa {
        std::vector<double> x, y, z, ux, uy, uz, q;
        std::vector<int32_t> id;
    };
    std::vector<Soa> sp(QK_MAX_SPECIES);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                for (particle *a = sr.cp[i][j][k].pl.head; a; a = a->next) {
                    const int s = species_of(a->cmr);
                    if (s < 0) {
                        std::cerr << "quill_gpu_create: particle of unknown cmr " << a->cmr << '\n';
                        qk_destroy(ctx);
                        return nullptr;
                    }
                    Soa &d = sp[s];
                    d.x.push_back(a->x); d.y.push_back(a->y); d.z.push_back(a->z);
                    d.ux.push_back(a->ux); d.uy.push_back(a->uy); d.uz.push_back(a->uz);
                    d.q.push_back(a->q);
                    d.id.push_back(a->trn);
                }
    for (int s = 0; s < QK_MAX_SPECIES; ++s) {
        Soa &d = sp[s];
        if (d.x.empty()) continue;
        if (qk_add_particles(ctx, s, (int64_t)d.x.size(), d.x.data(), d.y.data(), d.z.data(), d.ux.data(),
                             d.uy.data(), d.uz.data(), d.q.data(), d.id.data(), nullptr) != QK_OK) {
            qk_destroy(ctx);
            return nullptr;
        }
    }
    if (qk_init_boundaries(ctx) != QK_OK) {   // f_init_boundaries + interpolate_be
        qk_destroy(ctx);
        return nullptr;
    }
    return ctx;
}

int quill_gpu_step(qk_ctx *ctx, spatial_region &sr)
{
    qk_diag d;
    const int rc = qk_step(ctx, 1, &d);
    if (rc != QK_OK) return rc;
    sr.N_e = d.N_e;
    sr.N_p = d.N_p;
    sr.N_ph = d.N_ph;
    sr.N_qp_e = (int)d.n_e;
    sr.N_qp_p = (int)d.n_p;
    sr.N_qp_g = (int)d.n_ph;
    sr.energy_f = d.energy_f;
    sr.energy_e = d.energy_e;
    sr.energy_p = d.energy_p;
    sr.energy_ph = d.energy_ph;
    for (int n = 0; n < n_ion_populations && n < QK_MAX_ION; ++n) sr.ienergy[n] = d.energy_ions[n];
    return QK_OK;
}

int quill_gpu_fields_to_region(qk_ctx *ctx, spatial_region &sr)
{
    const int nx = sr.get_nx(), ny = sr.get_ny(), nz = sr.get_nz();
    const size_t n = 3 * (size_t)nx * ny * nz;
    std::vector<double> E(n), B(n), J(n), Be(n);
    const int rc = qk_get_fields(ctx, E.data(), B.data(), J.data(), Be.data());
    if (rc != QK_OK) return rc;
    size_t o = 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k, o += 3) {
                sr.ce[i][j][k].ex = E[o]; sr.ce[i][j][k].ey = E[o + 1]; sr.ce[i][j][k].ez = E[o + 2];
                sr.cb[i][j][k].bx = B[o]; sr.cb[i][j][k].by = B[o + 1]; sr.cb[i][j][k].bz = B[o + 2];
                sr.cj[i][j][k].jx = J[o]; sr.cj[i][j][k].jy = J[o + 1]; sr.cj[i][j][k].jz = J[o + 2];
                sr.cbe[i][j][k].bex = Be[o]; sr.cbe[i][j][k].bey = Be[o + 1]; sr.cbe[i][j][k].bez = Be[o + 2];
            }
    return QK_OK;
}
