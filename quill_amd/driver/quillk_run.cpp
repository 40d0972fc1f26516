// quillk_run.cpp -- command-line driver: QUILL's main() for one subdomain
// (src/main.cpp:2180-2340 at n_sr = 1) on libquillk.so.  Reads a .conf
// file, sets up fields and particles as init(), init_fields() and
// init_films() do, runs the step loop and writes the per-step N and energy
// records in the reference's text format (write_N / write_energy,
// src/main.cpp:160-187).
//
//   quillk-run CONF [--steps N] [--fp64] [--deposit tiled|global]
//              [--sort-every N] [--sort-fallback F] [--device D] [--out DIR] [--dry-run]
//
// Output iterations write the reference's plane files (ex .. bz, w, inv,
// rho, rho_p, rho_ph, jx .. jz as the config asks) unless --no-output.
// --dry-run prints the derived run as JSON and touches no GPU; --write-log
// DIR writes only the log file (src/main.cpp:3611-3787) and touches no GPU;
// --print-conf prints the parsed variables as JSON; --dump-setup DIR writes
// the initial particles of every species (raw little-endian doubles, rows
// of x y z ux uy uz q, file species<slot>.f64) and the laser calls
// (lasers.f64, rows of 16) for an independent run of the same set-up;
// --mw-fills NMW prints the moving window's column fills after the shift
// that made nmw (species, cells, densities) as JSON.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "qk_conf.h"
#include "qk_output.h"
#include "qk_setup.h"
#include "quillk.h"

using namespace qkd;

namespace {

constexpr double kPi = 3.14159265358979323846;   // include/physical_constants.h:3

int usage()
{
    std::fprintf(stderr,
                 "usage: quillk-run CONF [--steps N] [--fp64] [--deposit tiled|global] [--sort-every N]\n"
                 "                  [--sort-fallback F] [--device D] [--out DIR] [--no-output] [--dry-run]\n"
                 "                  [--print-conf]\n"
                 "                  [--dump-setup DIR] [--mw-fills NMW]\n"
                 "       quillk-run --write-planes IN DIR\n");
    return 2;
}

int check(int rc, const char *what)
{
    if (rc != QK_OK) std::fprintf(stderr, "quillk-run: %s failed (%d): %s\n", what, rc, qk_last_error());
    return rc;
}

const char *solver_name(int s) { return s == QK_SOLVER_FDTD ? "fdtd" : (s == QK_SOLVER_NDFX ? "ndfx" : "five_point"); }

// particles of every film, per species slot
struct Species {
    int slot;
    double cmr;
    ParticleSet set;
};

std::vector<Species> make_particles(const Run &run)
{
    std::vector<Species> out;
    out.push_back({QK_ELECTRON, -1, {}});
    out.push_back({QK_POSITRON, 1, {}});
    for (size_t n = 0; n < run.ion_mcr.size(); ++n) out.push_back({QK_ION0 + (int)n, run.params.ion_cmr[n], {}});
    for (const Film &f : run.films) {
        film_particles(run, f, -1, out[0].set);
        if (f.ions == 2) film_particles(run, f, 1, out[1].set);
        if (f.ions == 1)
            for (size_t n = 2; n < out.size(); ++n)
                if (out[n].cmr == f.cmr) film_particles(run, f, f.cmr, out[n].set);
    }
    return out;
}

void print_json(const Run &run, const std::vector<Species> &sp, int steps)
{
    const qk_params &p = run.params;
    std::printf("{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"dx\": %.17g, \"dy\": %.17g, \"dz\": %.17g, \"dt\": %.17g,\n",
                p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, p.dt);
    std::printf(" \"steps\": %d, \"solver\": \"%s\", \"pusher\": \"%s\", \"qed\": %d, \"lambda_cm\": %.17g,\n", steps,
                solver_name(p.solver), p.pusher == QK_PUSHER_VAY ? "vay" : "boris", p.qed, run.lambda_cm);
    std::printf(" \"e_s\": %.17g, \"q_vac\": %.17g, \"external_bz\": %.17g,\n", p.e_s, p.q_vac, p.external_bz);
    std::printf(" \"norm_n\": %.17g, \"norm_ef\": %.17g, \"norm_ep\": %.17g,\n", p.norm_n, p.norm_ef, p.norm_ep);
    std::printf(" \"merging\": %d, \"crnp\": %.17g, \"outputs\": %zu,\n", p.merging, p.crnp, run.out_iter.size());
    std::printf(" \"mwindow\": %d, \"mwspeed\": %.17g, \"mwtolerance\": %.17g, \"t_add_mw\": %.17g, \"mw_fills\": [",
                run.mwindow, run.mwspeed, run.mwtolerance, run.t_add_mw);
    // the column fills after the first shift (nmw = 2): species, cells, sum of densities
    const std::vector<CellFill> fills = moving_window_fills(run, 2);
    for (size_t n = 0; n < fills.size(); ++n) {
        double sum = 0;
        for (double d : fills[n].density) sum += d;
        std::printf("%s[%d, %zu, %.17g]", n ? ", " : "", fills[n].species, fills[n].density.size(), sum);
    }
    std::printf("],\n");
    std::printf(" \"ion_cmr\": [");
    for (int n = 0; n < p.n_ion; ++n) std::printf("%s%.17g", n ? ", " : "", p.ion_cmr[n]);
    std::printf("],\n \"lasers\": [");
    for (size_t n = 0; n < run.lasers.size(); ++n) {
        std::printf("%s[", n ? ", " : "");
        for (int m = 0; m < 16; ++m) std::printf("%s%.17g", m ? ", " : "", run.lasers[n].L[m]);
        std::printf("]");
    }
    // film() arguments of the set-up, newest first (src/init.cpp:537-541):
    // x0 x1 ne_y0 ne_y1 ions cmr gradwidth y0 y1 z0 z1 T vx xnpic ynpic znpic gradwidth_y
    std::printf("],\n \"film_args\": [");
    for (size_t n = 0; n < run.films.size(); ++n) {
        const Film &f = run.films[n];
        const double a[17] = {f.x0, f.x1, f.ne_y0, f.ne_y1, (double)f.ions, f.cmr, f.gradwidth, f.y0, f.y1, f.z0,
                              f.z1, f.T, f.vx, (double)f.npic[0], (double)f.npic[1], (double)f.npic[2], f.gradwidth_y};
        std::printf("%s[", n ? ", " : "");
        for (int m = 0; m < 17; ++m) std::printf("%s%.17g", m ? ", " : "", a[m]);
        std::printf("]");
    }
    std::printf("],\n \"films\": %zu, \"particles\": {", run.films.size());
    for (size_t n = 0; n < sp.size(); ++n) {
        double qsum = 0;
        for (double q : sp[n].set.q) qsum += q;
        std::printf("%s\"%d\": {\"n\": %zu, \"q_sum\": %.17g}", n ? ", " : "", sp[n].slot, sp[n].set.q.size(), qsum);
    }
    std::printf("}}\n");
}

// --mw-fills NMW: the column fills of add_moving_window_particles after the
// shift that made nmw, as JSON
void print_fills(const Run &run, int nmw)
{
    const std::vector<CellFill> fills = moving_window_fills(run, nmw);
    std::printf("[");
    for (size_t n = 0; n < fills.size(); ++n) {
        const CellFill &f = fills[n];
        std::printf("%s\n {\"species\": %d, \"npic\": [%d, %d, %d], \"ux0\": %.17g, \"cells\": [", n ? "," : "",
                    f.species, f.npic[0], f.npic[1], f.npic[2], f.ux0);
        for (size_t m = 0; m < f.density.size(); ++m)
            std::printf("%s[%d, %d, %d]", m ? ", " : "", f.ijk[3 * m], f.ijk[3 * m + 1], f.ijk[3 * m + 2]);
        std::printf("], \"density\": [");
        for (size_t m = 0; m < f.density.size(); ++m) std::printf("%s%.17g", m ? ", " : "", f.density[m]);
        std::printf("]}");
    }
    std::printf("\n]\n");
}

void print_conf(const std::vector<ConfVar> &vars)
{
    std::printf("[");
    for (size_t n = 0; n < vars.size(); ++n) {
        const ConfVar &v = vars[n];
        std::printf("%s\n {\"name\": \"%s\", \"value\": %.17g, \"units\": \"%s\", \"array\": [", n ? "," : "",
                    v.name.c_str(), v.value, v.units.c_str());
        for (size_t m = 0; m < v.array.size(); ++m) std::printf("%s%.17g", m ? ", " : "", v.array[m]);
        std::printf("]}");
    }
    std::printf("\n]\n");
}

bool dump_setup(const std::string &dir, const Run &run, const std::vector<Species> &sp)
{
    mkdir(dir.c_str(), 0755);
    for (const Species &s : sp) {
        std::ofstream f(dir + "/species" + std::to_string(s.slot) + ".f64", std::ios::binary);
        for (size_t n = 0; n < s.set.q.size(); ++n) {
            const double r[7] = {s.set.x[n], s.set.y[n], s.set.z[n], s.set.ux[n], s.set.uy[n], s.set.uz[n], s.set.q[n]};
            f.write((const char *)r, sizeof r);
        }
        if (!f) return false;
    }
    std::ofstream f(dir + "/lasers.f64", std::ios::binary);
    for (const LaserCall &l : run.lasers) f.write((const char *)l.L, sizeof l.L);
    return (bool)f;
}

// --write-planes IN DIR: the writers on given arrays (a check of the file
// formats against the reference's own writers).  IN holds int32 nx ny nz
// i_yz binary, then E, B, Be and rho as double[nx][ny][nz][3].
int write_planes(const char *in, const char *dir)
{
    std::ifstream f(in, std::ios::binary);
    int32_t h[5];
    if (!f.read((char *)h, sizeof h)) return 1;
    OutputGrid g{h[0], h[1], h[2], h[3], h[4] != 0};
    const size_t n = 3 * (size_t)g.nx * g.ny * g.nz;
    std::vector<double> E(n), B(n), Be(n), rho(n);
    for (std::vector<double> *a : {&E, &B, &Be, &rho})
        if (!f.read((char *)a->data(), (std::streamsize)(n * sizeof(double)))) return 1;
    OutputSet s;
    for (int c = 0; c < 3; ++c) s.e[c] = s.b[c] = true;
    s.rho_p = s.rho_ph = true;
    mkdir(dir, 0755);
    write_output_set(dir, "", g, s, E.data(), B.data(), Be.data(), rho.data(), nullptr, 1.0);
    return 0;
}

// --write-particles IN DIR: the spectrum and phase-space writer on given
// particles (a check of its file formats).  IN holds int32 n_ion, the
// SpectraConf numbers (4 doubles deps.., 8 int32 neps.. enthp.., 2 int32
// write_p write_ph), doubles dx dy dz lambda_cm, n_ion doubles icmr, then per
// species (3 + n_ion) an int64 count and count x 10 doubles (x y z ux uy uz
// g q cmr chi).
int write_particles(const char *in, const char *dir)
{
    std::ifstream f(in, std::ios::binary);
    int32_t n_ion = 0;
    if (!f.read((char *)&n_ion, 4) || n_ion < 0 || n_ion > QK_MAX_ION) return 1;
    SpectraConf c;
    double d4[4];
    int32_t i10[10];
    if (!f.read((char *)d4, sizeof d4) || !f.read((char *)i10, sizeof i10)) return 1;
    c.deps = d4[0]; c.deps_p = d4[1]; c.deps_ph = d4[2]; c.deps_i = d4[3];
    c.neps = i10[0]; c.neps_p = i10[1]; c.neps_ph = i10[2]; c.neps_i = i10[3];
    c.enthp = i10[4]; c.enthp_p = i10[5]; c.enthp_ph = i10[6]; c.enthp_i = i10[7];
    c.write_p = i10[8] != 0;
    c.write_ph = i10[9] != 0;
    double g4[4];
    if (!f.read((char *)g4, sizeof g4)) return 1;
    std::vector<double> icmr(n_ion);
    if (n_ion && !f.read((char *)icmr.data(), (std::streamsize)(n_ion * sizeof(double)))) return 1;
    std::vector<std::vector<double>> recs(3 + n_ion);
    std::vector<SpeciesRecords> sp(3 + n_ion);
    for (int s = 0; s < 3 + n_ion; ++s) {
        int64_t n = 0;
        if (!f.read((char *)&n, 8) || n < 0) return 1;
        recs[s].resize((size_t)n * 10 + 1);
        if (n && !f.read((char *)recs[s].data(), (std::streamsize)(n * 10 * sizeof(double)))) return 1;
        sp[s] = SpeciesRecords{recs[s].data(), n};
    }
    mkdir(dir, 0755);
    PhasespaceCounters k;
    write_spectrum_phasespace(dir, "", c, sp.data(), n_ion, icmr.data(), g4[0], g4[1], g4[2], g4[3], k);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) return usage();
    if (argc == 4 && std::strcmp(argv[1], "--write-planes") == 0) return write_planes(argv[2], argv[3]);
    if (argc == 4 && std::strcmp(argv[1], "--write-particles") == 0) return write_particles(argv[2], argv[3]);
    std::string conf = argv[1], out_dir, setup_dir, log_dir;
    int steps_override = -1, device = 0, sort_every = -1, fills_nmw = 0;
    double sort_fallback = -1;
    bool fp64 = false, dry = false, global_dep = false, show_conf = false, write_outputs = true;
    for (int a = 2; a < argc; ++a) {
        const std::string s = argv[a];
        auto next = [&]() -> const char * { return a + 1 < argc ? argv[++a] : nullptr; };
        if (s == "--steps") { const char *v = next(); if (!v) return usage(); steps_override = std::atoi(v); }
        else if (s == "--fp64") fp64 = true;
        else if (s == "--dry-run") dry = true;
        else if (s == "--print-conf") show_conf = true;
        else if (s == "--no-output") write_outputs = false;
        else if (s == "--dump-setup") { const char *v = next(); if (!v) return usage(); setup_dir = v; }
        else if (s == "--deposit") { const char *v = next(); if (!v) return usage(); global_dep = std::strcmp(v, "global") == 0; }
        else if (s == "--sort-every") { const char *v = next(); if (!v) return usage(); sort_every = std::atoi(v); }
        else if (s == "--sort-fallback") { const char *v = next(); if (!v) return usage(); sort_fallback = std::atof(v); }
        else if (s == "--device") { const char *v = next(); if (!v) return usage(); device = std::atoi(v); }
        else if (s == "--out") { const char *v = next(); if (!v) return usage(); out_dir = v; }
        else if (s == "--mw-fills") { const char *v = next(); if (!v) return usage(); fills_nmw = std::atoi(v); }
        else if (s == "--write-log") { const char *v = next(); if (!v) return usage(); log_dir = v; }
        else return usage();
    }

    std::vector<ConfVar> vars;
    std::string err;
    Run run;
    if (!read_conf_file(conf, vars, err)) {
        std::fprintf(stderr, "quillk-run: %s\n", err.c_str());
        return 1;
    }
    if (show_conf) {
        print_conf(vars);
        return 0;
    }
    if (!build_run(vars, run, err)) {
        std::fprintf(stderr, "quillk-run: %s\n", err.c_str());
        return 1;
    }
    qk_params p = run.params;
    if (fp64) p.precision = QK_FP64;
    if (global_dep) p.deposit = QK_DEPOSIT_GLOBAL;
    if (sort_every >= 0) p.sort_every = sort_every;
    if (sort_fallback >= 0) p.sort_fallback = sort_fallback;
    p.device = device;
    const int steps = steps_override >= 0 ? steps_override : run.last_iteration + 1;
    const std::vector<Species> sp = make_particles(run);
    if (!setup_dir.empty() && !dump_setup(setup_dir, run, sp)) {
        std::fprintf(stderr, "quillk-run: cannot write %s\n", setup_dir.c_str());
        return 1;
    }
    if (dry) {
        print_json(run, sp, steps);
        return 0;
    }
    if (!log_dir.empty()) {   // the log file alone, no GPU
        mkdir(log_dir.c_str(), 0755);
        std::ofstream flog(log_dir + "/log");
        flog << run.log;
        return flog ? 0 : 1;
    }
    if (fills_nmw > 0) {
        print_fills(run, fills_nmw);
        return 0;
    }
    if (out_dir.empty()) out_dir = run.data_folder;
    mkdir(out_dir.c_str(), 0755);

    qk_ctx *ctx = nullptr;
    if (check(qk_create(&p, &ctx), "qk_create")) return 1;
    int rc = 0;
    for (const LaserCall &l : run.lasers)
        if ((rc = check(qk_init_laser_cos(ctx, l.L), "qk_init_laser_cos"))) break;
    for (size_t n = 0; !rc && n < sp.size(); ++n) {
        const ParticleSet &s = sp[n].set;
        if (s.q.empty()) continue;
        rc = check(qk_add_particles(ctx, sp[n].slot, (int64_t)s.q.size(), s.x.data(), s.y.data(), s.z.data(),
                                    s.ux.data(), s.uy.data(), s.uz.data(), s.q.data(), nullptr, nullptr),
                   "qk_add_particles");
    }
    if (!rc) rc = check(qk_init_boundaries(ctx), "qk_init_boundaries");

    if (write_outputs) {
        std::ofstream flog(out_dir + "/log");
        flog << run.log;
    }
    std::ofstream fN(out_dir + "/N"), fE(out_dir + "/energy"), fMW;
    if (run.mwindow) fMW.open(out_dir + "/mwcoordinate");
    const int batch = 100;
    std::vector<qk_diag> diag(batch);
    double seconds = 0;
    int64_t pushed = 0;
    size_t next_out = 0;   // output iterations are written after their step (src/main.cpp:2344-2430)
    const size_t nf = 3 * (size_t)p.nx * p.ny * p.nz;
    int nmw = 1;           // src/main.cpp:122
    PhasespaceCounters ps_counters;
    // check_moving_window (src/main.cpp:1953-1978) after step l; nm = 4
    auto window_moves = [&](int l, bool &moves) -> int {
        moves = false;
        if (run.mwindow == 1) {
            moves = (l + 1) * p.dt * run.mwspeed > nmw * p.dx;
        } else if (run.mwindow == 2) {
            double max_w = 0, front = 0;
            int r = check(qk_max_w(ctx, 0, p.nx, &max_w), "qk_max_w");
            if (!r) r = check(qk_max_w(ctx, p.nx - 4 - 2, p.nx - 4, &front), "qk_max_w");
            if (r) return r;
            const double ratio = max_w > 0 ? front / max_w : 0.0;
            moves = ratio > run.mwtolerance;
        }
        return 0;
    };
    for (int done = 0; !rc && done < steps;) {
        int nb = std::min(batch, steps - done);
        if (write_outputs && next_out < run.out_iter.size()) nb = std::min(nb, run.out_iter[next_out] + 1 - done);
        // the window is checked after every step: "on" batches up to the next
        // shift (known in advance), "auto" steps one at a time
        if (run.mwindow == 2) nb = 1;
        if (run.mwindow == 1)
            for (int l = done; l < done + nb; ++l)
                if ((l + 1) * p.dt * run.mwspeed > nmw * p.dx) {
                    nb = l - done + 1;
                    break;
                }
        const auto t0 = std::chrono::steady_clock::now();
        rc = check(qk_step(ctx, nb, diag.data()), "qk_step");
        seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int l = 0; !rc && l < nb; ++l) {
            const qk_diag &d = diag[l];
            if (d.n_nonfinite > 0)
                std::fprintf(stderr, "quillk-run: step %d dropped %lld particles whose momentum was not finite\n",
                             done + l, (long long)d.n_nonfinite);
            fN << d.N_e << '\t' << d.N_p << '\t' << d.N_ph << std::endl;
            fE << d.energy_f << '\t' << d.energy_e << '\t' << d.energy_p << '\t' << d.energy_ph;
            for (int n = 0; n < p.n_ion; ++n) fE << '\t' << d.energy_ions[n];
            fE << std::endl;
            pushed += d.n_e + d.n_p + d.n_ph + d.n_ion;
        }
        // moving_window after the batch's last step (src/main.cpp:2333-2337);
        // write_mwcoordinate each step (:154-158, 2381-2383)
        bool moves = false;
        const int last = done + nb - 1;
        if (!rc && run.mwindow) {
            rc = window_moves(last, moves);
            if (!rc && moves) {
                rc = check(qk_moving_window(ctx), "qk_moving_window");
                ++nmw;
            }
            for (int l = 0; l < nb - 1; ++l) fMW << ((nmw - (moves ? 2 : 1)) * p.dx / 2 / kPi) << std::endl;
            fMW << ((nmw - 1) * p.dx / 2 / kPi) << std::endl;
        }
        done += nb;
        if (!rc && write_outputs && next_out < run.out_iter.size() && done == run.out_iter[next_out] + 1) {
            std::vector<double> E(nf), B(nf), J(nf), Be(nf), rho(nf);
            rc = check(qk_get_fields(ctx, E.data(), B.data(), J.data(), Be.data()), "qk_get_fields");
            if (!rc) rc = check(qk_compute_rho(ctx, rho.data()), "qk_compute_rho");
            if (!rc)
                write_output_set(out_dir, iteration_suffix(run.out_time[next_out]), run.out_grid, run.out_set,
                                 E.data(), B.data(), Be.data(), rho.data(), J.data(), p.dx / p.dt);
            // spectra and phase space of every species (src/main.cpp:2355)
            std::vector<std::vector<double>> recs(3 + p.n_ion);
            std::vector<SpeciesRecords> sp_recs(3 + p.n_ion);
            for (int s = 0; !rc && s < 3 + p.n_ion; ++s) {
                int64_t n = 0;
                rc = check(qk_particle_count(ctx, s, &n), "qk_particle_count");
                recs[s].resize((size_t)std::max<int64_t>(n, 1) * 10);
                if (!rc) rc = check(qk_get_particles(ctx, s, n, recs[s].data(), nullptr, &n), "qk_get_particles");
                sp_recs[s] = SpeciesRecords{recs[s].data(), n};
            }
            if (!rc)
                write_spectrum_phasespace(out_dir, iteration_suffix(run.out_time[next_out]), run.spectra,
                                          sp_recs.data(), p.n_ion, p.ion_cmr, p.dx, p.dy, p.dz, run.lambda_cm,
                                          ps_counters);
            ++next_out;
        }
        // add_moving_window_particles (src/main.cpp:1750-1835, 2396-2398)
        if (!rc && moves && (last + 1) * p.dt < run.t_add_mw)
            for (const CellFill &f : moving_window_fills(run, nmw)) {
                rc = check(qk_fill_cells(ctx, f.species, (int64_t)f.density.size(), f.ijk.data(), f.density.data(),
                                         f.npic, f.ux0, 0, nullptr),
                           "qk_fill_cells");
                if (rc) break;
            }
    }
    if (!rc)
        std::printf("{\"steps\": %d, \"seconds\": %.6f, \"ms_per_step\": %.6f, \"particle_steps_per_s\": %.6g}\n", steps,
                    seconds, steps ? seconds * 1e3 / steps : 0.0, seconds > 0 ? pushed / seconds : 0.0);
    qk_destroy(ctx);
    return rc ? 1 : 0;
}
