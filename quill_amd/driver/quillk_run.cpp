// quillk_run.cpp -- command-line driver: QUILL's main() for one subdomain
// (src/main.cpp:2180-2340 at n_sr = 1) on libquillk.so.  Reads a .conf
// file, sets up fields and particles as init(), init_fields() and
// init_films() do, runs the step loop and writes the per-step N and energy
// records in the reference's text format (write_N / write_energy,
// src/main.cpp:160-187).
//
//   quillk-run CONF [--steps N] [--fp64] [--deposit tiled|global]
//              [--sort-every N] [--device D] [--out DIR] [--dry-run]
//
// --dry-run prints the derived run as JSON and touches no GPU;
// --print-conf prints the parsed variables as JSON; --dump-setup DIR writes
// the initial particles of every species (raw little-endian doubles, rows
// of x y z ux uy uz q, file species<slot>.f64) and the laser calls
// (lasers.f64, rows of 16) for an independent run of the same set-up.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "qk_conf.h"
#include "qk_setup.h"
#include "quillk.h"

using namespace qkd;

namespace {

int usage()
{
    std::fprintf(stderr,
                 "usage: quillk-run CONF [--steps N] [--fp64] [--deposit tiled|global] [--sort-every N]\n"
                 "                  [--device D] [--out DIR] [--dry-run] [--print-conf] [--dump-setup DIR]\n");
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
    for (size_t n = 0; n < run.ion_mcr.size(); ++n) out.push_back({QK_ION0 + (int)n, run.params.ion_cmr[n], {}});
    for (const Film &f : run.films) {
        film_particles(run, f, -1, out[0].set);
        if (f.ions)
            for (size_t n = 1; n < out.size(); ++n)
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
    std::printf(" \"ion_cmr\": [");
    for (int n = 0; n < p.n_ion; ++n) std::printf("%s%.17g", n ? ", " : "", p.ion_cmr[n]);
    std::printf("],\n \"lasers\": [");
    for (size_t n = 0; n < run.lasers.size(); ++n) {
        std::printf("%s[", n ? ", " : "");
        for (int m = 0; m < 16; ++m) std::printf("%s%.17g", m ? ", " : "", run.lasers[n].L[m]);
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

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) return usage();
    std::string conf = argv[1], out_dir, setup_dir;
    int steps_override = -1, device = 0, sort_every = -1;
    bool fp64 = false, dry = false, global_dep = false, show_conf = false;
    for (int a = 2; a < argc; ++a) {
        const std::string s = argv[a];
        auto next = [&]() -> const char * { return a + 1 < argc ? argv[++a] : nullptr; };
        if (s == "--steps") { const char *v = next(); if (!v) return usage(); steps_override = std::atoi(v); }
        else if (s == "--fp64") fp64 = true;
        else if (s == "--dry-run") dry = true;
        else if (s == "--print-conf") show_conf = true;
        else if (s == "--dump-setup") { const char *v = next(); if (!v) return usage(); setup_dir = v; }
        else if (s == "--deposit") { const char *v = next(); if (!v) return usage(); global_dep = std::strcmp(v, "global") == 0; }
        else if (s == "--sort-every") { const char *v = next(); if (!v) return usage(); sort_every = std::atoi(v); }
        else if (s == "--device") { const char *v = next(); if (!v) return usage(); device = std::atoi(v); }
        else if (s == "--out") { const char *v = next(); if (!v) return usage(); out_dir = v; }
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

    std::ofstream fN(out_dir + "/N"), fE(out_dir + "/energy");
    const int batch = 100;
    std::vector<qk_diag> diag(batch);
    double seconds = 0;
    int64_t pushed = 0;
    for (int done = 0; !rc && done < steps;) {
        const int nb = std::min(batch, steps - done);
        const auto t0 = std::chrono::steady_clock::now();
        rc = check(qk_step(ctx, nb, diag.data()), "qk_step");
        seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int l = 0; !rc && l < nb; ++l) {
            const qk_diag &d = diag[l];
            fN << d.N_e << '\t' << d.N_p << '\t' << d.N_ph << std::endl;
            fE << d.energy_f << '\t' << d.energy_e << '\t' << d.energy_p << '\t' << d.energy_ph;
            for (int n = 0; n < p.n_ion; ++n) fE << '\t' << d.energy_ions[n];
            fE << std::endl;
            pushed += d.n_e + d.n_p + d.n_ph + d.n_ion;
        }
        done += nb;
    }
    if (!rc)
        std::printf("{\"steps\": %d, \"seconds\": %.6f, \"ms_per_step\": %.6f, \"particle_steps_per_s\": %.6g}\n", steps,
                    seconds, steps ? seconds * 1e3 / steps : 0.0, seconds > 0 ? pushed / seconds : 0.0);
    qk_destroy(ctx);
    return rc ? 1 : 0;
}
