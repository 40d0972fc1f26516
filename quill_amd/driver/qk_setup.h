// qk_setup.h -- from parsed .conf variables to a run of libquillk.so: the
// part of QUILL's init() (src/main.cpp:2558-3600), init_fields()
// (:1041-1228) and init_films() (:1241-1290) that sets up one subdomain's
// per-step loop.
#pragma once

#include <string>
#include <vector>

#include "quillk.h"
#include "qk_conf.h"

namespace qkd {

// one call of f_init_cos (src/init.cpp:6-210) in the argument order of
// qk_init_laser_cos
struct LaserCall {
    double L[16];
};

// film() (src/init.cpp:537-646) after init_films' argument conversion
struct Film {
    double x0, x1;          // c/omega units, subdomain origin
    double ne_y0, ne_y1;    // n_cr units
    bool ions;
    double cmr;             // of the ions
    double gradwidth, gradwidth_y;
    double y0, y1, z0, z1;
    double T, vx;
    int npic[3];
};

struct Run {
    qk_params params;
    double lambda_cm = 0;           // laser wavelength, cm
    int last_iteration = 0;         // the loop runs last_iteration + 1 steps
    std::vector<LaserCall> lasers;
    std::vector<Film> films;
    std::vector<double> ion_mcr;    // ion populations: mass / (proton mass * charge)
    std::string data_folder;
};

// Fill `run` from the config.  Options whose physics is outside this
// library (moving window, merging, beams, focused envelopes, thermal films,
// neutral flows, particle catching) are rejected with a message in err
// rather than silently ignored.
bool build_run(const std::vector<ConfVar> &vars, Run &run, std::string &err);

// lattice particles of one film for one species (fill_cell_by_particles,
// src/init.cpp:402-478, with T = 0): absolute positions in cells
struct ParticleSet {
    std::vector<double> x, y, z, ux, uy, uz, q;
};
void film_particles(const Run &run, const Film &f, double cmr, ParticleSet &out);

}  // namespace qkd
