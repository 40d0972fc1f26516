// qk_setup.h -- from parsed .conf variables to a run of libquillk.so: the
// part of QUILL's init() (src/main.cpp:2558-3600), init_fields()
// (:1041-1228) and init_films() (:1241-1290) that sets up one subdomain's
// per-step loop.
#pragma once

#include <string>
#include <vector>

#include "quillk.h"
#include "qk_conf.h"
#include "qk_output.h"

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
    int ions;               // film() ions argument: 0 none, 1 ions of cmr, 2 positrons
    double cmr;             // of the ions
    double gradwidth, gradwidth_y;
    double y0, y1, z0, z1;
    double T, vx;
    int npic[3];
    // the values the reference's film object keeps (printed in the log)
    double log_width = 0, log_y1 = 0, log_z1 = 0, log_ne = 0, log_mcr = 1;
    int log_npic[3] = {0, 0, 0};
};

struct Run {
    qk_params params;
    double lambda_cm = 0;           // laser wavelength, cm
    int last_iteration = 0;         // the loop runs last_iteration + 1 steps
    std::vector<LaserCall> lasers;
    std::vector<Film> films;
    std::vector<double> ion_mcr;    // ion populations: mass / (proton mass * charge)
    std::string data_folder;
    // output iterations and their times (OutputIterations, src/parameters.cpp:6-38)
    std::vector<int> out_iter;
    std::vector<double> out_time;
    OutputSet out_set;
    OutputGrid out_grid{};
    SpectraConf spectra;
    std::vector<double> t_end_list, period_list;   // c/omega, in config order
    double x0fout = 0;                             // c/omega
    std::string log;                               // the text of the log file (src/main.cpp:3611-3787)
    // moving window (src/main.cpp:2665-2679, 2932-2958, 3496-3525)
    int mwindow = 0;                 // 0 off, 1 on (mwspeed), 2 auto (mwtolerance)
    double mwspeed = 0, mwtolerance = 0;
    double t_add_mw = 0;             // c/omega; columns are added while (l + 1) dt < t_add_mw
    bool mwseed = false;             // electron column of the plasma density
    bool mwseed_ions = false;        // ion column of the same density
    double mw_cmr = 0;               // charge-to-mass ratio of those ions
    double ne_ncr = 0;               // 1 / k0^2: the plasma density in n_cr
    int npic[3] = {1, 1, 1};         // xnpic ynpic znpic
    double ylength = 0, zlength = 0; // c/omega
    std::vector<double> ne_x_coords, ne_x_values, ne_r_coords, ne_r_values;   // density profiles
};

// Fill `run` from the config.  Options whose physics is outside this
// library (beams, focused envelopes, thermal films, neutral flows, particle
// catching) are rejected with a message in err rather than silently ignored.
bool build_run(const std::vector<ConfVar> &vars, Run &run, std::string &err);

// lattice particles of one film for one species (fill_cell_by_particles,
// src/init.cpp:402-478, with T = 0): absolute positions in cells
struct ParticleSet {
    std::vector<double> x, y, z, ux, uy, uz, q;
};
void film_particles(const Run &run, const Film &f, double cmr, ParticleSet &out);

// cells of one fill_cell_by_particles pass (qk_fill_cells arguments)
struct CellFill {
    int species;                   // QK_ELECTRON, QK_ION0 + n
    std::vector<int32_t> ijk;      // cell triples in the reference's loop order
    std::vector<double> density;   // n per cell
    int npic[3];
    double ux0;                    // initial ux of every particle
};
// the cells film() visits (src/init.cpp:537-646, T = 0, no transverse
// profile) with the film's edges at x0, x1 (c/omega from the subdomain
// origin); column = append_for_moving_window (only i = nx - 3)
void film_cells(const Run &run, const Film &f, double x0, double x1, bool column, std::vector<int32_t> &ijk,
                std::vector<double> &density);
// lin_interpolation (src/low-level.cpp:492-510)
double lin_interpolation(double x, const std::vector<double> &xs, const std::vector<double> &vs);
// add_moving_window_particles (src/main.cpp:1750-1835) after the shift that
// made nmw: the electron (and ion) seed column, then each film's column
std::vector<CellFill> moving_window_fills(const Run &run, int nmw);

}  // namespace qkd
