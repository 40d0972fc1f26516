// qk_output.h -- the reference's output files for one subdomain (n_sr = 1):
// field planes ex ey ez bx by bz (write_fields, src/main.cpp:745-928 with
// spatial_region::fout_ex ..., src/fout.cpp:180-435), the field energy
// density w and invariant inv (:437-561), and the density planes rho,
// rho_p, rho_ph, irho_<cmr>_, jx jy jz (write_density, src/main.cpp:369-476,
// fout_rho, src/fout.cpp:563-660).
//
// Every plane file holds, for i = 0 .. nx-1, the row j = 0 .. ny-1 at
// k = nz/2 and then the row k = 0 .. nz-1 at j = ny/2, followed by the yz
// plane at the x0fout index; values as text (ostream default formatting,
// one per line) or as raw doubles (output_mode = binary).  w and inv are
// always text.
#pragma once

#include <cstdint>
#include <string>

namespace qkd {

struct OutputGrid {
    int nx, ny, nz;
    int i_yz;        // x index of the yz plane, int((xlength - x0fout) / dx)
    bool binary;     // output_mode = binary
};

// one component (0..2) of an AoS array double[nx][ny][nz][3]
void write_component(const std::string &path, const OutputGrid &g, const double *aos, int comp);
// w = |E|^2 + |Be|^2 and inv = |E|^2 - |Be|^2 at the nodes (e_to_particle /
// b_to_particle at integer positions), edge nodes clamped as fout_w does
void write_w_inv(const std::string &path, const OutputGrid &g, const double *E, const double *Be, bool inv);

// Which files an output iteration writes (e/b_components_for_output,
// particles_for_output, j_components_for_output)
struct OutputSet {
    bool e[3] = {false, false, false}, b[3] = {false, false, false};
    bool rho_p = false, rho_ph = false;
    bool j[3] = {false, false, false};
};

// write the files of one output iteration into dir with the given suffix:
// ex.. bz, w, inv, rho / rho_p / rho_ph (slots 0, 1, 2 of rho), jx jy jz
// (J scaled by dx / dt as scale_j does); any array may be null when its
// files are not asked for
void write_output_set(const std::string &dir, const std::string &suffix, const OutputGrid &g, const OutputSet &s,
                      const double *E, const double *B, const double *Be, const double *rho, const double *J,
                      double j_scale);

// Spectra and phase space (write_spectrum_phasespace, src/main.cpp:478-743)
struct SpectraConf {
    double deps = 0, deps_p = 0, deps_ph = 0, deps_i = 0;   // bin widths, MeV (MeV per nucleon for ions)
    int neps = 0, neps_p = 0, neps_ph = 0, neps_i = 0;       // bins
    int enthp = 0, enthp_p = 0, enthp_ph = 0, enthp_i = 0;   // every (enthp + 1)-th particle goes to phasespace
    bool write_p = false, write_ph = false;                  // particles_for_output has p / ph
    std::string particles_for_output = "e";
};

// the reference's running counters of the phase-space thinning (global in
// the reference, src/main.cpp:66, so they carry over between outputs)
struct PhasespaceCounters {
    int e = 0, p = 0, ph = 0, i = 0;
};

// the live particles of one species as qk_get_particles gives them:
// n x (x y z ux uy uz g q cmr chi), positions in cells
struct SpeciesRecords {
    const double *rec10 = nullptr;
    int64_t n = 0;
};

// spectrum<suffix>, spectrum_p, spectrum_ph, spectrum_<cmr>_ and
// phasespace<suffix>, phasespace_p, phasespace_ph, phasespace_<cmr>_ for one
// subdomain (n_sr = 1): every particle in the reference's traversal order
// (cells i, j, k; within a cell the order of the records given), binned by
// energy and thinned to every (enthp + 1)-th into 9-line records
// (q, x y z in lambda, ux uy uz, g, chi).  species: electrons, positrons,
// photons, then the n_ion ion populations of charge-to-mass ratios icmr.
void write_spectrum_phasespace(const std::string &dir, const std::string &suffix, const SpectraConf &c,
                               const SpeciesRecords *species, int n_ion, const double *icmr, double dx, double dy,
                               double dz, double lambda_cm, PhasespaceCounters &k);

// file-name suffix of an output time t (c/omega units):
// OutputIterations::get_current_iteration_string (src/parameters.cpp:42-50)
std::string iteration_suffix(double t);

}  // namespace qkd
