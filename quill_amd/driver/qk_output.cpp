// qk_output.cpp -- plane writers in the reference's file formats (see
// qk_output.h for the layout and the reference functions they follow).
#include "qk_output.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <algorithm>
#include <vector>

namespace qkd {

namespace {

inline size_t at(const OutputGrid &g, int i, int j, int k, int c)
{
    return 3 * (((size_t)i * g.ny + j) * g.nz + k) + c;
}

// values of one plane file in order: xy rows and xz rows for every i, then
// the yz plane
template <class F>
void for_each_value(const OutputGrid &g, F &&f)
{
    for (int i = 0; i < g.nx; ++i) {
        const int k = g.nz / 2;
        for (int j = 0; j < g.ny; ++j) f(i, j, k, 0);
        const int j = g.ny / 2;
        for (int kk = 0; kk < g.nz; ++kk) f(i, j, kk, 1);
    }
    for (int j = 0; j < g.ny; ++j)
        for (int k = 0; k < g.nz; ++k) f(g.i_yz, j, k, 2);
}

}  // namespace

void write_component(const std::string &path, const OutputGrid &g, const double *aos, int comp)
{
    if (g.binary) {
        // write_binary / write_binary_yzplane (src/fout.cpp:180-203): doubles
        std::vector<double> buf;
        buf.reserve((size_t)g.nx * (g.ny + g.nz) + (size_t)g.ny * g.nz);
        for_each_value(g, [&](int i, int j, int k, int) { buf.push_back(aos[at(g, i, j, k, comp)]); });
        std::ofstream f(path, std::ios::out | std::ios::binary);
        f.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)(buf.size() * sizeof(double)));
    } else {
        std::ofstream f(path, std::ios::out);
        for_each_value(g, [&](int i, int j, int k, int) { f << aos[at(g, i, j, k, comp)] << "\n"; });
    }
}

void write_w_inv(const std::string &path, const OutputGrid &g, const double *E, const double *Be, bool inv)
{
    std::ofstream f(path, std::ios::out);
    for_each_value(g, [&](int i0, int j0, int k0, int part) {
        // clamped node (fout_w / fout_inv / *_yzplane, src/fout.cpp:437-561)
        int i = i0, j = j0, k = k0;
        if (i == g.nx - 1) i -= 1;
        if (part == 0 && j == g.ny - 1) j -= 1;
        if (part == 1 && k == g.nz - 1) k -= 1;
        if (part == 2) {
            if (j == g.ny - 1) j -= 1;
            if (k == g.nz - 1) k -= 1;
        }
        // e_to_particle / b_to_particle (src/deposition.cpp:431-471) at an
        // integer position: weights (1, 0)
        const double xa = 0, xb = 1, ya = 0, yb = 1, za = 0, zb = 1;
        auto gather = [&](const double *F, double v[3]) {
            v[0] = F[at(g, i, j, k, 0)] * yb * zb + F[at(g, i, j + 1, k, 0)] * ya * zb +
                   F[at(g, i, j, k + 1, 0)] * yb * za + F[at(g, i, j + 1, k + 1, 0)] * ya * za;
            v[1] = F[at(g, i, j, k, 1)] * xb * zb + F[at(g, i + 1, j, k, 1)] * xa * zb +
                   F[at(g, i, j, k + 1, 1)] * xb * za + F[at(g, i + 1, j, k + 1, 1)] * xa * za;
            v[2] = F[at(g, i, j, k, 2)] * xb * yb + F[at(g, i + 1, j, k, 2)] * xa * yb +
                   F[at(g, i, j + 1, k, 2)] * xb * ya + F[at(g, i + 1, j + 1, k, 2)] * xa * ya;
        };
        double e[3], b[3];
        gather(E, e);
        gather(Be, b);
        const double w = inv ? e[0] * e[0] + e[1] * e[1] + e[2] * e[2] - b[0] * b[0] - b[1] * b[1] - b[2] * b[2]
                             : e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        f << w << "\n";
    });
}

void write_output_set(const std::string &dir, const std::string &suffix, const OutputGrid &g, const OutputSet &s,
                      const double *E, const double *B, const double *Be, const double *rho, const double *J,
                      double j_scale)
{
    const char *xyz = "xyz";
    for (int c = 0; c < 3; ++c) {
        if (s.e[c] && E) write_component(dir + "/e" + xyz[c] + suffix, g, E, c);
        if (s.b[c] && B) write_component(dir + "/b" + xyz[c] + suffix, g, B, c);
    }
    if (E && Be) {
        write_w_inv(dir + "/w" + suffix, g, E, Be, false);
        write_w_inv(dir + "/inv" + suffix, g, E, Be, true);
    }
    if (rho) {
        write_component(dir + "/rho" + suffix, g, rho, 0);
        if (s.rho_p) write_component(dir + "/rho_p" + suffix, g, rho, 1);
        if (s.rho_ph) write_component(dir + "/rho_ph" + suffix, g, rho, 2);
    }
    if (J && (s.j[0] || s.j[1] || s.j[2])) {
        std::vector<double> js(J, J + 3 * (size_t)g.nx * g.ny * g.nz);
        for (double &v : js) v *= j_scale;
        for (int c = 0; c < 3; ++c)
            if (s.j[c]) write_component(dir + "/j" + xyz[c] + suffix, g, js.data(), c);
    }
}

std::string iteration_suffix(double t)
{
    const double pi = 3.14159265358979323846;
    const double value = static_cast<int>(std::round(t / 2 / pi * 100.)) / 100.;
    char buf[100];
    std::snprintf(buf, sizeof buf, "%g", value);
    return buf;
}

namespace {

constexpr double kPiOut = 3.14159265358979323846;
constexpr double kProtonMassOut = 1836.1526721;   // include/physical_constants.h:4

// %g of the charge-to-mass ratio, as the ion file names use it
std::string cmr_name(double cmr)
{
    char b[32];
    std::snprintf(b, sizeof b, "%g", cmr);
    return b;
}

}  // namespace

void write_spectrum_phasespace(const std::string &dir, const std::string &suffix, const SpectraConf &c,
                               const SpeciesRecords *species, int n_ion, const double *icmr, double dx, double dy,
                               double dz, double lambda_cm, PhasespaceCounters &k)
{
    // every stream is opened, written or not, as the reference opens them
    std::ofstream ps(dir + "/phasespace" + suffix), ps_p(dir + "/phasespace_p" + suffix),
        ps_ph(dir + "/phasespace_ph" + suffix);
    std::vector<std::ofstream> ps_i(n_ion);
    for (int m = 0; m < n_ion; ++m) ps_i[m].open(dir + "/phasespace_" + cmr_name(icmr[m]) + "_" + suffix);
    std::vector<double> sp(c.neps, 0.0), sp_p(c.neps_p, 0.0), sp_ph(c.neps_ph, 0.0);
    std::vector<std::vector<double>> sp_i(n_ion, std::vector<double>(c.neps_i, 0.0));
    auto record = [&](std::ofstream &f, const double *r) {
        f << r[7] << "\n";
        f << dx * r[0] / (2 * kPiOut) << "\n";
        f << dy * r[1] / (2 * kPiOut) << "\n" << dz * r[2] / (2 * kPiOut) << "\n";
        f << r[3] << "\n" << r[4] << "\n" << r[5] << "\n";
        f << r[6] << "\n";
        f << r[9] << "\n";
    };
    // the traversal order: cells i, j, k (the reference walks cp[i][j][k]),
    // the records of one cell in the order given
    const int ns = 3 + n_ion;
    struct Ref {
        int64_t cell;
        int s;
        int64_t n;
    };
    std::vector<Ref> order;
    for (int s = 0; s < ns; ++s)
        for (int64_t n = 0; n < species[s].n; ++n) {
            const double *r = species[s].rec10 + 10 * n;
            const int64_t ci = (int64_t)std::floor(r[0]), cj = (int64_t)std::floor(r[1]), ck = (int64_t)std::floor(r[2]);
            order.push_back(Ref{(ci << 40) | (cj << 20) | ck, s, n});
        }
    std::stable_sort(order.begin(), order.end(), [](const Ref &a, const Ref &b) { return a.cell < b.cell; });
    for (const Ref &o : order) {
        const double *r = species[o.s].rec10 + 10 * o.n;
        const double g = r[6], q = r[7];
        int i_eps;
        if (o.s == 0) {   // electrons
            i_eps = (int)((g - 1) * 0.511 / c.deps);
            if (i_eps > -1 && i_eps < c.neps) sp[i_eps] -= q;   // q < 0 for electrons
            if (++k.e > c.enthp) {
                k.e = 0;
                record(ps, r);
            }
        } else if (o.s == 1 && c.write_p) {
            i_eps = (int)((g - 1) * 0.511 / c.deps_p);
            if (i_eps > -1 && i_eps < c.neps_p) sp_p[i_eps] += q;
            if (++k.p > c.enthp_p) {
                k.p = 0;
                record(ps_p, r);
            }
        } else if (o.s == 2 && c.write_ph) {
            i_eps = (int)(g * 0.511 / c.deps_ph);
            if (i_eps > -1 && i_eps < c.neps_ph) sp_ph[i_eps] += q;
            if (++k.ph > c.enthp_ph) {
                k.ph = 0;
                record(ps_ph, r);
            }
        } else if (o.s >= 3) {
            const int m = o.s - 3;
            i_eps = (int)((g - 1) * 0.511 * kProtonMassOut / c.deps_i);
            if (i_eps > -1 && i_eps < c.neps_i) sp_i[m][i_eps] += q;
            if (++k.i > c.enthp_i) {
                k.i = 0;
                record(ps_i[m], r);
            }
        }
    }
    // dN/deps in 1/MeV (nucleons per MeV for ions)
    const double base = 1.11485e13 * lambda_cm * dx * dy * dz / (8 * kPiOut * kPiOut * kPiOut);
    std::ofstream fs(dir + "/spectrum" + suffix), fs_p(dir + "/spectrum_p" + suffix),
        fs_ph(dir + "/spectrum_ph" + suffix);
    std::vector<std::ofstream> fs_i(n_ion);
    for (int m = 0; m < n_ion; ++m) fs_i[m].open(dir + "/spectrum_" + cmr_name(icmr[m]) + "_" + suffix);
    for (double v : sp) fs << v * (base / c.deps) << "\n";
    if (c.write_p)
        for (double v : sp_p) fs_p << v * (base / c.deps_p) << "\n";
    if (c.write_ph)
        for (double v : sp_ph) fs_ph << v * (base / c.deps_ph) << "\n";
    for (int m = 0; m < n_ion; ++m)
        for (double v : sp_i[m]) fs_i[m] << v * (base / (icmr[m] * c.deps_i)) << "\n";
}

}  // namespace qkd
