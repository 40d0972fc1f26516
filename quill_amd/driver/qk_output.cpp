// qk_output.cpp -- plane writers in the reference's file formats (see
// qk_output.h for the layout and the reference functions they follow).
#include "qk_output.h"

#include <cmath>
#include <cstdio>
#include <fstream>
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

}  // namespace qkd
