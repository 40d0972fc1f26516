// qk_setup.cpp -- config variables -> libquillk.so run description.  Each
// block follows the part of QUILL's init() named in its comment; units are
// converted exactly as there (lengths to c/omega by 2 pi, "um" and "fs"
// conversions with the laser wavelength in cm).
#include "qk_setup.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <sstream>

namespace qkd {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kNcrCoef = 1.11485e13;          // n_cr lambda^2, cm^-1 (src/main.cpp:2589)
constexpr double kLightCm = 2.99792458e10;       // cm/s
constexpr double kProtonMass = 1836.1526721;     // m_p / m_e (include/physical_constants.h:4)
constexpr double kComptonCm = 2.4263086e-10;     // lambda_C, cm (src/main.cpp:2276)

struct Ctx {
    const std::vector<ConfVar> &v;
    double lambda = 0;   // cm
    const ConfVar &get(const char *name, size_t from = 0) const { return find(v, name, from); }
};

// a length in lambda units (optionally um or fs) to c/omega
double length(const Ctx &c, const ConfVar &var, bool allow_fs)
{
    double x = var.value;
    if (var.units == "um") x = x * 1e-4 / c.lambda;
    else if (allow_fs && var.units == "fs") x = x * 1e-15 * kLightCm / c.lambda;
    return x * 2 * kPi;
}

// convert_units(value, units, "1/k") (src/main.cpp:2465-2491), with "#"
// (a plain number) read as `dflt` as get_units does (:2493-2495)
bool time_to_k(const Ctx &c, const ConfVar &var, const char *dflt, double &out, std::string &err)
{
    const std::string u = var.units == "#" ? std::string(dflt) : var.units;
    double x = var.value;
    if (u == "um") x *= 2 * kPi * 1e-4 / c.lambda;
    else if (u == "mm") x *= 2 * kPi * 0.1 / c.lambda;
    else if (u == "cm") x *= 2 * kPi / c.lambda;
    else if (u == "lambda") x *= 2 * kPi;
    else {
        err = "unknown time units [" + u + "] for " + var.name;
        return false;
    }
    out = x;
    return true;
}

bool find_boolean(const Ctx &c, const char *name, bool dflt)
{
    const std::string &u = c.get(name).units;
    if (u == "on" || u == "true") return true;
    if (u == "off" || u == "false") return false;
    return dflt;
}

// the output iterations and times of OutputIterations (src/parameters.cpp:6-38)
void output_iterations(const std::vector<double> &t_end, const std::vector<double> &period, double dt,
                       std::vector<int> &iters, std::vector<double> &times)
{
    int previous = -1;
    double t_begin = 0;
    for (size_t n = 0; n < t_end.size(); ++n) {
        const int count = (int)((t_end[n] + 0.5 * dt - t_begin) / period[n]) + 1;
        for (int j = 0; j < count; ++j) {
            const double time = t_begin + j * period[n];
            const int it = (int)std::round(time / dt);
            if (it > previous) {
                iters.push_back(it);
                times.push_back(time);
                previous = it;
            }
        }
        t_begin = t_end[n];
    }
    const int final_it = (int)std::round(t_end.back() / dt);
    if (final_it > previous) {
        iters.push_back(final_it);
        times.push_back(t_end.back());
    }
}

std::vector<std::string> split_amp(const std::string &s)
{
    std::vector<std::string> out(3);
    size_t n = 0, b = 0;
    for (size_t i = 0; i <= s.size() && n < 3; ++i)
        if (i == s.size() || s[i] == '&') {
            out[n++] = s.substr(b, i - b);
            b = i + 1;
        }
    return out;
}

}  // namespace

bool build_run(const std::vector<ConfVar> &vars, Run &run, std::string &err)
{
    Ctx c{vars};
    qk_params &p = run.params;
    p = qk_params{};

    // grid steps and wavelength (src/main.cpp:2573-2614)
    double dx = c.get("dx").value * 2 * kPi, dy = c.get("dy").value * 2 * kPi, dz = c.get("dz").value * 2 * kPi;
    double dt = c.get("dt").value * 2 * kPi;
    const ConfVar &lam = c.get("lambda");
    c.lambda = lam.units == "um" ? lam.value * 1e-4 : lam.value;
    const ConfVar &nev = c.get("ne");
    double ne = nev.units == "ncr" ? nev.value * kNcrCoef / c.lambda / c.lambda : nev.value;
    if (lam.units == "lambda_p") c.lambda = lam.value * std::sqrt(kNcrCoef / ne);
    if (c.lambda <= 0) {
        err = "lambda must be positive";
        return false;
    }
    const double k0 = std::sqrt(kNcrCoef / c.lambda / c.lambda / ne);
    if (dt == 0) {
        dt = 0.5 * k0 * std::sqrt(4.04);
        dx = 0;
    }
    if (dx == 0) {
        dx = (1 + 1e-4) * dt / (1 - 1 / (k0 * k0) * dt * dt / 4.04);
        if (dx < 0) {
            err = "cannot calculate stable dx, reduce dt";
            return false;
        }
    }
    run.lambda_cm = c.lambda;

    // box (src/main.cpp:2616-2644)
    const double xlength = length(c, c.get("xlength"), true);
    const double ylength = length(c, c.get("ylength"), false);
    const double zlength = length(c, c.get("zlength"), false);

    // run length (src/main.cpp:2646-2663)
    {
        std::vector<double> t_end, period;
        size_t from = 0;
        do {
            const size_t it = find_index(vars, "t_end", from);
            if (it == vars.size()) {
                err = "t_end is not set";
                return false;
            }
            double te, op;
            if (!time_to_k(c, vars[it], "lambda", te, err)) return false;
            from = it + 1;
            const ConfVar &opv = c.get("output_period", from);
            if (opv.units == "t_end") op = opv.value * te;
            else if (!time_to_k(c, opv, "lambda", op, err)) return false;
            if (!(op > 0)) {
                err = "output_period must be positive";
                return false;
            }
            t_end.push_back(te);
            period.push_back(op);
        } while (c.get("t_end", from).value != 0);
        output_iterations(t_end, period, dt, run.out_iter, run.out_time);
        run.t_end_list = t_end;
        run.period_list = period;
        run.last_iteration = run.out_iter.back();
    }

    // laser (src/main.cpp:2681-2905)
    const double xsigma = length(c, c.get("xsigma"), true);
    const double ysigma = length(c, c.get("ysigma"), false);
    const double zsigma = length(c, c.get("zsigma"), false);
    double x0 = length(c, c.get("x0"), true);
    double a0y = 0, a0z = 0;
    {
        const ConfVar &P = c.get("P");
        double pw = P.value * (P.units == "PW" ? 1e22 : (P.units == "TW" ? 1e19 : 1));
        if (pw != 0) a0y = 5.27281e+17 * std::sqrt(2 * pw / (kPi * 2.42161e+52 * ysigma * zsigma));
        const ConfVar &I = c.get("I");
        double iv = I.value * (I.units == "W/cm^2" ? 1e7 : 1);
        if (iv != 0) a0y = 5.27281e+17 * c.lambda * std::sqrt(iv / (kPi * 2.42161e+52));
        const ConfVar &W = c.get("W");
        double wv = W.value * (W.units == "J" ? 1e7 : 1);
        if (wv != 0) {
            a0y = std::sqrt(wv / (3.691e4 * c.lambda * kPi * std::sqrt(kPi / 2) * xsigma * ysigma * zsigma / 4));
            if (c.get("polarization").units == "circular") {
                a0y = a0y / std::sqrt(2);
                a0z = a0y;
            }
        }
        if (c.get("a0").value != 0) a0y = c.get("a0").value;
        const ConfVar &pol = c.get("polarization");
        if (pol.units == "circular") a0z = a0y;
        if (pol.units == "elliptic") a0z = pol.value * a0y;
        if (c.get("a0y").value != 0) {
            a0y = c.get("a0y").value;
            a0z = c.get("a0z").value;
        } else if (c.get("a0z").value != 0) {
            a0z = c.get("a0z").value;
            a0y = 0;
        }
    }
    p.external_bz = c.get("external_bz").value;
    std::string f_envelope = c.get("f_envelope").units;
    if (f_envelope.empty()) f_envelope = "cos";
    if (f_envelope != "cos") {
        err = "f_envelope = " + f_envelope + " is not supported (only the cos envelope with sscos = 0)";
        return false;
    }
    const double b_sign = c.get("b_sign").value != -1 ? 1 : 0;
    auto angle = [&](const char *name) {
        const ConfVar &v = c.get(name);
        return v.units == "pi" ? v.value * kPi : v.value;
    };
    const double phase = angle("phase"), phi = angle("phi"), phi_rotate = angle("phi_rotate");
    double y00 = length(c, c.get("y0"), false), z00 = length(c, c.get("z0"), false);
    {
        const ConfVar &r0v = c.get("r0");
        if (r0v.units != "off" && r0v.value != 0) {
            const double r0 = length(c, r0v, false);
            const ConfVar &th = c.get("theta");
            const double theta = th.units == "deg" ? th.value * kPi / 180 : th.value;
            x0 = -r0 * std::cos(theta) + 0.00001;
            y00 = -r0 * std::sin(theta);
        }
    }
    double target[3] = {0, 0, 0};
    {
        const char *names[3] = {"xtarget", "ytarget", "ztarget"};
        for (int a = 0; a < 3; ++a) {
            const ConfVar &t = c.get(names[a]);
            if (t.units != "off" && t.value != 0) target[a] = length(c, t, false);
        }
    }

    // moving window (src/main.cpp:2665-2679, 2932-2958): on unless "off"
    {
        const ConfVar &mw = c.get("mwindow");
        run.mwindow = mw.units == "off" ? 0 : (mw.units == "auto" ? 2 : 1);
        run.mwtolerance = c.get("mwtolerance").value;
        if (run.mwtolerance == 0) run.mwtolerance = 1e-6;
        run.mwspeed = c.get("mwspeed").value;
        if (run.mwindow == 1 && run.mwspeed == 0) run.mwspeed = 1;
        run.mwseed = c.get("mwseed").units != "off";
        run.mwseed_ions = c.get("mwseed_ions").units == "on";
        double mw_mcr = c.get("mw_mcr").value;
        if (mw_mcr == 0) mw_mcr = 1;
        run.mw_cmr = 1 / (kProtonMass * mw_mcr);
        const ConfVar &ta = c.get("t_add_mw");
        double tv = ta.value;
        if (ta.units == "um") tv = tv * 1e-4 / c.lambda;
        else if (ta.units == "mm") tv = tv / (10 * c.lambda);
        else if (ta.units == "cm") tv = tv / c.lambda;
        run.t_add_mw = tv != 0 ? tv * 2 * kPi : run.out_time.back();
        run.ne_ncr = 1 / (k0 * k0);
        run.ylength = ylength;
        run.zlength = zlength;
        // find_array(first, name, "1/k", "lambda") / (name, "") (src/main.cpp:3496-3525)
        auto coords = [&](const char *name, std::vector<double> &out) {
            const ConfVar &v = c.get(name);
            ConfVar one = v;
            out.clear();
            for (double a : v.array) {
                one.value = a;
                double x;
                if (!time_to_k(c, one, "lambda", x, err)) return false;
                out.push_back(x);
            }
            return true;
        };
        if (!coords("ne_profile_x_coords", run.ne_x_coords) || !coords("ne_profile_r_coords", run.ne_r_coords))
            return false;
        run.ne_x_values = c.get("ne_profile_x_values").array;
        run.ne_r_values = c.get("ne_profile_r_values").array;
        for (int a = 0; a < 2; ++a) {
            const std::vector<double> &xs = a ? run.ne_r_coords : run.ne_x_coords;
            const std::vector<double> &vs = a ? run.ne_r_values : run.ne_x_values;
            if (xs.size() != vs.size()) {
                err = std::string("the sizes of ne_profile_") + (a ? "r" : "x") + "_coords and _values differ";
                return false;
            }
            for (size_t m = 1; m < xs.size(); ++m)
                if (xs[m] < xs[m - 1]) {
                    err = std::string("ne_profile_") + (a ? "r" : "x") + "_coords is not sorted";
                    return false;
                }
        }
    }

    // options outside this library's per-step path are refused
    {
        const bool flows = c.get("nelflow").value != 0 || c.get("nerflow").value != 0;
        if (flows) {
            err = "neutral flows (nelflow / nerflow) are not supported";
            return false;
        }
        if (c.get("beam").units == "on") {
            err = "beam = on is not supported";
            return false;
        }
        const std::string &pm = c.get("pmerging").units;
        if (!pm.empty() && pm != "off" && pm != "ti" && pm != "nl") {
            err = "pmerging = " + pm + " is not a merging algorithm (ti, nl)";
            return false;
        }
        if (c.get("catching").units == "on" || c.get("dump_photons").units == "on") {
            err = "catching / dump_photons are not supported";
            return false;
        }
        if (c.get("shenergy").value != 0) {
            err = "second harmonic (shenergy) is not supported";
            return false;
        }
    }

    // films (src/main.cpp:3053-3231) and ion populations (:3251-3330)
    const std::string ions = c.get("ions").units;
    {
        std::vector<Film> films;   // in config order
        size_t from = 0;
        for (;;) {
            size_t at = from;
            for (;;) {
                at = find_index(vars, "film", at);
                if (at == vars.size() || vars[at].units == "on") break;
                ++at;
            }
            if (at == vars.size()) break;
            from = at + 1;
            Film f{};
            auto len = [&](const char *name, bool fs) {
                const ConfVar &v = c.get(name, from);
                if (v.units == "dx" && std::string(name) == "x0film") return v.value * dx;
                return length(c, v, fs);
            };
            f.x0 = len("x0film", true);
            f.x1 = f.x0 + len("filmwidth", true);
            f.gradwidth = len("gradwidth", true);
            f.gradwidth_y = len("gradwidth_y", true);
            f.y0 = len("y0film", true);
            f.y1 = c.get("y1film", from).value == 0 ? ylength : len("y1film", true);
            f.z0 = len("z0film", true);
            f.z1 = c.get("z1film", from).value == 0 ? zlength : len("z1film", true);
            auto dens = [&](const char *name) {
                const ConfVar &v = c.get(name, from);
                if (v.units == "ncr") return v.value * kNcrCoef / c.lambda / c.lambda;
                if (v.units == "ne") return v.value * ne;
                return v.value;
            };
            double nf = dens("nfilm"), n0 = dens("nfilm_y0"), n1 = dens("nfilm_y1");
            if (nf != 0) n0 = n1 = nf;
            const double to_ncr = 1.0 / (kNcrCoef / c.lambda / c.lambda);
            f.ne_y0 = n0 * to_ncr;
            f.ne_y1 = n1 * to_ncr;
            const double mcr = c.get("mcr", from).value == 0 ? 1 : c.get("mcr", from).value;
            f.cmr = 1 / (kProtonMass * mcr);
            f.ions = ions == "on" ? 1 : (ions == "positrons" ? 2 : 0);
            f.T = c.get("Tfilm", from).value;
            f.vx = c.get("vxfilm", from).value;
            f.npic[0] = (int)c.get("xnpic_film", from).value;
            f.npic[1] = (int)c.get("ynpic_film", from).value;
            f.npic[2] = (int)c.get("znpic_film", from).value;
            if (f.T != 0) {
                err = "thermal films (Tfilm != 0) are not supported";
                return false;
            }
            // as the reference's film object keeps them, for the log
            f.log_width = len("filmwidth", true);
            f.log_y1 = len("y1film", true);
            f.log_z1 = len("z1film", true);
            f.log_ne = nf;
            f.log_mcr = mcr;
            for (int a = 0; a < 3; ++a) f.log_npic[a] = f.npic[a];
            films.push_back(f);
            run.ion_mcr.push_back(mcr);
        }
        // the reference keeps films newest first (p_last_film, prev links)
        run.films.assign(films.rbegin(), films.rend());
        std::vector<double> mcrs;
        for (auto it = run.ion_mcr.rbegin(); it != run.ion_mcr.rend(); ++it) {
            bool seen = false;
            for (double m : mcrs) seen = seen || m == *it;
            if (!seen) mcrs.push_back(*it);
        }
        // the ions of mwseed_ions (src/main.cpp:3310-3317)
        if (run.mwseed_ions) {
            double mw_mcr = c.get("mw_mcr").value;
            if (mw_mcr == 0) mw_mcr = 1;
            bool seen = false;
            for (double m : mcrs) seen = seen || m == mw_mcr;
            if (!seen) mcrs.push_back(mw_mcr);
        }
        run.ion_mcr = (ions == "on" || ions == "positrons") ? mcrs : std::vector<double>{};
    }
    const int xnpic = (int)c.get("xnpic").value, ynpic = (int)c.get("ynpic").value, znpic = (int)c.get("znpic").value;
    run.npic[0] = xnpic;
    run.npic[1] = ynpic;
    run.npic[2] = znpic;
    for (Film &f : run.films)
        if (f.npic[0] == 0 || f.npic[1] == 0 || f.npic[2] == 0) {
            f.npic[0] = xnpic;
            f.npic[1] = ynpic;
            f.npic[2] = znpic;
        }
    if ((int)run.ion_mcr.size() > QK_MAX_ION) {
        err = "more than " + std::to_string(QK_MAX_ION) + " ion populations";
        return false;
    }
    run.data_folder = c.get("data_folder").units.empty() ? "results" : c.get("data_folder").units;

    // what the output iterations write (src/main.cpp:2720-2731, 2954-2972, 3387-3395, 3551-3555)
    {
        auto comps = [&](const char *name, bool out[3]) {
            const std::string v = c.get(name).units;
            out[0] = v == "x" || v == "xy" || v == "xz" || v == "xyz";
            out[1] = v == "y" || v == "xy" || v == "yz" || v == "xyz";
            out[2] = v == "z" || v == "xz" || v == "yz" || v == "xyz";
        };
        comps("e_components_for_output", run.out_set.e);
        comps("b_components_for_output", run.out_set.b);
        const std::string jc = c.get("j_components_for_output").units;
        for (int a = 0; a < 3; ++a) run.out_set.j[a] = jc.find("xyz"[a]) != std::string::npos && jc != "none";
        const std::string pf = c.get("particles_for_output").units;
        run.out_set.rho_p = pf == "ep" || pf == "epph";
        run.out_set.rho_ph = pf == "eph" || pf == "epph";
        // spectra and phase space (src/main.cpp:3337-3395): bin widths in MeV
        // (given in MeV or mc^2), 0 meaning the electrons' value
        SpectraConf &sc = run.spectra;
        auto mev = [&](const char *name) {
            const ConfVar &v = c.get(name);
            return v.units == "mc^2" ? v.value * 0.511 : v.value;
        };
        sc.deps = mev("deps");
        sc.deps_p = mev("deps_p");
        if (sc.deps_p == 0) sc.deps_p = sc.deps;
        sc.deps_ph = mev("deps_ph");
        if (sc.deps_ph == 0) sc.deps_ph = sc.deps;
        sc.deps_i = c.get("deps_i").value;   // MeV per nucleon
        if (sc.deps_i == 0) sc.deps_i = sc.deps;
        sc.neps = (int)c.get("neps").value;
        sc.neps_p = (int)c.get("neps_p").value;
        if (sc.neps_p == 0) sc.neps_p = sc.neps;
        sc.neps_ph = (int)c.get("neps_ph").value;
        if (sc.neps_ph == 0) sc.neps_ph = sc.neps;
        sc.neps_i = (int)c.get("neps_i").value;
        if (sc.neps_i == 0) sc.neps_i = sc.neps;
        sc.enthp = (int)c.get("enthp").value;
        sc.enthp_p = (int)c.get("enthp_p").value;
        if (sc.enthp_p == 0) sc.enthp_p = sc.enthp;
        sc.enthp_ph = (int)c.get("enthp_ph").value;
        if (sc.enthp_ph == 0) sc.enthp_ph = sc.enthp;
        sc.enthp_i = (int)c.get("enthp_i").value;
        if (sc.enthp_i == 0) sc.enthp_i = sc.enthp;
        sc.particles_for_output = pf.empty() ? "e" : pf;
        sc.write_p = run.out_set.rho_p;
        sc.write_ph = run.out_set.rho_ph;
        double x0fout = length(c, c.get("x0fout"), true);
        if (x0fout == 0) x0fout = xlength / 2;
        run.x0fout = x0fout;
        run.out_grid.i_yz = (int)((xlength - x0fout) / dx);
        run.out_grid.binary = c.get("output_mode").units == "binary";
    }

    // physics options (src/main.cpp:3531-3598)
    p.qed = find_boolean(c, "qed", true) ? 1 : 0;
    std::string pusher = c.get("pusher").units, solver = c.get("solver").units;
    if (pusher.empty()) pusher = "vay";
    if (solver.empty()) solver = "ndfx";
    if (pusher == "boris") p.pusher = QK_PUSHER_BORIS;
    else if (pusher == "vay") p.pusher = QK_PUSHER_VAY;
    else {
        err = "pusher unknown: " + pusher;
        return false;
    }
    if (solver == "ndfx") p.solver = QK_SOLVER_NDFX;
    else if (solver == "fdtd") p.solver = QK_SOLVER_FDTD;
    else if (solver == "five_point") p.solver = QK_SOLVER_FIVE_POINT;
    else {
        err = "solver unknown: " + solver;
        return false;
    }

    // subdomain (src/main.cpp:2241-2243, n_sr = 1) and the per-step constants
    // of the loop body (:2323-2327)
    p.nx = (int)(xlength / dx);
    p.ny = (int)(ylength / dy);
    p.nz = (int)(zlength / dz);
    run.out_grid.nx = p.nx;
    run.out_grid.ny = p.ny;
    run.out_grid.nz = p.nz;
    if (run.out_grid.i_yz < 0 || run.out_grid.i_yz >= p.nx) {
        err = "x0fout puts the yz output plane outside the box";
        return false;
    }
    p.dx = dx;
    p.dy = dy;
    p.dz = dz;
    p.dt = dt;
    p.e_s = c.lambda / kComptonCm;
    p.q_vac = 8 * kPi * kPi / (dx * dy * dz) * 2.818e-13 / c.lambda;
    p.norm_n = dx * dy * dz * kNcrCoef * c.lambda / (8 * kPi * kPi * kPi);
    p.norm_ef = 0.5 * dx * dy * dz * 3.691e4 * c.lambda / 1e7;
    p.norm_ep = 8.2e-14 * dx * dy * dz * kNcrCoef * c.lambda / (8 * kPi * kPi * kPi);
    p.n_ion = (int)run.ion_mcr.size();
    for (int n = 0; n < p.n_ion; ++n) p.ion_cmr[n] = 1 / (kProtonMass * run.ion_mcr[n]);
    p.precision = QK_FP32;
    p.deposit = QK_DEPOSIT_TILED;
    p.sort_every = 20;
    p.sort_fallback = 0.01;   // and earlier when a laser sweeps particles out of their tiles' windows
    // merging and its limit crnp = crpc x the box in cells (src/main.cpp:1699-1701, 3397-3401)
    {
        const std::string &pm = c.get("pmerging").units;
        p.merging = pm == "ti" ? 1 : (pm == "nl" ? 2 : 0);
        p.crnp = c.get("crpc").value * xlength * ylength * zlength / (dx * dy * dz);
        if (p.merging && !(p.crnp > 0)) {
            err = "pmerging needs crpc > 0";
            return false;
        }
    }

    // laser calls of init_fields, cos envelope (src/main.cpp:1150-1227)
    auto call = [&](double ay, double az, double xc, double xs, double y, double z, double append, double ph) {
        LaserCall l{{ay, az, xsigma, ysigma, zsigma, xc, b_sign, xs, phase, y, z, append, ph, target[0], target[1],
                     target[2]}};
        run.lasers.push_back(l);
    };
    if (a0y != 0 || a0z != 0) {
        if (phi_rotate != 0) {
            call(a0y, a0z, xlength / 2 + x0, x0, y00, z00, 1, phi_rotate);
        } else {
            call(a0y, a0z, xlength / 2 + x0, x0, y00, z00, 1, 0);
            std::string lp = c.get("lp_reflection").units, fr = c.get("f_reflection").units;
            if (lp.empty()) lp = "off";
            if (fr.empty()) fr = "off";
            const std::vector<std::string> L = split_amp(lp), F = split_amp(fr);
            // mirrored copy: amplitude signs from the y / z flags, +x0 or -x0
            auto mirror = [&](const std::string &fy, const std::string &fz, int sx, int sy, int sz) {
                call((1 - 2 * (fy == "y")) * a0y, (1 - 2 * (fz == "z")) * a0z, xlength / 2 + sx * x0, sx * x0,
                     sy * y00, sz * z00, 1, 0);
            };
            if (phi != 0) {
                call(a0y, a0z, xlength / 2 - x0, -x0, -y00, -z00, 1, phi);
            } else if (L[0] == "xy") {
                mirror(F[0], F[0], +1, +1, -1);
                if (L[1] == "xz") {
                    mirror(F[1], F[1], +1, -1, +1);
                    mirror(F[1], F[1], +1, -1, -1);
                    if (L[2] == "yz") {
                        mirror(F[2], F[1], -1, +1, +1);
                        mirror(F[2], F[1], -1, +1, -1);
                        mirror(F[2], F[1], -1, -1, +1);
                        mirror(F[2], F[1], -1, -1, -1);
                    }
                } else if (L[1] == "yz") {
                    mirror(F[1], F[1], -1, +1, +1);
                    mirror(F[1], F[1], -1, +1, -1);
                    if (L[2] == "xz") {
                        mirror(F[2], F[1], -1, -1, +1);
                        mirror(F[2], F[1], +1, -1, +1);
                        mirror(F[2], F[1], -1, -1, -1);
                        mirror(F[2], F[1], +1, -1, -1);
                    }
                }
            } else if (L[0] == "xz") {
                mirror(F[0], F[0], +1, -1, +1);
                if (L[1] == "yz") {
                    mirror(F[1], F[1], -1, -1, +1);
                    mirror(F[1], F[1], -1, +1, +1);
                }
            } else if (L[0] == "yz") {
                mirror(F[0], F[0], -1, +1, +1);
            }
        }
    }
    // the log file main() writes after init() (src/main.cpp:3611-3787):
    // key and value lines up to "$", then the run switches and a summary;
    // lengths and times in lambda units, densities in cm^-3
    {
        std::ostringstream L;
        const double tp = 2 * kPi;
        auto kv = [&](const char *k, auto v) { L << k << "\n" << v << "\n"; };
        auto str = [&](const char *name, const char *dflt) {
            const std::string &u = c.get(name).units;
            return u.empty() ? std::string(dflt) : u;
        };
        kv("dx", dx / 2 / kPi);
        kv("dy", dy / 2 / kPi);
        kv("dz", dz / 2 / kPi);
        kv("dt", dt / 2 / kPi);
        kv("nx", int(xlength / dx));
        kv("ny", int(ylength / dy));
        kv("nz", int(zlength / dz));
        kv("lambda", c.lambda);
        kv("ne", ne);
        for (int i = (int)run.t_end_list.size() - 1; i >= 0; i--) {
            kv("t_end", run.t_end_list[i] / 2 / kPi);
            kv("output_period", run.period_list[i] / 2 / kPi);
        }
        L << "t_add_mw\n" << run.t_add_mw / tp << '\n';
        kv("xsigma", xsigma / 2 / kPi);
        kv("ysigma", ysigma / 2 / kPi);
        kv("zsigma", zsigma / 2 / kPi);
        kv("x0fout", run.x0fout / 2 / kPi);
        kv("a0y", a0y);
        kv("a0z", a0z);
        kv("external_bz", p.external_bz);
        kv("mwindow", run.mwindow == 0 ? "off" : (run.mwindow == 1 ? "on" : "auto"));
        if (run.mwindow != 0) {
            kv("mwseed", run.mwseed ? "on" : "off");
            L << "mwseed_ions\n" << (run.mwseed_ions ? "on" : str("mwseed_ions", "off")) << '\n';
            if (run.mwseed_ions) {
                double mw_mcr = c.get("mw_mcr").value;
                L << "mw_mcr\n" << (mw_mcr == 0 ? 1.0 : mw_mcr) << '\n';
            }
        }
        kv("e_components_for_output", str("e_components_for_output", "none"));
        kv("b_components_for_output", str("b_components_for_output", "none"));
        kv("f_envelope", f_envelope);
        kv("b_sign", 2.0 * b_sign - 1);
        kv("x0", x0 / 2 / kPi);
        kv("y0", y00 / 2 / kPi);
        kv("z0", z00 / 2 / kPi);
        kv("xtarget", target[0] / 2 / kPi);
        kv("ytarget", target[1] / 2 / kPi);
        kv("ztarget", target[2] / 2 / kPi);
        kv("phase", phase);
        kv("lp_reflection", str("lp_reflection", "off"));
        kv("f_reflection", str("f_reflection", "off"));
        kv("phi", phi);
        kv("phi_rotate", phi_rotate);
        kv("shenergy", c.get("shenergy").value);
        kv("shphase", c.get("shphase").units == "pi" ? c.get("shphase").value * kPi : c.get("shphase").value);
        kv("beam", str("beam", "off"));
        kv("beam_particles", str("beam_particles", "e"));
        kv("Nb", c.get("Nb").value);
        kv("epsb", c.get("epsb").units == "mc^2" ? c.get("epsb").value * 0.511 : c.get("epsb").value);
        kv("xb", length(c, c.get("xb"), true) / 2 / kPi);
        kv("rb", length(c, c.get("rb"), false) / 2 / kPi);
        kv("x0b", length(c, c.get("x0b"), true) / 2 / kPi);
        kv("y0b", length(c, c.get("y0b"), false) / 2 / kPi);
        kv("phib", c.get("phib").value);
        kv("mwspeed", run.mwspeed);
        kv("nelflow", 0.0);
        kv("vlflow", c.get("vlflow").value);
        kv("mcrlflow", c.get("mcrlflow").value);
        kv("Tlflow", c.get("Tlflow").value);
        kv("nerflow", 0.0);
        kv("vrflow", c.get("vrflow").value);
        kv("mcrrflow", c.get("mcrrflow").value);
        kv("Trflow", c.get("Trflow").value);
        kv("ions", ions);
        for (const Film &f : run.films) {   // newest first, as p_last_film -> prev
            kv("film", "on");
            kv("x0film", f.x0 / 2 / kPi);
            kv("filmwidth", f.log_width / 2 / kPi);
            kv("gradwidth", f.gradwidth / 2 / kPi);
            kv("gradwidth_y", f.gradwidth_y / 2 / kPi);
            kv("y0film", f.y0 / 2 / kPi);
            kv("y1film", f.log_y1 / 2 / kPi);
            kv("z0film", f.z0 / 2 / kPi);
            kv("z1film", f.log_z1 / 2 / kPi);
            kv("nfilm", f.log_ne);
            kv("mcr", f.log_mcr);
            kv("Tfilm", f.T);
            kv("vxfilm", f.vx);
            kv("xnpic_film", f.log_npic[0]);
            kv("ynpic_film", f.log_npic[1]);
            kv("znpic_film", f.log_npic[2]);
        }
        kv("n_ion_populations", p.n_ion);
        for (int i = 0; i < p.n_ion; i++) kv("icmr", p.ion_cmr[i]);
        kv("xnpic", xnpic);
        kv("ynpic", ynpic);
        kv("znpic", znpic);
        const SpectraConf &sc = run.spectra;
        kv("particles_for_output", sc.particles_for_output);
        kv("deps", sc.deps);
        kv("neps", sc.neps);
        kv("enthp", sc.enthp);
        kv("deps_p", sc.deps_p);
        kv("neps_p", sc.neps_p);
        kv("enthp_p", sc.enthp_p);
        kv("deps_ph", sc.deps_ph);
        kv("neps_ph", sc.neps_ph);
        kv("enthp_ph", sc.enthp_ph);
        kv("deps_i", sc.deps_i);
        kv("neps_i", sc.neps_i);
        kv("enthp_i", sc.enthp_i);
        kv("n_sr", 1);
        kv("n_tracks", (int)c.get("n_tracks").value);
        std::string ptt = str("particles_to_track", "epgi");
        if (ptt == "none") ptt = "";
        kv("particles_to_track", ptt);
        {   // src/main.cpp:3419-3435: um, mm, cm or lambda
            const ConfVar &v = c.get("tr_start");
            double t = v.value;
            if (v.units == "um") t = t * 1e-4 / c.lambda;
            else if (v.units == "mm") t = t / 10 / c.lambda;
            else if (v.units == "cm") t = t / c.lambda;
            kv("tr_start", t * 2 * kPi / 2 / kPi);
        }
        kv("tr_init", c.get("tr_init").units == "volume" ? 1 : 0);
        // the track box, kept in c/omega (:3436-3483)
        for (const char *k : {"xtr1", "ytr1", "ztr1", "xtr2", "ytr2", "ztr2"})
            kv(k, length(c, c.get(k), k[0] == 'x'));
        kv("pmerging", str("pmerging", "off"));
        kv("crpc", c.get("crpc").value);
        L << "$\n";
        L << "catching" << std::endl << "off" << std::endl;
        L << "dump_photons" << std::endl << "off" << std::endl;
        L << "qed" << std::endl << (p.qed ? "on" : "off") << std::endl;
        L << "balancing" << std::endl << "off" << std::endl;
        L << "output_mode\n" << (run.out_grid.binary ? 1 : 0) << '\n';
        L << "#------------------------------\n";
        const std::string polarization = str("polarization", "");
        L << "polarization = " << polarization << "\n";
        const double am = a0y * (a0y > a0z) + a0z * (a0z >= a0y);
        L << "P = " << am * am / 8 * ysigma * zsigma * 8.75e9 / 1e12 << " TW\n";
        L << "I = " << kPi * am * am * 8.75e9 / (c.lambda * c.lambda) << " W/cm^2\n";
        L << "W = " << (a0y * a0y + a0z * a0z) * xsigma * ysigma * zsigma * kPi * std::sqrt(kPi / 2) / 4 * c.lambda *
                           3.691e4 / 1e7
          << " J\n";
        L << "lambda = " << c.lambda * 1e4 << " um\n";
        L << "xsigma = " << xsigma / 2 / kPi * c.lambda * 1e4 << " um = c*" << xsigma / 2 / kPi * c.lambda / kLightCm * 1e15
          << " fs\n";
        L << "ysigma = " << ysigma / 2 / kPi * c.lambda * 1e4 << " um\n";
        L << "zsigma = " << zsigma / 2 / kPi * c.lambda * 1e4 << " um\n";
        L << "last_t_end = " << run.out_time.back() / 2 / kPi * c.lambda * 10 << " mm\n";
        L << "n_cr = " << kPi / (2.818e-13 * c.lambda * c.lambda) << " cm^{-3}\n";
        L << "ne/n_cr = " << ne / (kPi / (2.818e-13 * c.lambda * c.lambda)) << "\n";
        L << "#------------------------------\n";
        run.log = L.str();
    }
    return true;
}

void film_cells(const Run &run, const Film &f, double x0, double x1, bool column, std::vector<int32_t> &ijk,
                std::vector<double> &density)
{
    const qk_params &p = run.params;
    const double dx = p.dx, dy = p.dy, dz = p.dz;
    int i0 = (int)(x0 / dx), i1 = (int)(x1 / dx);
    if (column) {
        // append_for_moving_window: only the column i = nx - 3, when the film reaches it
        i0 = p.nx - 3;
        i1 = (i0 < p.nx - 2 && i1 >= p.nx - 2) ? p.nx - 2 : i0;
    } else {
        i0 = std::min(std::max(i0, 0), p.nx);
        i1 = std::min(std::max(i1, 0), p.nx);
    }
    const int jy0 = (int)(f.y0 / dy), jy1 = (int)(f.y1 / dy);
    const int kz0 = (int)(f.z0 / dz), kz1 = (int)(f.z1 / dz);
    const double film_ylen = double(jy1 - jy0 - 3);
    for (int i = i0; i < i1; ++i)
        for (int j = jy0 + 1; j < jy1 - 1; ++j) {
            double nes = f.ne_y1 * (j - jy0 - 1) / film_ylen + f.ne_y0 * (jy1 - 2 - j) / film_ylen;
            for (int k = kz0 + 1; k < kz1 - 1; ++k) {
                // the y ramps scale nes once per k, cumulatively, as the reference does
                if (f.gradwidth_y >= 0 && j - jy0 - 1 < f.gradwidth_y / dy) nes *= (j - jy0 - 1) / (f.gradwidth_y / dy);
                if (f.gradwidth_y >= 0 && jy1 - 2 - j < f.gradwidth_y / dy) nes *= (jy1 - 2 - j) / (f.gradwidth_y / dy);
                double n;
                if (f.gradwidth >= 0) n = i >= (int)((x0 + f.gradwidth) / dx) ? nes : nes * (i * dx - x0) / f.gradwidth;
                else n = i >= (int)((x0 - f.gradwidth) / dx) ? nes : nes * (1 + (i * dx - x0) / f.gradwidth);
                ijk.insert(ijk.end(), {i, j, k});
                density.push_back(n);
            }
        }
}

void film_particles(const Run &run, const Film &f, double cmr, ParticleSet &out)
{
    // film() (src/init.cpp:537-646) with T = 0: fill_cell_by_particles
    // (src/init.cpp:402-478, dsplmt = 0) on each of its cells
    std::vector<int32_t> ijk;
    std::vector<double> dens;
    film_cells(run, f, f.x0, f.x1, false, ijk, dens);
    const double ux0 = f.vx / std::sqrt(1 - f.vx * f.vx);
    const int bi = f.npic[0], bj = f.npic[1], bk = f.npic[2];
    for (size_t m = 0; m < dens.size(); ++m) {
        const int i = ijk[3 * m], j = ijk[3 * m + 1], k = ijk[3 * m + 2];
        double q0 = 1 / double(bi * bj * bk) * dens[m];
        if (cmr < 0) q0 = -q0;
        for (int ip = 0; ip < bi; ++ip)
            for (int jp = 0; jp < bj; ++jp)
                for (int kp = 0; kp < bk; ++kp) {
                    out.x.push_back(0.5 / bi + double(i) + double(ip) / double(bi));
                    out.y.push_back(0.5 / bj + double(j) + double(jp) / double(bj));
                    out.z.push_back(0.5 / bk + double(k) + double(kp) / double(bk));
                    out.ux.push_back(ux0);
                    out.uy.push_back(0);
                    out.uz.push_back(0);
                    out.q.push_back(q0);
                }
    }
}

double lin_interpolation(double x, const std::vector<double> &xs, const std::vector<double> &vs)
{
    if (xs.empty()) return 1.0;
    if (x < xs.front()) return vs.front();
    if (x > xs.back()) return vs.back();
    const size_t i = std::lower_bound(xs.begin(), xs.end(), x) - xs.begin();
    // x == xs[0]: the reference reads vs[-1] here; the limit from the right is vs[0]
    if (i == 0) return vs.front();
    const double left = vs[i - 1], right = vs[i];
    return left + (right - left) * (x - xs[i - 1]) / (xs[i] - xs[i - 1]);
}

namespace {

int ion_slot(const Run &run, double cmr)
{
    for (int n = 0; n < run.params.n_ion; ++n)
        if (run.params.ion_cmr[n] == cmr) return QK_ION0 + n;
    return -1;
}

}  // namespace

std::vector<CellFill> moving_window_fills(const Run &run, int nmw)
{
    const qk_params &p = run.params;
    std::vector<CellFill> out;
    // the seed column at i = nx - 3 (src/main.cpp:1756-1823)
    auto seed = [&](int species) {
        CellFill f{species, {}, {}, {run.npic[0], run.npic[1], run.npic[2]}, 0.0};
        double n;
        n = run.ne_ncr;
        n *= lin_interpolation((nmw - 1) * p.dx, run.ne_x_coords, run.ne_x_values);
        for (int j = 0; j < int(run.ylength / p.dy); j++)
            for (int k = 0; k < int(run.zlength / p.dz); k++) {
                const double zcell = k * p.dz, ycell = j * p.dy;
                const double ycenter = run.ylength / 2.0, zcenter = run.zlength / 2.0;
                const double zrel = zcell - zcenter, yrel = ycell - ycenter;
                const double r = std::sqrt(zrel * zrel + yrel * yrel);
                const double modifier = lin_interpolation(r, run.ne_r_coords, run.ne_r_values);
                if (modifier != 0.0) {
                    f.ijk.insert(f.ijk.end(), {p.nx - 3, j, k});
                    f.density.push_back(n * modifier);
                }
            }
        if (!f.density.empty()) out.push_back(std::move(f));
    };
    if (run.mwseed) seed(QK_ELECTRON);
    if (run.mwseed_ions) {
        const int s = ion_slot(run, run.mw_cmr);
        if (s >= 0) seed(s);
    }
    // every film's column, newest film first (src/main.cpp:1822-1833)
    for (const Film &f : run.films) {
        const double ux0 = f.vx / std::sqrt(1 - f.vx * f.vx);
        CellFill e{QK_ELECTRON, {}, {}, {f.npic[0], f.npic[1], f.npic[2]}, ux0};
        film_cells(run, f, f.x0 - p.dx * nmw, f.x1 - p.dx * nmw, true, e.ijk, e.density);
        if (e.density.empty()) continue;
        if (f.ions == 1) {
            const int s = ion_slot(run, f.cmr);
            if (s >= 0) out.push_back(CellFill{s, e.ijk, e.density, {f.npic[0], f.npic[1], f.npic[2]}, ux0});
        } else if (f.ions == 2) {
            out.push_back(CellFill{QK_POSITRON, e.ijk, e.density, {f.npic[0], f.npic[1], f.npic[2]}, ux0});
        }
        out.push_back(std::move(e));
    }
    return out;
}

}  // namespace qkd
