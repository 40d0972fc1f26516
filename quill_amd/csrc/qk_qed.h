// qk_qed.h -- device functions of the QED step on gfx950: the quantum
// parameter chi, the photon-emission and Breit-Wheeler pair spectra
// (src/mc.cpp:26-97), the Schwinger rate (src/mc.cpp:99-116) and the
// counter-based random numbers that replace get_rand (src/mc.cpp:12-23).
// The spectra are closed forms evaluated per particle in the particle
// precision; the acceptance test compares against a 53-bit uniform so that
// event probabilities far below 2^-24 keep their value.
#pragma once

#include "qk_device.h"

namespace qk {

// chi = |gamma E + u x B|_perp / e_s (src/mc.cpp:26-38)
template <typename real>
__device__ inline real qed_chi(real inv_es, const real e[3], const real b[3], const real u[3], real g)
{
    real a = g * e[0] + u[1] * b[2] - u[2] * b[1];
    real s2 = a * a;
    a = g * e[1] - u[0] * b[2] + u[2] * b[0];
    s2 += a * a;
    a = g * e[2] + u[0] * b[1] - u[1] * b[0];
    s2 += a * a;
    a = u[0] * e[0] + u[1] * e[1] + u[2] * e[2];
    s2 -= a * a;
    return s2 > real(0) ? sqrt(s2) * inv_es : real(0);
}

__device__ inline float qk_cbrt(float x) { return cbrtf(x); }
__device__ inline double qk_cbrt(double x) { return cbrt(x); }

// the bracket shared by both spectra (src/mc.cpp:55-59, 91-95):
// (kap + 0.80049)^(-3/4) - c * (kap + 0.70861 (1 - 0.65 kap / (1 + kap^2)))^(1/4)
template <typename real>
__device__ inline real qed_bracket(real kap, real c)
{
    const real k1 = kap + real(0.80049);
    const real t1 = real(1) / sqrt(sqrt(k1 * k1 * k1));
    const real t2 = c * sqrt(sqrt(kap + real(0.70861) * (real(1) - real(0.65) * kap / (real(1) + kap * kap))));
    return t1 - t2;
}

// photon emission spectrum dW/dr (src/mc.cpp:40-74).  For chi <= 0.133 the
// fraction r is rescaled in place by r_max = 1 / (1 + 2 / (15 chi)) and the
// rescaled value is the one the photon gets.
template <typename real>
__device__ inline real qed_rate_w(real e_s, real g, real &r, real chi)
{
    if ((real(1) - r) * g < real(1) || r == real(0) || chi == real(0)) return real(0);
    real scale = real(1);
    if (chi <= real(0.133)) {
        const real r_max = real(1) / (real(1) + real(2) / (real(15) * chi));
        r = r_max * r;
        scale = r_max;
    }
    const real d = r / ((real(1) - r) * chi);
    const real a = qk_cbrt(d);
    const real c = (real(2) * (real(1) - r) + r * r) * chi * a / r;
    return -scale * real(0.0020591) * e_s / (g * g) * exp(real(-2) * d / real(3)) * qed_bracket(a * a, c);
}

// Breit-Wheeler pair spectrum (src/mc.cpp:76-97)
template <typename real>
__device__ inline real qed_rate_tilde_w(real e_s, real g, real r, real chi)
{
    if (chi == real(0) || r * g < real(1) || (real(1) - r) * g < real(1)) return real(0);
    const real d = real(1) / (r * (real(1) - r) * chi);
    const real a = qk_cbrt(d);
    const real c = (real(2) * r * (real(1) - r) - real(1)) * chi * a;
    return real(0.0020591) * e_s / (g * g) * exp(real(-2) * d / real(3)) * qed_bracket(a * a, c);
}

// Schwinger pair rate per unit volume and time (src/mc.cpp:99-116), FP64
__device__ inline double qed_vacuum_rate(double e_s, const double e[3], const double b[3])
{
    const double pi = 3.14159265358979323846;
    const double f = 0.5 * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2] - b[0] * b[0] - b[1] * b[1] - b[2] * b[2]);
    const double gg = e[0] * b[0] + e[1] * b[1] + e[2] * b[2];
    const double a = sqrt(f * f + gg * gg);
    const double xi = sqrt(a + f), th = sqrt(fmax(a - f, 0.0));
    if (xi == 0) return 0;
    if (th != 0) return e_s * e_s / (4 * pi) * xi * th / tanh(pi * th / xi) * exp(-pi * e_s / xi);
    return e_s * e_s / (4 * pi) * xi * xi / pi * exp(-pi * e_s / xi);
}

// Random numbers of one event draw: Philox4x32-10 keyed by the run seed,
// counter (slot, step, stream tag); r in (0, 1) from 24 bits (the spectrum
// fraction), acc in [0, 1) from 53 bits (the acceptance test).
struct QedDraw {
    double acc;
    float r;
};

__device__ inline QedDraw qed_draw(uint64_t seed, uint32_t step, uint32_t slot, uint32_t tag)
{
    const u32x4 v = philox(u32x4{{slot, step, tag, 0x0ED5A17Du}}, (uint32_t)seed, (uint32_t)(seed >> 32));
    QedDraw d;
    d.r = (float)u24_open0(v.v[0]);
    d.acc = ((double)(v.v[1] >> 11) * 4294967296.0 + (double)v.v[2]) * (1.0 / 9007199254740992.0);
    return d;
}

}  // namespace qk
