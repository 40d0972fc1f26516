// qk_kernels.h -- host-callable launchers of the gfx950 kernels.
// Every launcher enqueues on the given stream and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qk_device.h"

namespace qk {

// solver coefficients (src/maxwell.cpp:24-33, 238-249); FDTD uses a = 0, b = 1
struct SolverCoef {
    double a[3], b[3];
};

// per-species accumulators of the fused diagnostics, all in FP64:
// [0] sum q, [1] sum energy, [2] count
constexpr int kDiagPerSpecies = 3;

struct PushArgs {
    int solver_unused;
    int pusher;              // 0 Vay, 1 Boris
    int kind;                // 0 lepton / ion (charged), 1 photon
    int deposit;             // QK_DEPOSIT_*
    double cmr;              // charge-to-mass ratio
    double external_bz;
    int64_t n;               // slots to visit
    double *diag;            // kDiagPerSpecies doubles, accumulated atomically
    int energy_mode;         // 0: (g-1) q ; 1: g q (photons) ; 2: (g-1) q / cmr (ions)
};

// clear n doubles with a kernel (kept out of hipMemsetAsync so the step can be
// captured in a HIP graph)
hipError_t launch_zero(hipStream_t st, double *p, size_t n);

template <typename real>
hipError_t launch_push(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                       const Species<real> &s, const PushArgs &a);

template <typename real>
hipError_t launch_maxwell_b(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, const Fields<real> &f);
template <typename real>
hipError_t launch_maxwell_e(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, const Fields<real> &f,
                            const Current<real> &j, const Current<real> &j_next);
template <typename real>
hipError_t launch_interpolate_be(hipStream_t st, const Grid &g, const Fields<real> &f);
template <typename real>
hipError_t launch_init_boundaries(hipStream_t st, int solver, const Grid &g, const Fields<real> &f);
template <typename real>
hipError_t launch_field_energy(hipStream_t st, const Grid &g, const Fields<real> &f, double *acc);
template <typename real>
hipError_t launch_laser_cos(hipStream_t st, const Grid &g, const Fields<real> &f, const double *L);
template <typename real>
hipError_t launch_load_uniform(hipStream_t st, const Grid &g, const Species<real> &s, int64_t first, int ppc,
                               double q, double sigma, uint64_t seed, int ilo, int ihi);
template <typename real>
hipError_t launch_compute_rho(hipStream_t st, const Grid &g, const Species<real> &s, int64_t n, double *rho, int slot);

}  // namespace qk
