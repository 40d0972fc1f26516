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
// [0] sum q, [1] sum energy, [2] live count, [3] particles deposited through
// the global fallback of the tiled kernel
constexpr int kDiagPerSpecies = 4;

// most particles per workgroup of the tiled kernel (a chunk never spans two
// tiles; a tile is cut into equal chunks).  Twice the mean tile population at
// 8 per cell, so a uniform plasma's tile is one chunk: 4096 measured 15 %
// slower on the push, 2048 30 %, from staging tiles more than once
constexpr int kChunk = 8192;

struct PushArgs {
    int pusher;              // 0 Vay, 1 Boris
    int kind;                // 0 charged (lepton / ion), 1 photon
    double cmr;              // charge-to-mass ratio
    double external_bz;
    int64_t first, n;        // slots [first, first + n) (global kernel)
    double *diag;            // kDiagPerSpecies doubles, accumulated atomically
    int energy_mode;         // 0: (g-1) q ; 1: g q (photons) ; 2: (g-1) q / cmr (ions)
    double *nonfinite;       // FP64 count of particles dropped for a non-finite momentum (may be null)
    // tiled kernel: work list of (tile, first slot, count) chunks
    const int4 *chunks;
    const int *nchunks;      // device counter
    int max_chunks;          // launch size (upper bound of *nchunks)
    TileGeom tg;
};

// clear n doubles / ints with a kernel (kept out of hipMemsetAsync so the
// step can be captured in a HIP graph)
hipError_t launch_zero(hipStream_t st, double *p, size_t n);
hipError_t launch_fill_int(hipStream_t st, int *p, size_t n, int v);

// one thread per particle, every current contribution a global atomic
template <typename real>
hipError_t launch_push_global(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                              const Species<real> &s, const PushArgs &a);
// one workgroup per chunk of a cell tile: the current is accumulated in an
// LDS window around the tile and flushed to global memory once
template <typename real>
hipError_t launch_push_tiled(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                             const Species<real> &s, const PushArgs &a);

// ---- QED -------------------------------------------------------------------

// constants of one QED step
struct QedArgs {
    double e_s;          // Sauter field (mc omega / e units)
    uint64_t seed;       // run seed of the counter-based streams
    uint32_t step;       // global step index (counter word)
};

// a species buffer that kernels append new particles to: slot = atomic
// increment of *count; an append at or beyond cap is dropped and sets
// *overflow (the host reserves enough capacity before each step)
template <typename real>
struct AppendTarget {
    Species<real> s;
    int *count;
    int cap;
    int *overflow;
    double *diag;        // kDiagPerSpecies accumulators of the target species
};

// electrons / positrons with photon emission (src/advance.cpp:253-359);
// every current contribution a global atomic; emitted photons go to ph
template <typename real>
hipError_t launch_push_qed_lepton(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                                  const Species<real> &s, const PushArgs &a, const QedArgs &q,
                                  const AppendTarget<real> &ph);
// photons with Breit-Wheeler pair creation (src/advance.cpp:155-252); the
// pairs go to el / po
template <typename real>
hipError_t launch_photon_qed(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                             const Species<real> &s, const PushArgs &a, const QedArgs &q,
                             const AppendTarget<real> &el, const AppendTarget<real> &po);
// birth_from_vacuum (src/mc.cpp:142-196): one draw per cell 2 .. n-3
template <typename real>
hipError_t launch_vacuum_pairs(hipStream_t st, const Grid &g, const Fields<real> &f, int pusher, double q_vac,
                               const QedArgs &q, const AppendTarget<real> &el, const AppendTarget<real> &po);

// ---- macro-particle merging (pmerging, src/mc.cpp:202-462) ---------------

// acc[0] += sum |q|, acc[2] += live count over slots [0, n) (acc: 4 doubles)
template <typename real>
hipError_t launch_charge_sum(hipStream_t st, const Species<real> &s, int64_t n, double *acc);
// mode 1 ("ti"): delete each live particle with probability p, acc[0] +=
// deleted |q|; mode 2 ("nl"): delete with p / (1 + (2 |q| / av_q)^4), divide
// the survivors' q by one minus that probability, acc[0] += their new |q|.
// Draws: Philox keyed by (seed, step, slot, tag).
template <typename real>
hipError_t launch_merge_delete(hipStream_t st, const Species<real> &s, int64_t n, int mode, double p, double av_q,
                               uint64_t seed, uint32_t step, uint32_t tag, double *acc);
// q *= a for every live particle
template <typename real>
hipError_t launch_scale_charge(hipStream_t st, const Species<real> &s, int64_t n, double a);

// ---- moving window (src/advance.cpp:389-454) --------------------------------

// E, B and Be move one plane toward -x (F[i] = F[i + 1]); the plane nx - 1 is zeroed
template <typename real>
hipError_t launch_shift_fields(hipStream_t st, const Grid &g, const Fields<real> &f);
// packed cells of slots [0, n): plane i = 0 is erased, the rest move to i - 1
hipError_t launch_shift_particles(hipStream_t st, int *cell, int64_t n);
// *out = max(*out, max over planes [ilo, ihi) of E^2 + B^2) as the bits of a
// non-negative double (get_max_w, src/low-level.cpp:512-532)
template <typename real>
hipError_t launch_max_w(hipStream_t st, const Grid &g, const Fields<real> &f, int ilo, int ihi,
                        unsigned long long *out);
// fill_cell_by_particles (src/init.cpp:402-478) with T = 0 on ncells listed
// cells (i j k triples), npic[0] x npic[1] x npic[2] particles each, written to
// slots first, first + 1, ...; ids idbase[m] + particle index when s.id is set
template <typename real>
hipError_t launch_fill_cells(hipStream_t st, const Species<real> &s, int64_t first, int64_t ncells, const int *cells,
                             const double *density, const int *idbase, const int npic[3], double sign, double ux0);

// ---- adjoint of the deterministic step --------------------------------------

// the transposes of fadvance on the FP64 adjoint arrays lam = lE[3], lB[3],
// lBe[3], lJ[3]: interpolate_be, B half step, E step (which sets lJ), B half
// step, in that order
hipError_t launch_adj_fields(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, double *const lam[12]);
// the transpose of interpolate_be alone (for Be of the first step)
hipError_t launch_adj_interp_be(hipStream_t st, const Grid &g, double *const lam[12]);
// the tangent of one species' push at the current state s, fields f (E, Be)
// and their tangents df: tpart (x y z ux uy uz q, FP64) advances one step,
// dJ (FP64) receives the tangent of the current
template <typename real>
hipError_t launch_tlm_particles(hipStream_t st, const Grid &g, const Fields<real> &f, const Fields<double> &df,
                                const Species<real> &s, int64_t n, const PushArgs &a, double *const dJ[3],
                                double *const tpart[7]);
// the adjoint of one species' push from its recorded state s and fields f
// (E, Be) of the step: lpart = lx[3], lu[3], lq (FP64, per slot) go from after
// the step to before it; lE and lBe receive the gather's node adjoints
template <typename real>
hipError_t launch_adj_particles(hipStream_t st, const Grid &g, const Fields<real> &f, const Species<real> &s,
                                int64_t n, const PushArgs &a, double *const lam[12], double *const lpart[7]);

// the field-energy sum is spread over this many FP64 slots (one atomic per
// workgroup lands on slot blockIdx mod kEnergySlots) and summed on the host
constexpr int kEnergySlots = 256;

// half step of B; with energy != null also accumulates sum(E^2 + B^2) of the
// incoming fields into energy[0 .. kEnergySlots) (FP64)
template <typename real>
hipError_t launch_maxwell_b(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, const Fields<real> &f,
                            double *energy);
template <typename real>
hipError_t launch_maxwell_e(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, const Fields<real> &f,
                            const Current<real> &j, const Current<real> &j_next);
template <typename real>
hipError_t launch_interpolate_be(hipStream_t st, const Grid &g, const Fields<real> &f);
template <typename real>
hipError_t launch_init_boundaries(hipStream_t st, int solver, const Grid &g, const Fields<real> &f);
template <typename real>
hipError_t launch_laser_cos(hipStream_t st, const Grid &g, const Fields<real> &f, const double *L);

// the pulse of f_init_cos (src/init.cpp:6-210): envelope half lengths,
// centre, phase and propagation frame
struct LaserFrame {
    double a0y, a0z, xs, ys, zs, x0, phase, xt, yt, zt, y12, z12;
    double r0[3], ey[3], ez[3];
    double y0, z0;
    int b_sign, append;
};
// frame of the 16 f_init_cos parameters on grid g; false if degenerate
bool laser_frame(const Grid &g, const double *P, LaserFrame &L);

// laser sources entering through one x plane: corrections of one sub-step
constexpr int kMaxSources = 4;
constexpr int kInjGap = 4;   // empty planes between the two slabs of the stack
struct InjectArgs {
    int nsrc;
    LaserFrame src[kMaxSources];
    double shift[kMaxSources];   // distance each pulse has travelled at this sub-step
    int ishift[kMaxSources];     // moving-window shifts since the source was added
    int side;                    // 0: total field at i >= plane; 1: at i < plane
    int plane;
    int W;                       // planes per slab
    int main_lo;                 // grid plane of slab plane 0
    int b_field;                 // 1: B half step (input E_inc), 0: E step (input B_inc)
};
// the stack grid gs (2 W + kInjGap planes) holds the incident field and the
// correction passes; j0, j1 zero current buffers of the stack's size
template <typename real>
hipError_t launch_inject(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, const Fields<real> &f,
                         const Grid &gs, const Fields<real> &stack, const Current<real> &j0, const Current<real> &j1,
                         const InjectArgs &a);
template <typename real>
hipError_t launch_load_uniform(hipStream_t st, const Grid &g, const Species<real> &s, int64_t first, int ppc,
                               double q, double sigma, uint64_t seed, int ilo, int ihi);
template <typename real>
hipError_t launch_compute_rho(hipStream_t st, const Grid &g, const Species<real> &s, int64_t n, double *rho, int slot);

// ---- cell-tile counting sort ----------------------------------------------

// scratch for the exclusive scan of n ints: at most scan_scratch_ints(n) ints
size_t scan_scratch_ints(size_t n);
// out[i] = sum of in[0..i) for i in [0, n]; out has n + 1 entries
hipError_t launch_exclusive_scan(hipStream_t st, const int *in, int *out, int n, int *scratch);

struct SortBuffers {
    int *counts;      // nkeys + 1: particles per key (first sort); per tile, those that stay (re-sort)
    int *imm;         // nkeys + 1: re-sort: per tile, those that arrive, then all, then the arrivals' counters
    int *offsets;     // nkeys + 1: first slot per key (first sort)
    int *tile_start;  // ntiles + 1
    int *tile_mult;   // ntiles: inverse spread multiplier of the tile's storage order
    int *chunk_n;     // ntiles + 1
    int *chunk_off;   // ntiles + 1
    int4 *chunks;     // max_chunks
    int *scratch;     // scan scratch
    int max_chunks;
    const int *old_start;   // ntiles + 1: the tiles of the source's slots [0, n_sorted), or null
    int64_t n_sorted;
};

// sort the live particles of src[0, n) into dst by tile: the first sort by
// (tile, cell) with each tile in its spread storage order (consecutive slots
// hold particles of distinct cells), later sorts keeping each tile's order
// (old_start: the tiles of the previous sort); writes tile_start, the chunk
// work list and its length (chunk_off[ntiles])
template <typename real>
hipError_t launch_sort(hipStream_t st, const TileGeom &tg, const Species<real> &src, const Species<real> &dst,
                       int64_t n, const SortBuffers &b);

}  // namespace qk
