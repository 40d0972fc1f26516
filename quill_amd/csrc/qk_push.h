// qk_push.h -- per-particle pieces shared by the particle kernels of
// qk_particles.hip (QED off) and qk_qed.hip (QED on): step constants, the
// global-atomic current accumulator, particle records, re-binning, the
// QED-off advance of one charged particle and the fused diagnostics.
// Internal to the kernel translation units.
#pragma once

#include "qk_kernels.h"

namespace qk {
namespace {

constexpr int kBlock = 256;

template <typename real>
struct StepConst {
    real cmr_dt, ext_bz, inv_d[3], dt;
    double inv_cmr;
    int energy_mode;
    double *nonfinite;   // count of particles dropped for a non-finite momentum (FP64 sum, may be null)
};

// a particle whose momentum or displacement stopped being finite (corrupt
// input) is dropped and counted, so the step reports it
__device__ inline void count_nonfinite(double *nonfinite)
{
    if (nonfinite) atomicAdd(nonfinite, 1.0);
}

template <typename real>
StepConst<real> step_const(const Grid &g, const PushArgs &a)
{
    StepConst<real> sc;
    sc.cmr_dt = (real)(a.cmr * g.dt);
    sc.ext_bz = (real)a.external_bz;
    sc.inv_d[0] = (real)(1.0 / g.dx);
    sc.inv_d[1] = (real)(1.0 / g.dy);
    sc.inv_d[2] = (real)(1.0 / g.dz);
    sc.dt = (real)g.dt;
    sc.inv_cmr = a.cmr != 0 ? 1.0 / a.cmr : 0.0;
    sc.energy_mode = a.energy_mode;
    sc.nonfinite = a.nonfinite;
    return sc;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// every contribution is a global atomic
template <typename real>
struct GlobalAcc {
    real *jx, *jy, *jz;
    int sx, sy;
    int base;   // linear index of the particle's start cell
    int cur;    // linear index of the current piece's cell
    __device__ inline void set_cell(int wi, int wj, int wk) { cur = base + wi * sx + wj * sy + wk; }
    __device__ inline void operator()(int comp, int di, int dj, int dk, real v)
    {
        const int id = cur + di * sx + dj * sy + dk;
        real *a = comp == 0 ? jx : (comp == 1 ? jy : jz);
        unsafeAtomicAdd(a + id, v);
    }
};

// Sum four per-thread values over the workgroup and add them to dst[0..4)
// with one atomic each (skipped when the count and fallback sums are zero).
// Every thread of the workgroup must call it; calls may follow each other.
__device__ inline void block_accumulate(double *dst, double v0, double v1, double v2, double v3)
{
    __shared__ double part[4][16];   // up to 1024 threads
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    v2 = wave_sum(v2);
    v3 = wave_sum(v3);
    const int w = threadIdx.x >> 6;
    __syncthreads();   // a previous call may still be reading part
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = v0;
        part[1][w] = v1;
        part[2][w] = v2;
        part[3][w] = v3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[4] = {0, 0, 0, 0};
        for (int m = 0; m < (int)(blockDim.x >> 6); ++m)
            for (int a = 0; a < 4; ++a) s[a] += part[a][m];
        if (s[2] != 0 || s[3] != 0)
            for (int a = 0; a < 4; ++a) atomicAdd(dst + a, s[a]);
    }
}

struct DiagAcc {
    double q = 0, e = 0, n = 0, fallback = 0;
};

// one particle in registers
template <typename real>
struct PRec {
    int cell;      // packed cell, negative: free slot
    real o[3];     // offset in the cell
    real u[3];     // u = p / m c
    real q;        // macro charge
    real chi;      // QED parameter (QED kernels only)
};

template <typename real>
__device__ inline void load_rec(const Species<real> &s, int64_t p, PRec<real> &r)
{
    r.cell = s.cell[p];
    r.o[0] = s.ox[p]; r.o[1] = s.oy[p]; r.o[2] = s.oz[p];
    r.u[0] = s.ux[p]; r.u[1] = s.uy[p]; r.u[2] = s.uz[p];
    r.q = s.q[p];
}

// the same with a 32-bit slot: each array is its base (scalar registers)
// plus one 32-bit byte offset shared by all of them
template <typename T>
__device__ inline T &at32(T *base, uint32_t byte_off)
{
    return *(T *)((char *)base + byte_off);
}

// The records stream through once per step: non-temporal loads and stores
// keep them from evicting the field arrays from the caches, which the field
// kernels then find there (the field phase measured 25 % faster).
template <typename T>
__device__ inline T ld_nt(T *base, uint32_t byte_off)
{
    return __builtin_nontemporal_load(&at32(base, byte_off));
}

template <typename T>
__device__ inline void st_nt(T *base, uint32_t byte_off, T v)
{
    __builtin_nontemporal_store(v, &at32(base, byte_off));
}

template <typename real>
__device__ inline void load_rec32(const Species<real> &s, uint32_t p, PRec<real> &r)
{
    const uint32_t o4 = p * 4u, orl = p * (uint32_t)sizeof(real);
    r.cell = ld_nt(s.cell, o4);
    r.o[0] = ld_nt(s.ox, orl); r.o[1] = ld_nt(s.oy, orl); r.o[2] = ld_nt(s.oz, orl);
    r.u[0] = ld_nt(s.ux, orl); r.u[1] = ld_nt(s.uy, orl); r.u[2] = ld_nt(s.uz, orl);
    r.q = ld_nt(s.q, orl);
}

template <typename real>
__device__ inline void store_rec32(const Species<real> &s, uint32_t p, const PRec<real> &r)
{
    const uint32_t o4 = p * 4u, orl = p * (uint32_t)sizeof(real);
    st_nt(s.cell, o4, r.cell);
    st_nt(s.ox, orl, r.o[0]); st_nt(s.oy, orl, r.o[1]); st_nt(s.oz, orl, r.o[2]);
    st_nt(s.ux, orl, r.u[0]); st_nt(s.uy, orl, r.u[1]); st_nt(s.uz, orl, r.u[2]);
}

// write back what a push changes (the charge does not change)
template <typename real>
__device__ inline void store_rec(const Species<real> &s, int64_t p, const PRec<real> &r)
{
    s.cell[p] = r.cell;
    s.ox[p] = r.o[0]; s.oy[p] = r.o[1]; s.oz[p] = r.o[2];
    s.ux[p] = r.u[0]; s.uy[p] = r.u[1]; s.uz[p] = r.u[2];
}

__device__ inline int cell_base(const Grid &g, int c) { return cell_i(c) * g.sx + cell_j(c) * g.sy + cell_k(c); }

// Move the cell (i, j, k) by floor(pos) so that pos returns to [0, 1).
template <typename real>
__device__ inline void rebin(int &i, int &j, int &k, real pos[3])
{
    int ci[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const real fl = floor(pos[m]);
        ci[m] = (int)fl;
        pos[m] -= fl;
        if (pos[m] >= real(1)) {   // rounding of a tiny negative position
            pos[m] = real(0);
            ci[m] += 1;
        }
    }
    i += ci[0];
    j += ci[1];
    k += ci[2];
}

// gamma - 1 as u^2 / (gamma + 1): no cancellation for slow particles
template <typename real>
__device__ inline double kinetic(const real u[3], real gam)
{
    return (double)((u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * qk_rcp(gam + real(1)));
}

// One charged particle with QED off, given E and Be interpolated to it
// (interp_corners of the field nodes around it: e_to_particle /
// b_to_particle, src/deposition.cpp:431-471), add external Bz, push (Vay or Boris), deposit
// the current of the straight path u dt / gamma split at cell faces, move,
// and mark the slot free if the new cell is not is_inside (place(),
// src/plist.cpp:7-40).  Diagnostics of compute_N / compute_energy
// (src/fout.cpp:41-138) come from the post-push state.  dep(i, j, k, base,
// q, pos, d, da) performs the deposit and advances pos.
template <typename real, int PUSHER, class Dep>
__device__ inline void advance_particle_eb(const Grid &g, const real (&ei)[3], const real (&bi)[3], PRec<real> &r,
                                           const StepConst<real> &sc, Dep &&dep, DiagAcc &da)
{
    const int c = r.cell;
    if (c < 0) return;
    int i = cell_i(c), j = cell_j(c), k = cell_k(c);
    // p_boundary (src/plist.cpp:73-102) erases the layers 0, n-2, n-1 before
    // the push; only a moving-window shift leaves particles there (place()
    // keeps is_inside cells only)
    if (!is_inside(g, i, j, k)) {
        r.cell = -1;
        return;
    }
    // local copies stay in registers (the deposit indexes them by axis)
    real pos[3] = {r.o[0], r.o[1], r.o[2]};
    real u[3] = {r.u[0], r.u[1], r.u[2]};
    const int base = i * g.sx + j * g.sy + k;
    real e[3] = {ei[0], ei[1], ei[2]}, b[3] = {bi[0], bi[1], bi[2]};
    b[2] += sc.ext_bz;
    real gam = qk_sqrt(real(1) + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    push<PUSHER>(sc.cmr_dt, e, b, u, gam);
    // displacement in cells (containers.cpp:58-67, advance.cpp:281-284)
    const real a = sc.dt * qk_rcp(gam);
    const real d[3] = {u[0] * a * sc.inv_d[0], u[1] * a * sc.inv_d[1], u[2] * a * sc.inv_d[2]};
    // |d| < dt/dx < 1 for any finite momentum; a non-finite one (corrupt
    // input) would index outside the stencil, so that particle is dropped
    r.u[0] = u[0];
    r.u[1] = u[1];
    r.u[2] = u[2];
    if (!(fabs(d[0]) < real(1) && fabs(d[1]) < real(1) && fabs(d[2]) < real(1))) {
        r.cell = -1;
        count_nonfinite(sc.nonfinite);
        return;
    }
    dep(i, j, k, base, r.q, pos, d, da);
    rebin(i, j, k, pos);
    r.o[0] = pos[0];
    r.o[1] = pos[1];
    r.o[2] = pos[2];
    if (is_inside(g, i, j, k)) {
        r.cell = pack_cell(i, j, k);
        const double ke = kinetic(u, gam);
        da.q += (double)r.q;
        da.n += 1;
        da.e += sc.energy_mode == 2 ? ke * (double)r.q * sc.inv_cmr : ke * (double)r.q;
    } else {
        r.cell = -1;
    }
}

// the same from the field nodes around the particle (load_corners)
template <typename real, int PUSHER, class Dep>
__device__ inline void advance_particle(const Grid &g, const Corners<real> &nodes, PRec<real> &r,
                                        const StepConst<real> &sc, Dep &&dep, DiagAcc &da)
{
    if (r.cell < 0) return;
    real e[3], b[3];
    interp_corners(nodes, r.o[0], r.o[1], r.o[2], e, b);
    advance_particle_eb<real, PUSHER>(g, e, b, r, sc, dep, da);
}

}  // namespace
}  // namespace qk
