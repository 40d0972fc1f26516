// qk_particles.hip -- particle kernels of the step on gfx950: the fused
// gather + push + charge-conserving deposit + re-binning kernels (the particle
// loop of src/advance.cpp:90-387) in two forms, photon transport, the device
// plasma loader and the node-charge deposit used for the Gauss-law check.
//
//   k_push_global  one thread per particle, each current contribution a global
//                  atomic (the plain method, used for the FP64 comparison)
//   k_push_tiled   one workgroup per chunk of an 8x8x8 cell tile; particles are
//                  sorted by (tile, cell), so the wavefront's gathers hit a few
//                  cache lines and its deposits land in an LDS window of the
//                  tile plus a halo, flushed once with one global atomic per
//                  nonzero window entry.  A particle that drifted out of the
//                  window since the last sort deposits through global atomics.
#include "qk_kernels.h"

namespace qk {

namespace {

constexpr int kBlock = 256;
// the window covers start cells within kHalo cells of the tile; a deposit
// reaches one cell below and two above the start cell
constexpr int kHalo = 1;
constexpr int kWin = kTile + 2 * kHalo + 3;
constexpr int kWin3 = kWin * kWin * kWin;

template <typename real>
struct StepConst {
    real cmr_dt, ext_bz, inv_d[3], dt;
    double inv_cmr;
    int energy_mode;
};

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

// contributions go to the workgroup's LDS window [3][kWin][kWin][kWin].  The
// window is FP64 for both particle precisions: on gfx950 an FP64 LDS atomic
// add sustains about twenty times the rate of an FP32 one, and the tile's
// partial sums keep more bits before the single FP32/FP64 flush.
template <typename real>
struct LdsAcc {
    double *win;
    int base;   // window index of the particle's start cell
    int cur;
    __device__ inline void set_cell(int wi, int wj, int wk) { cur = base + (wi * kWin + wj) * kWin + wk; }
    __device__ inline void operator()(int comp, int di, int dj, int dk, real v)
    {
        atomicAdd(win + comp * kWin3 + cur + (di * kWin + dj) * kWin + dk, (double)v);
    }
};

__device__ inline void block_accumulate(double *dst, double v0, double v1, double v2, double v3)
{
    __shared__ double part[4][kBlock / 64];
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    v2 = wave_sum(v2);
    v3 = wave_sum(v3);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = v0;
        part[1][w] = v1;
        part[2][w] = v2;
        part[3][w] = v3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[4] = {0, 0, 0, 0};
        for (int m = 0; m < kBlock / 64; ++m)
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
};

template <typename real>
__device__ inline void load_rec(const Species<real> &s, int64_t p, PRec<real> &r)
{
    r.cell = s.cell[p];
    r.o[0] = s.ox[p]; r.o[1] = s.oy[p]; r.o[2] = s.oz[p];
    r.u[0] = s.ux[p]; r.u[1] = s.uy[p]; r.u[2] = s.uz[p];
    r.q = s.q[p];
}

// write back what a push changes (the charge does not change)
template <typename real>
__device__ inline void store_rec(const Species<real> &s, int64_t p, const PRec<real> &r)
{
    s.cell[p] = r.cell;
    s.ox[p] = r.o[0]; s.oy[p] = r.o[1]; s.oz[p] = r.o[2];
    s.ux[p] = r.u[0]; s.uy[p] = r.u[1]; s.uz[p] = r.u[2];
}

// One charged particle with QED off: gather E and Be at the particle, add
// external Bz, push (Vay or Boris), deposit the current of the straight path
// u dt / gamma split at cell faces, move, and mark the slot free if the new
// cell is not is_inside (place(), src/plist.cpp:7-40).  Diagnostics of
// compute_N / compute_energy (src/fout.cpp:41-138) come from the post-push
// state.  dep(i, j, k, base, q, pos, d, da) performs the deposit and
// advances pos.
template <typename real, int PUSHER, class Dep>
__device__ inline void advance_particle(const Grid &g, const Fields<real> &f, PRec<real> &r, const StepConst<real> &sc,
                                        Dep &&dep, DiagAcc &da)
{
    const int c = r.cell;
    if (c < 0) return;
    const int i = cell_i(c), j = cell_j(c), k = cell_k(c);
    // local copies stay in registers (the deposit indexes them by axis)
    real pos[3] = {r.o[0], r.o[1], r.o[2]};
    real u[3] = {r.u[0], r.u[1], r.u[2]};
    const int base = i * g.sx + j * g.sy + k;
    real e[3], b[3];
    gather3(f.ex, f.ey, f.ez, base, g.sx, g.sy, pos[0], pos[1], pos[2], e);
    gather3(f.bex, f.bey, f.bez, base, g.sx, g.sy, pos[0], pos[1], pos[2], b);
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
        return;
    }
    dep(i, j, k, base, r.q, pos, d, da);
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
    r.o[0] = pos[0];
    r.o[1] = pos[1];
    r.o[2] = pos[2];
    const int ni = i + ci[0], nj = j + ci[1], nk = k + ci[2];
    if (is_inside(g, ni, nj, nk)) {
        r.cell = pack_cell(ni, nj, nk);
        const double gg = (double)gam;
        da.q += (double)r.q;
        da.n += 1;
        da.e += sc.energy_mode == 2 ? (gg - 1) * (double)r.q * sc.inv_cmr : (gg - 1) * (double)r.q;
    } else {
        r.cell = -1;
    }
}

template <typename real, int PUSHER>
__global__ void __launch_bounds__(kBlock) k_push_global(Grid g, Fields<real> f, Current<real> cur, Species<real> s,
                                                        int64_t first, int64_t n, StepConst<real> sc, double *diag)
{
    const int64_t p = first + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    DiagAcc da;
    if (p < first + n) {
        auto dep = [&](int, int, int, int base, real q, real pos[3], const real d[3], DiagAcc &) {
            GlobalAcc<real> acc{cur.jx, cur.jy, cur.jz, g.sx, g.sy, base, base};
            deposit_path<real>(acc, q, pos, d);
        };
        PRec<real> r;
        load_rec(s, p, r);
        if (r.cell >= 0) {
            advance_particle<real, PUSHER>(g, f, r, sc, dep, da);
            store_rec(s, p, r);
        }
    }
    block_accumulate(diag, da.q, da.e, da.n, da.fallback);
}

// Workgroup order inside a block of kBlock particles: lane t works on
// particle (kPerm * t) mod kBlock of the block.  Sorted particles share a cell
// in runs of about ppc; the odd stride spreads a wavefront over distinct cells
// (so its LDS atomics hit distinct addresses) and distinct LDS banks, while
// the workgroup as a whole still reads and writes whole cache lines.
constexpr int kPerm = 9;

template <typename real, int PUSHER>
__global__ void __launch_bounds__(kBlock) k_push_tiled(Grid g, Fields<real> f, Current<real> cur, Species<real> s,
                                                       StepConst<real> sc, const int4 *__restrict__ chunks,
                                                       const int *__restrict__ nchunks, TileGeom tg, double *diag)
{
    __shared__ double win[3 * kWin3];
    if ((int)blockIdx.x >= *nchunks) return;   // whole workgroup leaves together
    const int4 ch = chunks[blockIdx.x];
    const int tile = ch.x;
    const int tk = tile % tg.ntz, t2 = tile / tg.ntz;
    const int tj = t2 % tg.nty, ti = t2 / tg.nty;
    const int i0 = ti * kTile, j0 = tj * kTile, k0 = tk * kTile;
    for (int w = threadIdx.x; w < 3 * kWin3; w += kBlock) win[w] = 0.0;
    __syncthreads();

    DiagAcc da;
    auto dep = [&](int i, int j, int k, int base, real q, real pos[3], const real d[3], DiagAcc &acc_d) {
        const int ri = i - i0 + kHalo, rj = j - j0 + kHalo, rk = k - k0 + kHalo;
        const int lim = kTile + 2 * kHalo;
        if ((unsigned)ri < (unsigned)lim && (unsigned)rj < (unsigned)lim && (unsigned)rk < (unsigned)lim) {
            // window index of the start cell: one extra layer below for w = -1
            LdsAcc<real> acc{win, ((ri + 1) * kWin + (rj + 1)) * kWin + (rk + 1), 0};
            deposit_path<real>(acc, q, pos, d);
        } else {
            GlobalAcc<real> acc{cur.jx, cur.jy, cur.jz, g.sx, g.sy, base, base};
            deposit_path<real>(acc, q, pos, d);
            acc_d.fallback += 1;
        }
    };
    const int64_t end = (int64_t)ch.y + ch.z;
    const int sidx = (kPerm * (int)threadIdx.x) & (kBlock - 1);
    for (int64_t b0 = ch.y; b0 < end; b0 += kBlock) {
        const int64_t p = b0 + sidx;
        if (p < end) {
            PRec<real> r;
            load_rec(s, p, r);
            if (r.cell >= 0) {
                advance_particle<real, PUSHER>(g, f, r, sc, dep, da);
                store_rec(s, p, r);
            }
        }
    }
    __syncthreads();

    // flush the window: one global atomic per nonzero entry inside the grid
    const int wi0 = i0 - kHalo - 1, wj0 = j0 - kHalo - 1, wk0 = k0 - kHalo - 1;
    for (int w = threadIdx.x; w < kWin3; w += kBlock) {
        const int wk = w % kWin, t = w / kWin;
        const int wj = t % kWin, wi = t / kWin;
        const int gi = wi0 + wi, gj = wj0 + wj, gk = wk0 + wk;
        if ((unsigned)gi >= (unsigned)g.nx || (unsigned)gj >= (unsigned)g.ny || (unsigned)gk >= (unsigned)g.nz) continue;
        const int id = gi * g.sx + gj * g.sy + gk;
        const real vx = win[w], vy = win[kWin3 + w], vz = win[2 * kWin3 + w];
        if (vx != real(0)) unsafeAtomicAdd(cur.jx + id, vx);
        if (vy != real(0)) unsafeAtomicAdd(cur.jy + id, vy);
        if (vz != real(0)) unsafeAtomicAdd(cur.jz + id, vz);
    }
    block_accumulate(diag, da.q, da.e, da.n, da.fallback);
}

// Photons with QED off move on straight lines (src/advance.cpp:173-183) and
// carry no current.
template <typename real>
__global__ void __launch_bounds__(kBlock) k_photon_move(Grid g, Species<real> s, int64_t first, int64_t n,
                                                        StepConst<real> sc, double *diag)
{
    const int64_t p = first + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    DiagAcc da;
    if (p < first + n) {
        const int c = s.cell[p];
        if (c >= 0) {
            const real u[3] = {s.ux[p], s.uy[p], s.uz[p]};
            const real gam = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            const real a = sc.dt / gam;
            real pos[3] = {s.ox[p] + u[0] * a * sc.inv_d[0], s.oy[p] + u[1] * a * sc.inv_d[1],
                           s.oz[p] + u[2] * a * sc.inv_d[2]};
            int ci[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const real fl = floor(pos[m]);
                ci[m] = (int)fl;
                pos[m] -= fl;
                if (pos[m] >= real(1)) {
                    pos[m] = real(0);
                    ci[m] += 1;
                }
            }
            const int ni = cell_i(c) + ci[0], nj = cell_j(c) + ci[1], nk = cell_k(c) + ci[2];
            const bool finite = fabs(u[0] * a) < real(1e30) && fabs(u[1] * a) < real(1e30) && fabs(u[2] * a) < real(1e30);
            if (finite && is_inside(g, ni, nj, nk)) {
                s.cell[p] = pack_cell(ni, nj, nk);
                s.ox[p] = pos[0];
                s.oy[p] = pos[1];
                s.oz[p] = pos[2];
                const double q = s.q[p];
                da.q = q;
                da.n = 1;
                da.e = (double)gam * q;
            } else {
                s.cell[p] = -1;
            }
        }
    }
    block_accumulate(diag, da.q, da.e, da.n, 0.0);
}

// Uniform thermal plasma: ppc particles per is_inside cell with ilo <= i < ihi,
// Philox keyed by (seed, cell * ppc + slot), 24-bit uniform offsets and
// Box-Muller normal momenta; written in cell order after slot `first`.
template <typename real>
__global__ void __launch_bounds__(kBlock) k_load_uniform(Grid g, Species<real> s, int64_t first, int ppc, double q,
                                                         double sigma, uint64_t seed, int ilo, int ihi, int64_t total)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= total) return;
    const int slot = (int)(t % ppc);
    int64_t c = t / ppc;
    const int nk = g.nz - 3, nj = g.ny - 3;
    const int k = 1 + (int)(c % nk);
    c /= nk;
    const int j = 1 + (int)(c % nj);
    const int i = ilo + (int)(c / nj);
    const uint64_t cell = ((uint64_t)i * g.ny + j) * g.nz + k;
    const uint64_t gid = cell * (uint64_t)ppc + slot;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const u32x4 a = philox(u32x4{{(uint32_t)gid, (uint32_t)(gid >> 32), 0u, 0x51ADu}}, k0, k1);
    const u32x4 b = philox(u32x4{{(uint32_t)gid, (uint32_t)(gid >> 32), 1u, 0x51ADu}}, k0, k1);
    const double two_pi = 6.283185307179586;
    const double r1 = sqrt(-2.0 * log(u24_open0(a.v[3])));
    const double t1 = two_pi * u24(b.v[0]);
    const double r2 = sqrt(-2.0 * log(u24_open0(b.v[1])));
    const double t2 = two_pi * u24(b.v[2]);
    const int64_t o = first + t;
    s.cell[o] = pack_cell(i, j, k);
    s.ox[o] = (real)u24(a.v[0]);
    s.oy[o] = (real)u24(a.v[1]);
    s.oz[o] = (real)u24(a.v[2]);
    s.ux[o] = (real)(sigma * r1 * cos(t1));
    s.uy[o] = (real)(sigma * r1 * sin(t1));
    s.uz[o] = (real)(sigma * r2 * cos(t2));
    s.q[o] = (real)q;
    if (s.chi) s.chi[o] = real(0);
    if (s.id) s.id[o] = (int)gid;
}

// node charge with trilinear weights (rhodeposition, src/deposition.cpp:358-429)
template <typename real>
__global__ void __launch_bounds__(kBlock) k_compute_rho(Grid g, Species<real> s, int64_t n, double *rho, int slot)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const int c = s.cell[p];
    if (c < 0) return;
    const int i = cell_i(c), j = cell_j(c), k = cell_k(c);
    const double xa = s.ox[p], ya = s.oy[p], za = s.oz[p];
    const double xb = 1 - xa, yb = 1 - ya, zb = 1 - za;
    const double q = s.q[p];
    const int b = i * g.sx + j * g.sy + k;
    double *r = rho + slot;
    unsafeAtomicAdd(r + 3 * (int64_t)b, q * xb * yb * zb);
    unsafeAtomicAdd(r + 3 * (int64_t)(b + g.sx), q * xa * yb * zb);
    unsafeAtomicAdd(r + 3 * (int64_t)(b + g.sy), q * xb * ya * zb);
    unsafeAtomicAdd(r + 3 * (int64_t)(b + 1), q * xb * yb * za);
    unsafeAtomicAdd(r + 3 * (int64_t)(b + g.sx + g.sy), q * xa * ya * zb);
    unsafeAtomicAdd(r + 3 * (int64_t)(b + g.sx + 1), q * xa * yb * za);
    unsafeAtomicAdd(r + 3 * (int64_t)(b + g.sy + 1), q * xb * ya * za);
    unsafeAtomicAdd(r + 3 * (int64_t)(b + g.sx + g.sy + 1), q * xa * ya * za);
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

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
    return sc;
}

}  // namespace

template <typename real>
hipError_t launch_push_global(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                              const Species<real> &s, const PushArgs &a)
{
    if (a.n <= 0) return hipSuccess;
    const StepConst<real> sc = step_const<real>(g, a);
    if (a.kind == 1)
        k_photon_move<real><<<blocks_for(a.n), kBlock, 0, st>>>(g, s, a.first, a.n, sc, a.diag);
    else if (a.pusher == 1)
        k_push_global<real, 1><<<blocks_for(a.n), kBlock, 0, st>>>(g, f, j, s, a.first, a.n, sc, a.diag);
    else
        k_push_global<real, 0><<<blocks_for(a.n), kBlock, 0, st>>>(g, f, j, s, a.first, a.n, sc, a.diag);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_push_tiled(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                             const Species<real> &s, const PushArgs &a)
{
    if (a.max_chunks <= 0) return hipSuccess;
    const StepConst<real> sc = step_const<real>(g, a);
    if (a.pusher == 1)
        k_push_tiled<real, 1><<<a.max_chunks, kBlock, 0, st>>>(g, f, j, s, sc, a.chunks, a.nchunks, a.tg, a.diag);
    else
        k_push_tiled<real, 0><<<a.max_chunks, kBlock, 0, st>>>(g, f, j, s, sc, a.chunks, a.nchunks, a.tg, a.diag);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_load_uniform(hipStream_t st, const Grid &g, const Species<real> &s, int64_t first, int ppc,
                               double q, double sigma, uint64_t seed, int ilo, int ihi)
{
    const int64_t total = (int64_t)(ihi - ilo) * (g.ny - 3) * (g.nz - 3) * ppc;
    if (total <= 0) return hipSuccess;
    k_load_uniform<real><<<blocks_for(total), kBlock, 0, st>>>(g, s, first, ppc, q, sigma, seed, ilo, ihi, total);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_compute_rho(hipStream_t st, const Grid &g, const Species<real> &s, int64_t n, double *rho, int slot)
{
    if (n <= 0) return hipSuccess;
    k_compute_rho<real><<<blocks_for(n), kBlock, 0, st>>>(g, s, n, rho, slot);
    return hipGetLastError();
}

#define QK_INST(real)                                                                                            \
    template hipError_t launch_push_global<real>(hipStream_t, const Grid &, const Fields<real> &,               \
                                                 const Current<real> &, const Species<real> &, const PushArgs &); \
    template hipError_t launch_push_tiled<real>(hipStream_t, const Grid &, const Fields<real> &,                \
                                                const Current<real> &, const Species<real> &, const PushArgs &); \
    template hipError_t launch_load_uniform<real>(hipStream_t, const Grid &, const Species<real> &, int64_t, int, \
                                                  double, double, uint64_t, int, int);                           \
    template hipError_t launch_compute_rho<real>(hipStream_t, const Grid &, const Species<real> &, int64_t,       \
                                                 double *, int);
QK_INST(float)
QK_INST(double)
#undef QK_INST

}  // namespace qk
