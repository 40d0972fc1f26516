// qk_particles.hip -- particle kernels of the step on gfx950: the fused
// gather + push + charge-conserving deposit + re-binning kernels (the particle
// loop of src/advance.cpp:90-387) in two forms, photon transport, the device
// plasma loader and the node-charge deposit used for the Gauss-law check.
//
//   k_push_global  one thread per particle, each current contribution a global
//                  atomic (the plain method, used for the FP64 comparison)
//   k_push_tiled   one workgroup per chunk of an 8x8x8 cell tile; particles are
//                  sorted by tile, so the workgroup stages the tile's E and Be
//                  nodes in LDS once, gathers from there, and deposits into an
//                  LDS window of the tile plus a halo, flushed once with one
//                  global atomic per nonzero window entry.  Inside a tile the
//                  storage order puts particles of distinct cells in
//                  consecutive slots, so a wavefront's LDS atomics do not
//                  collide.  A particle that drifted out of the window since
//                  the last sort gathers and deposits through global memory.
#include "qk_push.h"

namespace qk {

namespace {

// A workgroup of the tiled kernel serves particles whose start cell lies
// within kHalo cells of its tile ("in the window").  Relative to the tile's
// lower corner such a cell has r in [-kHalo, kT + kHalo) on each axis.
#ifndef QK_HALO
#define QK_HALO 1
#endif
constexpr int kHalo = QK_HALO;
constexpr int kT[3] = {kTx, kTy, kTz};
// threads per workgroup of the tiled kernel: the LDS window and field stage
// (78 KB) admit two workgroups per CU, so the workgroup is made large enough
// to keep four wavefronts per SIMD in flight
#ifndef QK_TBLOCK
#define QK_TBLOCK 512
#endif
#ifdef QK_TWAVES_EU
#define QK_TATTR __attribute__((amdgpu_waves_per_eu(QK_TWAVES_EU)))
#else
#define QK_TATTR
#endif
constexpr int kTBlock = QK_TBLOCK;
// out-of-window particles listed per workgroup and pushed after its rounds:
// at most this many per chunk, the rest inline
constexpr int kDefer = 512;

// Current window, FP64, one block per component.  A deposit from start cell
// r touches cells r-1 .. r+1 (one face crossing per axis at most) and, for
// the two axes transverse to the component, the node above as well: the
// window spans r - 1 .. r + 1 along the component's axis and r - 1 .. r + 2
// across it, i.e. window index = r + kHalo + 1.
__host__ __device__ constexpr int jdim(int c, int a) { return kT[a] + 2 * kHalo + (a == c ? 2 : 3); }
__host__ __device__ constexpr int jsize(int c) { return jdim(c, 0) * jdim(c, 1) * jdim(c, 2); }
__host__ __device__ constexpr int joff(int c) { return c == 0 ? 0 : (c == 1 ? jsize(0) : jsize(0) + jsize(1)); }
constexpr int kJWin = jsize(0) + jsize(1) + jsize(2);

// Field stage: E and Be nodes a window particle interpolates from.  Each
// component needs its own axis at the start cell only and the node above
// across it (gather at the lower node, src/deposition.cpp:431-471): index =
// r + kHalo, extent kT + 2 kHalo along the component, one more across.
__host__ __device__ constexpr int fdim(int c, int a) { return kT[a] + 2 * kHalo + (a == c ? 0 : 1); }
__host__ __device__ constexpr int fsize(int c) { return fdim(c, 0) * fdim(c, 1) * fdim(c, 2); }
__host__ __device__ constexpr int foff(int c) { return c == 0 ? 0 : (c == 1 ? fsize(0) : fsize(0) + fsize(1)); }
constexpr int kFStage = fsize(0) + fsize(1) + fsize(2);   // per field (E or Be)

// contributions go to the workgroup's LDS current window (one block per
// component, dims jdim).  The window is FP64 for both particle precisions:
// on gfx950 an FP64 LDS atomic add sustains about twenty times the rate of
// an FP32 one, and the tile's partial sums keep more bits before the single
// FP32/FP64 flush.
template <typename real>
struct LdsAcc {
    double *win;
    int base[3];   // window index of the particle's start cell, per component
    int cur[3];    // ... of the current piece's cell
    __device__ inline LdsAcc(double *w, int ri, int rj, int rk) : win(w)
    {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            base[c] = ((ri + kHalo + 1) * jdim(c, 1) + (rj + kHalo + 1)) * jdim(c, 2) + (rk + kHalo + 1);
    }
    __device__ inline void set_cell(int wi, int wj, int wk)
    {
#pragma unroll
        for (int c = 0; c < 3; ++c) cur[c] = base[c] + (wi * jdim(c, 1) + wj) * jdim(c, 2) + wk;
    }
    __device__ inline void operator()(int comp, int di, int dj, int dk, real v)
    {
        const int d1 = jdim(comp, 1), d2 = jdim(comp, 2);
        atomicAdd(win + joff(comp) + cur[comp] + (di * d1 + dj) * d2 + dk, (double)v);
    }
};

// Corner nodes of a window particle from the field stage (E then Be, one
// block per component, dims fdim); the same squares as load_corners.
template <typename real>
__device__ inline void load_corners_stage(const real *st, int ri, int rj, int rk, Corners<real> &c)
{
    int b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        b[a] = foff(a) + ((ri + kHalo) * fdim(a, 1) + (rj + kHalo)) * fdim(a, 2) + (rk + kHalo);
    const int sx0 = fdim(0, 2), sx1 = fdim(1, 1) * fdim(1, 2), sy2 = fdim(2, 1) * fdim(2, 2), sz2 = fdim(2, 2);
    const int ox[4] = {0, sx0, 1, sx0 + 1};          // x components: (j, k) square
    const int oy[4] = {0, sx1, 1, sx1 + 1};          // y components: (i, k) square
    const int oz[4] = {0, sy2, sz2, sy2 + sz2};      // z components: (i, j) square
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        c.e[0][m] = st[b[0] + ox[m]];
        c.e[1][m] = st[b[1] + oy[m]];
        c.e[2][m] = st[b[2] + oz[m]];
        c.b[0][m] = st[kFStage + b[0] + ox[m]];
        c.b[1][m] = st[kFStage + b[1] + oy[m]];
        c.b[2][m] = st[kFStage + b[2] + oz[m]];
    }
}

// Component C of one field into its stage block, in two passes over a
// thread's share of the block: the loads (load_component), then the LDS
// writes (store_component).  Nodes outside the grid (never read by a
// particle) are set to zero.  With the loads of all six blocks issued before
// the first write, a thread waits for memory once while staging, not once
// per block and share (the push measured 5 % faster at 256x128x128).
constexpr int kStageIt = (fsize(0) + kTBlock - 1) / kTBlock;   // every block has fsize(0) nodes

template <int C, typename real>
__device__ inline void load_component(real (&v)[kStageIt], const real *__restrict__ src, const Grid &g, int gi0,
                                      int gj0, int gk0)
{
    static_assert(fsize(C) == fsize(0), "stage blocks of equal size");
    constexpr int d1 = fdim(C, 1), d2 = fdim(C, 2);
#pragma unroll
    for (int m = 0; m < kStageIt; ++m) {
        const int w = threadIdx.x + m * kTBlock;
        const int lk = w % d2, t = w / d2;
        const int lj = t % d1, li = t / d1;
        const int gi = gi0 + li, gj = gj0 + lj, gk = gk0 + lk;
        const bool in = w < fsize(C) && (unsigned)gi < (unsigned)g.nx && (unsigned)gj < (unsigned)g.ny &&
                        (unsigned)gk < (unsigned)g.nz;
        v[m] = in ? src[gi * g.sx + gj * g.sy + gk] : real(0);
    }
}

template <int C, typename real>
__device__ inline void store_component(real *dst, const real (&v)[kStageIt])
{
#pragma unroll
    for (int m = 0; m < kStageIt; ++m) {
        const int w = threadIdx.x + m * kTBlock;
        if (w < fsize(C)) dst[foff(C) + w] = v[m];
    }
}

// add component C of the current window to the grid: one global atomic per
// nonzero entry inside the grid.  v holds the thread's share of the window
// block (entry threadIdx.x + m kTBlock), read from LDS for all three
// components before the first atomic.
constexpr int kFlushIt = (jsize(0) + kTBlock - 1) / kTBlock;

template <int C, typename real>
__device__ inline void flush_component(real *__restrict__ dst, const double (&v)[kFlushIt], const Grid &g, int gi0,
                                       int gj0, int gk0)
{
    static_assert(jsize(C) <= kFlushIt * kTBlock, "window block larger than the thread shares");
    constexpr int d1 = jdim(C, 1), d2 = jdim(C, 2);
#pragma unroll
    for (int m = 0; m < kFlushIt; ++m) {
        const int w = threadIdx.x + m * kTBlock;
        const int wk = w % d2, t = w / d2;
        const int wj = t % d1, wi = t / d1;
        const int gi = gi0 + wi, gj = gj0 + wj, gk = gk0 + wk;
        if (w < jsize(C) && v[m] != 0.0 && (unsigned)gi < (unsigned)g.nx && (unsigned)gj < (unsigned)g.ny &&
            (unsigned)gk < (unsigned)g.nz)
            unsafeAtomicAdd(dst + gi * g.sx + gj * g.sy + gk, (real)v[m]);
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
            Corners<real> nodes;
            load_corners(f, cell_base(g, r.cell), g.sx, g.sy, nodes);
            advance_particle<real, PUSHER>(g, nodes, r, sc, dep, da);
            store_rec(s, p, r);
        }
    }
    block_accumulate(diag, da.q, da.e, da.n, da.fallback);
}


// A particle that starts outside the window is not pushed in its round but
// listed (its slot) in LDS, and the workgroup pushes the listed particles
// after its rounds, one per thread, through global memory.  Inline, every
// wavefront with one such lane would run the global gather and deposit for
// it; listed, they run densely (the push measured 3.8 % faster at
// 512x256x256, where they gather in the laser's wake, and 1.5 % slower at
// 256x128x128, where a few per chunk make a short serial tail).
template <typename real, int PUSHER>
__global__ void __launch_bounds__(kTBlock) QK_TATTR k_push_tiled(Grid g, Fields<real> f, Current<real> cur, Species<real> s,
                                                       StepConst<real> sc, const int4 *__restrict__ chunks,
                                                       const int *__restrict__ nchunks, TileGeom tg, double *diag)
{
    __shared__ double win[kJWin];
    __shared__ real stage[2 * kFStage];
    __shared__ uint32_t listed[kDefer];
    __shared__ int nlisted;
    // chunks in launch order: the round-robin deal of workgroups to the XCDs
    // spreads neighbouring tiles over all eight, which measured faster here
    // than grouping them on one XCD (unlike the scatter, see qk_sort.hip)
    const int vb = (int)blockIdx.x;
    // the work-list entry loads together with the count (the list has an
    // entry for every launched workgroup; those past the count are unused)
    const int4 ch = chunks[vb];
    if (vb >= *nchunks) return;   // whole workgroup leaves together
    const int tile = ch.x;
    const int tk = tile % tg.ntz, t2 = tile / tg.ntz;
    const int tj = t2 % tg.nty, ti = t2 / tg.nty;
    const int i0 = ti * kTx - tg.ishift, j0 = tj * kTy, k0 = tk * kTz;
    // Slots as 32-bit indices: the record arrays are addressed as scalar
    // base + one shared 32-bit byte offset (capacity keeps slot * size below
    // 2^32).  The first two rounds' records load while the tile is staged.
    const uint32_t end32 = (uint32_t)(ch.y + ch.z);
    uint32_t p = (uint32_t)ch.y + threadIdx.x;
    PRec<real> r1, r2;   // the records one and two rounds ahead
    if (p < end32) load_rec32(s, p, r1);
    if (p + kTBlock < end32) load_rec32(s, p + kTBlock, r2);
    for (int w = threadIdx.x; w < kJWin; w += kTBlock) win[w] = 0.0;
    if (threadIdx.x == 0) nlisted = 0;
    {
        const int si = i0 - kHalo, sj = j0 - kHalo, sk = k0 - kHalo;
        real v[6][kStageIt];
        load_component<0>(v[0], f.ex, g, si, sj, sk);
        load_component<1>(v[1], f.ey, g, si, sj, sk);
        load_component<2>(v[2], f.ez, g, si, sj, sk);
        load_component<0>(v[3], f.bex, g, si, sj, sk);
        load_component<1>(v[4], f.bey, g, si, sj, sk);
        load_component<2>(v[5], f.bez, g, si, sj, sk);
        store_component<0>(stage, v[0]);
        store_component<1>(stage, v[1]);
        store_component<2>(stage, v[2]);
        store_component<0>(stage + kFStage, v[3]);
        store_component<1>(stage + kFStage, v[4]);
        store_component<2>(stage + kFStage, v[5]);
    }
    __syncthreads();

    DiagAcc da;
    // start cell relative to the tile, and whether it lies in the window
    int ri = 0, rj = 0, rk = 0;
    bool in_win = false;
    auto dep = [&](int, int, int, int base, real q, real pos[3], const real d[3], DiagAcc &acc_d) {
        // one accumulator per branch: a shared one with per-lane atomics
        // (path split once, two masked atomics per term) measured 2.5 % slower
        if (in_win) {
            LdsAcc<real> acc(win, ri, rj, rk);
            deposit_path<real>(acc, q, pos, d);
        } else {
            GlobalAcc<real> acc{cur.jx, cur.jy, cur.jz, g.sx, g.sy, base, base};
            deposit_path<real>(acc, q, pos, d);
            acc_d.fallback += 1;
        }
    };
    // software pipeline over the rounds of kTBlock particles: while one
    // particle is pushed, the next round's record has its field nodes
    // gathered and the record two rounds ahead is loading (the first two
    // rounds' records were issued before the staging).  listed_q: the
    // particle was listed for the pass after the rounds.
    auto gather = [&](const PRec<real> &q, uint32_t slot, int &a, int &b, int &cc, bool &w, bool &listed_q,
                      Corners<real> &nd) {
        listed_q = false;
        if (q.cell < 0) return;
        a = cell_i(q.cell) - i0;
        b = cell_j(q.cell) - j0;
        cc = cell_k(q.cell) - k0;
        w = (unsigned)(a + kHalo) < (unsigned)(kTx + 2 * kHalo) && (unsigned)(b + kHalo) < (unsigned)(kTy + 2 * kHalo) &&
            (unsigned)(cc + kHalo) < (unsigned)(kTz + 2 * kHalo);
        if (w) {
            load_corners_stage(stage, a, b, cc, nd);
            return;
        }
        const int at = atomicAdd(&nlisted, 1);
        if (at < kDefer) {
            listed[at] = slot;
            listed_q = true;
            return;
        }
        load_corners(f, cell_base(g, q.cell), g.sx, g.sy, nd);
    };
    int ri1 = 0, rj1 = 0, rk1 = 0;
    bool w1 = false, l1 = false;
    // The pipeline carries the next particle's E and Be (six values), not
    // its 24 field nodes: the nodes' LDS reads are issued before the push of
    // the current particle and interpolated after it, so the nodes are dead
    // at the loop's end and only six values are copied there (about a
    // hundred fewer register moves per round; push 0.5 % faster).
    real e1[3] = {}, b1[3] = {};
    if (p < end32) {
        Corners<real> nd;
        gather(r1, p, ri1, rj1, rk1, w1, l1, nd);
        if (r1.cell >= 0 && !l1) interp_corners(nd, r1.o[0], r1.o[1], r1.o[2], e1, b1);
    }
    // Each round stores the record of the round before it, right after
    // issuing its own loads: gfx9 counts loads and stores in one vmcnt, so
    // the loop's wait for those loads (one round later) then waits for
    // stores that have had a whole push to complete, not for stores just
    // issued.  The first round rewrites its own record unchanged, and free
    // and listed slots are rewritten unchanged, which keeps the stores
    // unconditional.
    PRec<real> rs = r1;
    uint32_t ps = p;
    const bool started = p < end32;
    for (; p < end32; p += kTBlock) {
        PRec<real> r = r1;
        const real e[3] = {e1[0], e1[1], e1[2]}, b[3] = {b1[0], b1[1], b1[2]};
        ri = ri1;
        rj = rj1;
        rk = rk1;
        in_win = w1;
        const bool is_listed = l1;
        r1 = r2;
        const bool more = p + kTBlock < end32;
        Corners<real> nd;
        // The round's memory operations (the next particle's LDS reads, the
        // record two rounds ahead, the previous record's store) are issued
        // at a raised wavefront priority, so that they reach the memory
        // system ahead of the other wavefronts' arithmetic: push 1.3 %
        // faster at both 512x256x256 and 256x128x128 (priority 1 or 3, or
        // raised also for the staging or the flush, measured less).
        __builtin_amdgcn_s_setprio(2);
        if (more) gather(r1, p + kTBlock, ri1, rj1, rk1, w1, l1, nd);
        if (p + 2 * kTBlock < end32) load_rec32(s, p + 2 * kTBlock, r2);
        store_rec32(s, ps, rs);
        __builtin_amdgcn_s_setprio(0);
        if (r.cell >= 0 && !is_listed) advance_particle_eb<real, PUSHER>(g, e, b, r, sc, dep, da);
        if (more && r1.cell >= 0 && !l1) interp_corners(nd, r1.o[0], r1.o[1], r1.o[2], e1, b1);
        rs = r;
        ps = p;
    }
    if (started) store_rec32(s, ps, rs);
    // the listed slots' rewrites complete before the pass below stores them
    // again (vmcnt 0; expcnt and lgkmcnt left at their maxima)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    {
        auto gdep = [&](int, int, int, int base, real q, real pos[3], const real d[3], DiagAcc &acc_d) {
            GlobalAcc<real> acc{cur.jx, cur.jy, cur.jz, g.sx, g.sy, base, base};
            deposit_path<real>(acc, q, pos, d);
            acc_d.fallback += 1;
        };
        const int nl = min(nlisted, kDefer);
        for (int e = threadIdx.x; e < nl; e += kTBlock) {
            const uint32_t q = listed[e];
            PRec<real> rr;
            load_rec32(s, q, rr);
            Corners<real> nn;
            load_corners(f, cell_base(g, rr.cell), g.sx, g.sy, nn);
            advance_particle<real, PUSHER>(g, nn, rr, sc, gdep, da);
            store_rec32(s, q, rr);
        }
    }

    {
        const int wi0 = i0 - kHalo - 1, wj0 = j0 - kHalo - 1, wk0 = k0 - kHalo - 1;
        double v[3][kFlushIt];
#pragma unroll
        for (int m = 0; m < kFlushIt; ++m) {
            const int w = threadIdx.x + m * kTBlock;
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c][m] = w < jsize(c) ? win[joff(c) + w] : 0.0;
        }
        flush_component<0>(cur.jx, v[0], g, wi0, wj0, wk0);
        flush_component<1>(cur.jy, v[1], g, wi0, wj0, wk0);
        flush_component<2>(cur.jz, v[2], g, wi0, wj0, wk0);
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
        if (c >= 0 && !is_inside(g, cell_i(c), cell_j(c), cell_k(c))) {
            s.cell[p] = -1;   // p_boundary (src/plist.cpp:73-102)
        } else if (c >= 0) {
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
            if (!finite) count_nonfinite(sc.nonfinite);
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

// pmerging (src/mc.cpp:202-462), per species: total charge and count, the
// random deletion, the rescale of the survivors
template <typename real>
__global__ void __launch_bounds__(kBlock) k_charge_sum(Species<real> s, int64_t n, double *acc)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double q = 0, c = 0;
    if (p < n && s.cell[p] >= 0) {
        q = fabs((double)s.q[p]);
        c = 1;
    }
    block_accumulate(acc, q, 0.0, c, 0.0);
}

template <typename real>
__global__ void __launch_bounds__(kBlock) k_merge_delete(Species<real> s, int64_t n, int mode, double pd, double av_q,
                                                         uint64_t seed, uint32_t step, uint32_t tag, double *acc)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double sum = 0, c = 0;
    if (p < n && s.cell[p] >= 0) {
        const double q = (double)s.q[p], aq = fabs(q);
        const u32x4 v = philox(u32x4{{(uint32_t)p, step, tag, 0x3E76Eu}}, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double r = ((double)(v.v[0] >> 11) * 4294967296.0 + (double)v.v[1]) * (1.0 / 9007199254740992.0);
        if (mode == 1) {
            if (r < pd) {
                s.cell[p] = -1;
                sum = aq;
                c = 1;
            }
        } else {
            const double x = 2 * aq / av_q, pp = pd / (1 + x * x * x * x);   // _ppd, src/mc.cpp:198-200
            if (r < pp) {
                s.cell[p] = -1;
                c = 1;
            } else {
                const real qn = (real)(q / (1 - pp));
                s.q[p] = qn;
                sum = fabs((double)qn);
            }
        }
    }
    block_accumulate(acc, sum, 0.0, c, 0.0);
}

template <typename real>
__global__ void __launch_bounds__(kBlock) k_scale_charge(Species<real> s, int64_t n, double a)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < n && s.cell[p] >= 0) s.q[p] = (real)((double)s.q[p] * a);
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
        k_push_tiled<real, 1><<<a.max_chunks, kTBlock, 0, st>>>(g, f, j, s, sc, a.chunks, a.nchunks, a.tg, a.diag);
    else
        k_push_tiled<real, 0><<<a.max_chunks, kTBlock, 0, st>>>(g, f, j, s, sc, a.chunks, a.nchunks, a.tg, a.diag);
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

template <typename real>
hipError_t launch_charge_sum(hipStream_t st, const Species<real> &s, int64_t n, double *acc)
{
    if (n <= 0) return hipSuccess;
    k_charge_sum<real><<<blocks_for(n), kBlock, 0, st>>>(s, n, acc);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_merge_delete(hipStream_t st, const Species<real> &s, int64_t n, int mode, double p, double av_q,
                               uint64_t seed, uint32_t step, uint32_t tag, double *acc)
{
    if (n <= 0) return hipSuccess;
    k_merge_delete<real><<<blocks_for(n), kBlock, 0, st>>>(s, n, mode, p, av_q, seed, step, tag, acc);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_scale_charge(hipStream_t st, const Species<real> &s, int64_t n, double a)
{
    if (n <= 0) return hipSuccess;
    k_scale_charge<real><<<blocks_for(n), kBlock, 0, st>>>(s, n, a);
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
                                                 double *, int);                                                  \
    template hipError_t launch_charge_sum<real>(hipStream_t, const Species<real> &, int64_t, double *);           \
    template hipError_t launch_merge_delete<real>(hipStream_t, const Species<real> &, int64_t, int, double, double, \
                                                  uint64_t, uint32_t, uint32_t, double *);                        \
    template hipError_t launch_scale_charge<real>(hipStream_t, const Species<real> &, int64_t, double);
QK_INST(float)
QK_INST(double)
#undef QK_INST

}  // namespace qk
