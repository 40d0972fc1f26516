// qk_device.h -- device-side building blocks of the PIC step for gfx950:
// grid geometry, field views, gather, pushers, the charge-conserving segment
// deposit and the counter-based random generator.  Included by the kernel
// translation units only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qk {

// ---------------------------------------------------------------------------
// geometry

struct Grid {
    int nx, ny, nz;
    int sx;            // stride of i: ny * nz
    int sy;            // stride of j: nz
    int ncell;
    double dx, dy, dz, dt;
};

// packed cell coordinate of a particle: i in bits 20..30, j in 10..19, k in
// 0..9; a negative value marks a free slot
__host__ __device__ inline int pack_cell(int i, int j, int k) { return (i << 20) | (j << 10) | k; }
__host__ __device__ inline int cell_i(int c) { return c >> 20; }
__host__ __device__ inline int cell_j(int c) { return (c >> 10) & 1023; }
__host__ __device__ inline int cell_k(int c) { return c & 1023; }

// cell tiles of (1 << kTsx) x (1 << kTsy) x (1 << kTsz) cells, numbered
// i-major; particles are sorted by (tile, cell inside the tile) and stored
// inside a tile in the scatter's spread order (see qk_sort.hip)
#ifndef QK_TILE_SHIFTS
#define QK_TILE_SHIFTS 3, 3, 3
#endif
constexpr int kTileShifts[3] = {QK_TILE_SHIFTS};
constexpr int kTsx = kTileShifts[0], kTsy = kTileShifts[1], kTsz = kTileShifts[2];
constexpr int kTx = 1 << kTsx, kTy = 1 << kTsy, kTz = 1 << kTsz;
constexpr int kTileCells = kTx * kTy * kTz;

struct TileGeom {
    int ntx, nty, ntz;   // tiles per axis
    int ntiles;
    int nkeys;           // sort keys: ntiles * kTileCells
    int ishift;          // moving-window shifts since the sort: a tile's cells now start at i = 8 ti - ishift
};

__host__ __device__ inline int tile_of(const TileGeom &t, int i, int j, int k)
{
    return ((i >> kTsx) * t.nty + (j >> kTsy)) * t.ntz + (k >> kTsz);
}

// (tile, cell inside the tile): one key per cell, so the scatter's slot
// counters are spread over every cell rather than one per tile
__host__ __device__ inline int sort_key(const TileGeom &t, int c)
{
    const int i = c >> 20, j = (c >> 10) & 1023, k = c & 1023;
    return tile_of(t, i, j, k) * kTileCells + ((((i & (kTx - 1)) << kTsy | (j & (kTy - 1))) << kTsz) | (k & (kTz - 1)));
}

// Workgroups are dealt to the 8 XCDs of the chip round-robin by block index,
// and each XCD has its own L2.  xcd_block renumbers the blocks of a grid of
// nb so that XCD x runs the contiguous range [x q + min(x, r), ...) of
// virtual blocks (q, r = nb / 8, nb % 8): neighbouring work shares one L2.
__device__ inline int64_t xcd_block(int64_t b, int64_t nb)
{
    const int64_t q = nb / 8, r = nb % 8, x = b % 8;
    return x * q + (x < r ? x : r) + b / 8;
}

// is_inside (src/plist.cpp:42-48)
__device__ inline bool is_inside(const Grid &g, int i, int j, int k)
{
    return i > 0 && i < g.nx - 2 && j > 0 && j < g.ny - 2 && k > 0 && k < g.nz - 2;
}

template <typename real>
struct Fields {
    real *ex, *ey, *ez;
    real *bx, *by, *bz;
    real *bex, *bey, *bez;
};

template <typename real>
struct Current {
    real *jx, *jy, *jz;
};

// one particle species, structure of arrays
template <typename real>
struct Species {
    int *cell;
    real *ox, *oy, *oz;    // offsets inside the cell, [0, 1)
    real *ux, *uy, *uz;    // u = p / m c
    real *q;               // macro charge (n_cr units, signed)
    real *chi;             // QED parameter (may be null when QED is off)
    int *id;               // optional particle id
};

// ---------------------------------------------------------------------------
// gather (src/deposition.cpp:431-471): each component is bilinear in the two
// directions transverse to it, at the lower node along its own direction.
// Split into its loads and its interpolation, so a kernel can issue the loads
// of the next particle before it computes the current one.
template <typename real>
struct Corners {
    real e[3][4], b[3][4];   // the four nodes each component interpolates from
};

template <typename real>
__device__ inline void load_corners(const Fields<real> &f, int base, int sx, int sy, Corners<real> &c)
{
    const int ox[4] = {0, sy, 1, sy + 1};        // x components: (j, k) square
    const int oy[4] = {0, sx, 1, sx + 1};        // y components: (i, k) square
    const int oz[4] = {0, sx, sy, sx + sy};      // z components: (i, j) square
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        c.e[0][m] = f.ex[base + ox[m]];
        c.e[1][m] = f.ey[base + oy[m]];
        c.e[2][m] = f.ez[base + oz[m]];
        c.b[0][m] = f.bex[base + ox[m]];
        c.b[1][m] = f.bey[base + oy[m]];
        c.b[2][m] = f.bez[base + oz[m]];
    }
}

template <typename real>
__device__ inline void interp_corners(const Corners<real> &c, real xa, real ya, real za, real e[3], real b[3])
{
    const real xb = real(1) - xa, yb = real(1) - ya, zb = real(1) - za;
    const real w[3][4] = {{yb * zb, ya * zb, yb * za, ya * za},
                          {xb * zb, xa * zb, xb * za, xa * za},
                          {xb * yb, xa * yb, xb * ya, xa * ya}};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        e[a] = c.e[a][0] * w[a][0] + c.e[a][1] * w[a][1] + c.e[a][2] * w[a][2] + c.e[a][3] * w[a][3];
        b[a] = c.b[a][0] * w[a][0] + c.b[a][1] * w[a][1] + c.b[a][2] * w[a][2] + c.b[a][3] * w[a][3];
    }
}

// ---------------------------------------------------------------------------
// reciprocal and square root: the hardware instructions (1 ulp) in FP32,
// IEEE-exact in FP64

__device__ inline float qk_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double qk_rcp(double x) { return 1.0 / x; }
__device__ inline float qk_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double qk_sqrt(double x) { return sqrt(x); }

// ---------------------------------------------------------------------------
// pushers (src/pusher.cpp)

// Vay 2008 (pusher.cpp:5-42); g enters as gamma at n-1/2 and leaves as gamma at n+1/2
template <typename real>
__device__ inline void push_vay(real cmr_dt, const real ef[3], const real bf[3], real u[3], real &g)
{
    const real e0 = ef[0] * cmr_dt, e1 = ef[1] * cmr_dt, e2 = ef[2] * cmr_dt;
    const real h = real(0.5) * cmr_dt;
    const real b0 = bf[0] * h, b1 = bf[1] * h, b2 = bf[2] * h;
    const real ig = qk_rcp(g);
    const real w0 = u[0] + e0 + (u[1] * b2 - u[2] * b1) * ig;
    const real w1 = u[1] + e1 + (u[2] * b0 - u[0] * b2) * ig;
    const real w2 = u[2] + e2 + (u[0] * b1 - u[1] * b0) * ig;
    const real bb = b0 * b0 + b1 * b1 + b2 * b2;
    const real bw = b0 * w0 + b1 * w1 + b2 * w2;
    const real s = real(1) + w0 * w0 + w1 * w1 + w2 * w2 - bb;
    g = qk_sqrt(real(0.5) * (s + qk_sqrt(s * s + real(4) * (bb + bw * bw))));
    const real ign = qk_rcp(g);
    const real ig2 = ign * ign;
    const real den = qk_rcp(real(1) + bb * ig2);
    const real bwg = bw * ig2;
    u[0] = (w0 + (w1 * b2 - w2 * b1) * ign + b0 * bwg) * den;
    u[1] = (w1 + (w2 * b0 - w0 * b2) * ign + b1 * bwg) * den;
    u[2] = (w2 + (w0 * b1 - w1 * b0) * ign + b2 * bwg) * den;
}

// Boris (pusher.cpp:44-97)
template <typename real>
__device__ inline void push_boris(real cmr_dt, const real ef[3], const real bf[3], real u[3], real &g)
{
    const real h = real(0.5) * cmr_dt;
    const real e0 = ef[0] * h, e1 = ef[1] * h, e2 = ef[2] * h;
    real m0 = u[0] + e0, m1 = u[1] + e1, m2 = u[2] + e2;
    const real hg = h * qk_rcp(qk_sqrt(real(1) + m0 * m0 + m1 * m1 + m2 * m2));
    const real t0 = bf[0] * hg, t1 = bf[1] * hg, t2 = bf[2] * hg;
    const real w0 = m0 + (m1 * t2 - m2 * t1);
    const real w1 = m1 + (m2 * t0 - m0 * t2);
    const real w2 = m2 + (m0 * t1 - m1 * t0);
    const real f = real(2) * qk_rcp(real(1) + t0 * t0 + t1 * t1 + t2 * t2);
    m0 += (w1 * t2 - w2 * t1) * f;
    m1 += (w2 * t0 - w0 * t2) * f;
    m2 += (w0 * t1 - w1 * t0) * f;
    u[0] = m0 + e0;
    u[1] = m1 + e1;
    u[2] = m2 + e2;
    g = qk_sqrt(real(1) + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
}

template <int PUSHER, typename real>
__device__ inline void push(real cmr_dt, const real ef[3], const real bf[3], real u[3], real &g)
{
    if (PUSHER == 1) push_boris(cmr_dt, ef, bf, u, g);
    else push_vay(cmr_dt, ef, bf, u, g);
}

// ---------------------------------------------------------------------------
// charge-conserving deposit of one straight segment (simple_jdep,
// src/deposition.cpp:13-39).  p is the segment start relative to the cell
// whose lower corner is w, d the segment.  The Acc policy receives
// (component, di, dj, dk, value) with offsets relative to w.
//
// The weights are the reference's, written at the segment's midpoint m = p +
// d/2: for jx at the node (j + 1, k + 1) the reference's ya za + ya d2/2 +
// za d1/2 + d1 d2/3 (ya = p.y, za = p.z) equals t = m.y m.z + d1 d2/12, and
// the other three nodes of the square take m.z - t, m.y - t and
// (1 - m.y - m.z) + t (the four sum to 1).  Same for jy over (x, z) and jz
// over (x, y).  Fewer operations than the reference's form, equal up to
// rounding.
template <typename real, class Acc>
__device__ inline void deposit_segment(Acc &acc, real q, real px, real py, real pz, real d0, real d1, real d2)
{
    const real h = real(0.5), c12 = real(1) / real(12);
    const real xm = px + h * d0, ym = py + h * d1, zm = pz + h * d2;
    const real qx = d0 * q, qy = d1 * q, qz = d2 * q;
    {
        const real t = ym * zm + d1 * d2 * c12;
        acc(0, 0, 1, 1, qx * t);
        acc(0, 0, 0, 1, qx * (zm - t));
        acc(0, 0, 1, 0, qx * (ym - t));
        acc(0, 0, 0, 0, qx * ((real(1) - ym - zm) + t));
    }
    {
        const real t = xm * zm + d0 * d2 * c12;
        acc(1, 1, 0, 1, qy * t);
        acc(1, 0, 0, 1, qy * (zm - t));
        acc(1, 1, 0, 0, qy * (xm - t));
        acc(1, 0, 0, 0, qy * ((real(1) - xm - zm) + t));
    }
    {
        const real t = xm * ym + d0 * d1 * c12;
        acc(2, 1, 1, 0, qz * t);
        acc(2, 0, 1, 0, qz * (ym - t));
        acc(2, 1, 0, 0, qz * (xm - t));
        acc(2, 0, 0, 0, qz * ((real(1) - xm - ym) + t));
    }
}

// Split the straight path d, starting at p (relative to the start cell, so
// 0 <= p < 1 on each axis), at every crossed cell face and deposit each piece
// (jdeposition, src/deposition.cpp:41-356).  At most one face per axis is
// crossed (|d| < 1), at path fraction t = (face - p) / d.  The pieces are the
// path between consecutive crossings in increasing t (regulate,
// src/low-level.cpp:316-366, sorts the same fractions); piece n lies in the
// start cell with the first n crossed axes switched to their new index, which
// is the cell the reference assigns in each of its 0-3 crossing cases (for
// three crossings it takes the floor of the piece's midpoint, the same cell).
// Ties between crossing fractions only reorder a zero-length piece, which
// deposits nothing.  A wavefront runs as many pieces as its lane with the
// most crossings, not one code path per crossing count.  The Acc policy gets
// set_cell(wi, wj, wk) (offsets of the piece's cell from the start cell)
// before each piece.  On return p holds the end point, advanced piece by
// piece as the reference advances the particle.
template <typename real, class Acc>
__device__ inline void deposit_path(Acc &acc, real q, real p[3], const real d[3])
{
    int c2[3];
    real t[3];
    int ax[3] = {0, 1, 2};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        c2[m] = (int)floor(p[m] + d[m]);
        // non-crossing axes sort after every crossing one
        t[m] = c2[m] != 0 ? (real(c2[m] > 0 ? 1 : 0) - p[m]) * qk_rcp(d[m]) : real(2);
    }
    const int nc = (c2[0] != 0) + (c2[1] != 0) + (c2[2] != 0);
    auto cswap = [&](int a, int b) {
        if (t[a] > t[b]) {
            const real tt = t[a]; t[a] = t[b]; t[b] = tt;
            const int ta = ax[a]; ax[a] = ax[b]; ax[b] = ta;
        }
    };
    cswap(0, 1);
    cswap(1, 2);
    cswap(0, 1);
    int w[3] = {0, 0, 0};
    real done = real(0);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        if (n > nc) break;
        const real upto = n < nc ? t[n] : real(1);
        const real fr = upto - done;
        done = upto;
        const real s0 = d[0] * fr, s1 = d[1] * fr, s2 = d[2] * fr;
        acc.set_cell(w[0], w[1], w[2]);
        deposit_segment<real>(acc, q, p[0] - real(w[0]), p[1] - real(w[1]), p[2] - real(w[2]), s0, s1, s2);
        p[0] += s0;
        p[1] += s1;
        p[2] += s2;
        if (n < 3) {
#pragma unroll
            for (int m = 0; m < 3; ++m)
                if (ax[n] == m) w[m] = c2[m];
        }
    }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11)

struct u32x4 {
    uint32_t v[4];
};

__host__ __device__ inline u32x4 philox(u32x4 c, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
        u32x4 n;
        n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
        n.v[1] = (uint32_t)p1;
        n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
        n.v[3] = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

__host__ __device__ inline double u24(uint32_t r) { return (double)(r >> 8) * (1.0 / 16777216.0); }
__host__ __device__ inline double u24_open0(uint32_t r) { return (double)((r >> 8) + 1u) * (1.0 / 16777216.0); }

// ---------------------------------------------------------------------------
// wave64 reductions

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

}  // namespace qk
