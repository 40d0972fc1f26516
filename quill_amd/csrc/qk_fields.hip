// qk_fields.hip -- field kernels of the step on gfx950: the three Maxwell
// solvers (FDTD, NDFX, five-point) with the reference's boundary-plane sets,
// B interpolation to E positions, boundary zeroing, the laser initial
// condition and the FP64 field-energy reduction.
//
// Fields are structure-of-arrays, one array per component, [nx][ny][nz] with
// k fastest; a thread owns one cell and consecutive lanes walk k, so every
// stencil load of a wavefront is a coalesced run.
#include "qk_kernels.h"

namespace qk {

namespace {

constexpr int kBX = 64, kBY = 4;   // workgroup: 64 lanes along k, 4 rows along j

inline dim3 cell_grid(const Grid &g) { return dim3((g.nz + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, g.nx); }

// Block order of the field sweeps (grid x along k, y along j, z along the
// planes).  The hardware deals blocks to the 8 XCDs round-robin by linear
// index, which puts neighbouring rows of blocks on different XCDs (and
// L2s).  Here the planes still advance in lock-step over the XCDs, but each
// XCD takes one contiguous band of block rows in every plane, so a block
// finds the row above its own in its XCD's L2 (the neighbour plane is read
// by the same XCD a few blocks later, as before).  Field phase 1.18 -> 1.15
// ms at 512x256x256 and 0.145 -> 0.138 ms at 256x128x128; a contiguous
// range of whole planes per XCD was faster at the smaller grid only.
__device__ inline dim3 band_block()
{
    const int P = gridDim.x * gridDim.y;   // blocks per plane (run of planes)
    if (P % 8 != 0) return blockIdx;
    const int hb = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int x = hb % 8, l = hb / 8, q = P / 8;
    const int vin = x * q + l % q;
    return dim3((unsigned)(vin % gridDim.x), (unsigned)(vin / gridDim.x), (unsigned)(l / q));
}

__device__ inline bool in(int v, int lo, int hi) { return v >= lo && v <= hi; }

// which cells get component C of B updated
// (FDTD src/maxwell.cpp:445-484, NDFX :251-290, five-point :35-85)
template <int S, int C>
__device__ inline bool b_updates(const Grid &g, int i, int j, int k)
{
    const int c[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
    if (S == 2) {   // each index in [1, n-3] or at n-1, at most one at n-1
        int top = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (c[a] == n[a] - 1) ++top;
            else if (!in(c[a], 1, n[a] - 3)) return false;
        }
        return top <= 1;
    }
    const int lo = (S == 0) ? 0 : 1;   // lowest plane updated
    int nlow = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == C) {
            if (!in(c[a], lo + 1, n[a] - 2)) return false;
        } else {
            if (!in(c[a], lo, n[a] - 2)) return false;
            nlow += (c[a] == lo);
        }
    }
    return nlow <= 1;
}

// which cells get component C of E updated
// (FDTD src/maxwell.cpp:486-522, NDFX :292-328, five-point :87-125)
template <int S, int C>
__device__ inline bool e_updates(const Grid &g, int i, int j, int k)
{
    const int c[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
    const int lo = (S == 0) ? 0 : 1;
    const int top = (S == 2) ? 3 : 2;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int first = (a == C) ? lo : lo + 1;
        if (!in(c[a], first, n[a] - top)) return false;
    }
    return true;
}

template <typename real>
struct Coef {
    real a[3], b[3];
    real dtd[3];   // dt / d_alpha
    real d[3];     // d_alpha
};

template <int A>
__device__ inline int stride(const Grid &g)
{
    return A == 0 ? g.sx : (A == 1 ? g.sy : 1);
}

// zero outside the grid; used by the five-point planes at n-1 (src/maxwell.cpp:52-84)
template <typename real>
__device__ inline real rd0(const Grid &g, const real *F, int i, int j, int k)
{
    if ((unsigned)i >= (unsigned)g.nx || (unsigned)j >= (unsigned)g.ny || (unsigned)k >= (unsigned)g.nz) return real(0);
    return F[i * g.sx + j * g.sy + k];
}

// Difference of component C (stored in F) along axis A at cell (i,j,k);
// FWD: F[+1] - F[0] (B update), else F[0] - F[-1] (E update).
//   FDTD:  plain difference
//   NDFX:  smoothed along the component's own axis C with (b_C, a_C)
//   five:  b_A (F[+1]-F[0]) + a_A (F[+2]-F[+1]+F[0]-F[-1]), zero outside
template <int S, int C, int A, bool FWD, typename real>
__device__ inline real diff(const Grid &g, const Coef<real> &cf, const real *__restrict__ F, int i, int j, int k)
{
    const int ua = (A == 0), va = (A == 1), wa = (A == 2);
    const int li = FWD ? i : i - ua, lj = FWD ? j : j - va, lk = FWD ? k : k - wa;
    const int lo = li * g.sx + lj * g.sy + lk;
    const int sA = stride<A>(g);
    const int hi = lo + sA;
    if (S == 0) {
        return F[hi] - F[lo];
    } else if (S == 1) {
        const int sC = stride<C>(g);
        return cf.b[C] * (F[hi] - F[lo]) + cf.a[C] * ((F[hi + sC] - F[lo + sC]) + (F[hi - sC] - F[lo - sC]));
    } else {
        const real fm = rd0(g, F, li - ua, lj - va, lk - wa);
        const real f0 = rd0(g, F, li, lj, lk);
        const real f1 = rd0(g, F, li + ua, lj + va, lk + wa);
        const real f2 = rd0(g, F, li + 2 * ua, lj + 2 * va, lk + 2 * wa);
        return cf.b[A] * (f1 - f0) + cf.a[A] * (f2 - f1 + f0 - fm);
    }
}

// half step of B (maxwell_solver::advance_b, src/maxwell.cpp:10-22).  With
// energy != null the first half step also sums E^2 + B^2 of the fields it
// starts from, in FP64: compute_energy's field term, which the reference takes
// between padvance and fadvance (src/fout.cpp:102, src/main.cpp:2327-2331).
template <int S, typename real>
__global__ void __launch_bounds__(kBX * kBY) k_maxwell_b(Grid g, Coef<real> cf, const real *__restrict__ ex,
                                                         const real *__restrict__ ey, const real *__restrict__ ez,
                                                         real *__restrict__ bx, real *__restrict__ by,
                                                         real *__restrict__ bz, double *energy)
{
    __shared__ double part[kBX * kBY / 64];
    const dim3 vb = band_block();
    const int k = vb.x * kBX + threadIdx.x;
    const int j = vb.y * kBY + threadIdx.y;
    const int i = vb.z;
    double en = 0;
    if (k < g.nz && j < g.ny) {
        const int id = i * g.sx + j * g.sy + k;
        const real h = real(0.5);
        const real b0 = bx[id], b1 = by[id], b2 = bz[id];
        if (energy) {
            const double e0 = ex[id], e1 = ey[id], e2 = ez[id];
            en = e0 * e0 + e1 * e1 + e2 * e2 + (double)b0 * b0 + (double)b1 * b1 + (double)b2 * b2;
        }
        if (b_updates<S, 0>(g, i, j, k))
            bx[id] = b0 + h * (-cf.dtd[1] * diff<S, 2, 1, true>(g, cf, ez, i, j, k) + cf.dtd[2] * diff<S, 1, 2, true>(g, cf, ey, i, j, k));
        if (b_updates<S, 1>(g, i, j, k))
            by[id] = b1 + h * (cf.dtd[0] * diff<S, 2, 0, true>(g, cf, ez, i, j, k) - cf.dtd[2] * diff<S, 0, 2, true>(g, cf, ex, i, j, k));
        if (b_updates<S, 2>(g, i, j, k))
            bz[id] = b2 + h * (cf.dtd[1] * diff<S, 0, 1, true>(g, cf, ex, i, j, k) - cf.dtd[0] * diff<S, 1, 0, true>(g, cf, ey, i, j, k));
    }
    if (energy) {   // uniform across the launch
        en = wave_sum(en);
        const int t = threadIdx.y * kBX + threadIdx.x;
        if ((t & 63) == 0) part[t >> 6] = en;
        __syncthreads();
        if (t == 0) {
            double s = 0;
            for (int w = 0; w < kBX * kBY / 64; ++w) s += part[w];
            const int slot = (int)((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) % kEnergySlots;
            atomicAdd(energy + slot, s);
        }
    }
}

// E step with the current of this step; also clears the other current
// buffer, which the next step deposits into (replaces the J reset of
// src/advance.cpp:112-123 without a separate pass)
template <int S, typename real>
__global__ void __launch_bounds__(kBX * kBY) k_maxwell_e(Grid g, Coef<real> cf, real *__restrict__ ex, real *__restrict__ ey,
                                                         real *__restrict__ ez, const real *__restrict__ bx,
                                                         const real *__restrict__ by, const real *__restrict__ bz,
                                                         const real *__restrict__ jx, const real *__restrict__ jy,
                                                         const real *__restrict__ jz, real *__restrict__ njx,
                                                         real *__restrict__ njy, real *__restrict__ njz)
{
    const dim3 vb = band_block();
    const int k = vb.x * kBX + threadIdx.x;
    const int j = vb.y * kBY + threadIdx.y;
    const int i = vb.z;
    if (k >= g.nz || j >= g.ny) return;
    const int id = i * g.sx + j * g.sy + k;
    if (e_updates<S, 0>(g, i, j, k))
        ex[id] += cf.dtd[1] * diff<S, 2, 1, false>(g, cf, bz, i, j, k) - cf.dtd[2] * diff<S, 1, 2, false>(g, cf, by, i, j, k) - jx[id] * cf.d[0];
    if (e_updates<S, 1>(g, i, j, k))
        ey[id] += -cf.dtd[0] * diff<S, 2, 0, false>(g, cf, bz, i, j, k) + cf.dtd[2] * diff<S, 0, 2, false>(g, cf, bx, i, j, k) - jy[id] * cf.d[1];
    if (e_updates<S, 2>(g, i, j, k))
        ez[id] += cf.dtd[0] * diff<S, 1, 0, false>(g, cf, by, i, j, k) - cf.dtd[1] * diff<S, 0, 1, false>(g, cf, bx, i, j, k) - jz[id] * cf.d[2];
    njx[id] = real(0);
    njy[id] = real(0);
    njz[id] = real(0);
}

// Be = B averaged to E positions (src/advance.cpp:20-83): each component is
// the mean over a 2x2x2 stencil (Bx over i..i+1, j-1..j, k-1..k; By over
// i-1..i, j..j+1, k-1..k; Bz over i-1..i, j-1..j, k..k+1); on a boundary plane
// the half that leaves the grid is dropped (4-point mean); cells on two or
// more boundary planes keep their value.
//
// A workgroup owns a kBeK x kBeJ block of (j, k) columns over a run of
// planes.  It loads B of its block plus the one-column rim each component's
// square reaches (Bx: rows j-1, columns k-1; By: rows j+1, columns k-1; Bz:
// rows j-1, columns k+1), on the run's planes plus one (Bx one plane ahead,
// By and Bz one behind), with every load issued before the first LDS write
// and unconditional (indices clamped into the grid, entries outside it set
// to zero), then takes each cell's sums from LDS in the order and with the
// entries of a sum over the grid.  Each B value is loaded about 1.4 times
// instead of the thirteen L1 reads per cell of a thread summing its own
// squares from global memory (field phase 1.28 -> 1.18 ms at 512x256x256).
constexpr int kBeK = 64, kBeJ = 8;
constexpr int kBeW = kBeK + 1, kBeH = kBeJ + 1;   // rim-extended block, per component
constexpr int kBeCols = kBeW * kBeH;
// planes per run; the LDS holds the run plus one plane of three components
template <typename real>
struct BeRun {
    static constexpr int n = sizeof(real) == 4 ? 4 : 2;
};

// 2x2 sum of the LDS plane P (kBeH rows of kBeW) at (r, c) with the
// neighbour row r + dr (when vj) and column c + dc (when vk)
template <typename real>
__device__ inline real lds_sum(const real *P, int r, int c, int dr, int dc, bool vj, bool vk)
{
    real s = P[r * kBeW + c];
    if (vk) s += P[r * kBeW + c + dc];
    if (vj) s += P[(r + dr) * kBeW + c];
    if (vj && vk) s += P[(r + dr) * kBeW + c + dc];
    return s;
}

template <typename real>
__global__ void __launch_bounds__(kBeK * kBeJ) k_interpolate_be(Grid g, const real *__restrict__ bx,
                                                                const real *__restrict__ by,
                                                                const real *__restrict__ bz, real *__restrict__ bex,
                                                                real *__restrict__ bey, real *__restrict__ bez)
{
    constexpr int R = BeRun<real>::n;
    constexpr int kN = (R + 1) * kBeCols;   // LDS entries per component
    constexpr int kIt = (kN + kBeK * kBeJ - 1) / (kBeK * kBeJ);
    // [plane][component][row][column]: plane m holds Bx at x = i0 + m and By,
    // Bz at x = i0 - 1 + m; component c's (row, column) is (j, k) =
    // (j0 - 1 + row, k0 - 1 + col) for Bx, (j0 + row, k0 - 1 + col) for By,
    // (j0 - 1 + row, k0 + col) for Bz
    __shared__ real L[R + 1][3][kBeCols];
    const dim3 vb = band_block();
    const int k0 = vb.x * kBeK, j0 = vb.y * kBeJ, i0 = vb.z * R;
    const int tid = threadIdx.y * kBeK + threadIdx.x;
    real v[3][kIt];
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
        const int w = tid + it * kBeK * kBeJ;
        const int m = w / kBeCols, e = w - m * kBeCols;
        const int r = e / kBeW, q = e - r * kBeW;
        const int px = i0 + m, pyz = i0 - 1 + m;
        const bool live = w < kN;
        const int jx = j0 - 1 + r, kx = k0 - 1 + q, jy = j0 + r, kz = k0 + q;
        const bool inx = live && px < g.nx && (unsigned)jx < (unsigned)g.ny && (unsigned)kx < (unsigned)g.nz;
        const bool iny = live && pyz >= 0 && pyz < g.nx && (unsigned)jy < (unsigned)g.ny && (unsigned)kx < (unsigned)g.nz;
        const bool inz = live && pyz >= 0 && pyz < g.nx && (unsigned)jx < (unsigned)g.ny && (unsigned)kz < (unsigned)g.nz;
        const int cpx = min(px, g.nx - 1), cpyz = max(min(pyz, g.nx - 1), 0);
        const int cjx = max(min(jx, g.ny - 1), 0), cjy = min(jy, g.ny - 1);
        const int ckx = max(min(kx, g.nz - 1), 0), ckz = min(kz, g.nz - 1);
        const real ax = bx[cpx * g.sx + cjx * g.sy + ckx];
        const real ay = by[cpyz * g.sx + cjy * g.sy + ckx];
        const real az = bz[cpyz * g.sx + cjx * g.sy + ckz];
        v[0][it] = inx ? ax : real(0);
        v[1][it] = iny ? ay : real(0);
        v[2][it] = inz ? az : real(0);
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
        const int w = tid + it * kBeK * kBeJ;
        if (w < kN) {
            const int m = w / kBeCols, e = w - m * kBeCols;
            L[m][0][e] = v[0][it];
            L[m][1][e] = v[1][it];
            L[m][2][e] = v[2][it];
        }
    }
    __syncthreads();
    const int j = j0 + threadIdx.y, k = k0 + threadIdx.x;
    if (j >= g.ny || k >= g.nz) return;
    const int nbjk = (j == 0 || j == g.ny - 1) + (k == 0 || k == g.nz - 1);
    // entries of each component's 2x2 square inside the grid
    const bool xvj = j >= 1, xvk = k >= 1;           // Bx: rows j-1, j; columns k-1, k
    const bool yvj = j + 1 < g.ny, yvk = k >= 1;     // By: rows j, j+1; columns k-1, k
    const bool zvj = j >= 1, zvk = k + 1 < g.nz;     // Bz: rows j-1, j; columns k, k+1
    const int cx = (1 + xvj) * (1 + xvk), cy = (1 + yvj) * (1 + yvk), cz = (1 + zvj) * (1 + zvk);
    const int ty = threadIdx.y, tx = threadIdx.x;
#pragma unroll
    for (int m = 0; m < R; ++m) {
        // Be at x = i0 + m: Bx planes i, i + 1 (L[m], L[m + 1]), By and Bz
        // planes i - 1, i (L[m], L[m + 1])
        const int i = i0 + m;
        if (i < g.nx && nbjk + (i == 0 || i == g.nx - 1) < 2) {
            const bool up = i + 1 < g.nx;
            const real sx0 = lds_sum(L[m][0], ty + 1, tx + 1, -1, -1, xvj, xvk);
            const real px1 = up ? lds_sum(L[m + 1][0], ty + 1, tx + 1, -1, -1, xvj, xvk) : real(0);
            const real qy = i > 0 ? lds_sum(L[m][1], ty, tx + 1, 1, -1, yvj, yvk) : real(0);
            const real sy1 = lds_sum(L[m + 1][1], ty, tx + 1, 1, -1, yvj, yvk);
            const real rz = i > 0 ? lds_sum(L[m][2], ty + 1, tx, -1, 1, zvj, zvk) : real(0);
            const real sz1 = lds_sum(L[m + 1][2], ty + 1, tx, -1, 1, zvj, zvk);
            const int id = i * g.sx + j * g.sy + k;
            bex[id] = (sx0 + px1) * (cx * (1 + up) == 8 ? real(0.125) : real(0.25));
            bey[id] = (qy + sy1) * (cy * (1 + (i > 0)) == 8 ? real(0.125) : real(0.25));
            bez[id] = (rz + sz1) * (cz * (1 + (i > 0)) == 8 ? real(0.125) : real(0.25));
        }
    }
}

// init_boundaries: zero components on boundary planes, transverse indices in
// [0, n-2] (FDTD src/maxwell.cpp:532-597, NDFX :338-438, five-point :135-236)
template <typename real>
__global__ void __launch_bounds__(kBX * kBY) k_init_boundaries(Grid g, int solver, Fields<real> f)
{
    const int k = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const int c[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
    int em = 0, bm = 0;
    for (int a = 0; a < 3; ++a) {
        const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
        if (c[b1] > n[b1] - 2 || c[b2] > n[b2] - 2) continue;
        const int normal = 1 << a, tang = 7 & ~normal;
        const int p = c[a];
        if (solver == 0) {          // FDTD
            if (p == 0) { em |= tang; bm |= normal; }
            if (p == n[a] - 1) { em |= 7; bm |= 7; }
        } else if (solver == 1) {   // NDFX
            if (p == 0 || p == n[a] - 1) { em |= 7; bm |= 7; }
            if (p == 1) { em |= tang; bm |= normal; }
        } else {                    // five-point
            if (p == 0 || p == n[a] - 1) { em |= 7; bm |= 7; }
            if (p == 1) em |= tang;
            if (p == n[a] - 2) bm |= tang;
        }
    }
    const int id = i * g.sx + j * g.sy + k;
    if (em & 1) f.ex[id] = real(0);
    if (em & 2) f.ey[id] = real(0);
    if (em & 4) f.ez[id] = real(0);
    if (bm & 1) f.bx[id] = real(0);
    if (bm & 2) f.by[id] = real(0);
    if (bm & 4) f.bz[id] = real(0);
}

// The pulse of f_init_cos with the cos envelope (src/init.cpp:6-210) at the
// point (px, py, pz) (c/omega units from the subdomain origin), after it has
// travelled `shift` along its direction of propagation (+r0 when b_sign, -r0
// otherwise: the direction of E x B); shift = 0 is f_init_cos itself.  FP64.
// Returns false outside the envelope (ev, bv untouched).
__device__ inline bool pulse_at(const LaserFrame &L, double px, double py, double pz, double shift, double ev[3],
                                double bv[3])
{
    const double PI = 3.141592653589793;
    const double X = px - L.x0 - L.xt, Y = py - L.y12 - L.y0 - L.yt, Z = pz - L.z12 - L.z0 - L.zt;
    double x = L.r0[0] * X + L.r0[1] * Y + L.r0[2] * Z;
    const double y = L.ey[0] * X + L.ey[1] * Y + L.ey[2] * Z;
    const double z = L.ez[0] * X + L.ez[1] * Y + L.ez[2] * Z;
    if (shift != 0) x -= L.b_sign ? shift : -shift;
    if (!(x < L.xs && x > -L.xs && y < L.ys && y > -L.ys && z < L.zs && z > -L.zs)) return false;
    const double xi = x + L.phase;
    const double cx = cos(PI * x / 2 / L.xs), cy = cos(PI * y / 2 / L.ys), cz = cos(PI * z / 2 / L.zs);
    const double sx = sin(PI * x / 2 / L.xs), sy = sin(PI * y / 2 / L.ys), sz = sin(PI * z / 2 / L.zs);
    const double sxi = sin(xi), cxi = cos(xi);
    const double tr = cy * cy * cz * cz;
    const double fy = L.a0y * tr * (cxi * cx * cx - PI / L.xs * sxi * cx * sx);
    const double fz = L.a0z * tr * (sxi * cx * cx + PI / L.xs * cxi * cx * sx);
    const double fx = L.a0y * sxi * cx * cx * cz * cz * PI / L.ys * cy * sy - L.a0z * cxi * cx * cx * cy * cy * PI / L.zs * cz * sz;
    double gx = L.a0z * cxi * cx * cx * PI / L.ys * cz * cz * cy * sy + L.a0y * sxi * cx * cx * cy * cy * PI / L.zs * cz * sz;
    double gy = -fz, gz = fy;
    if (!L.b_sign) { gx = -gx; gy = -gy; gz = -gz; }
    for (int m = 0; m < 3; ++m) {
        ev[m] = fx * L.r0[m] + fy * L.ey[m] + fz * L.ez[m];
        bv[m] = gx * L.r0[m] + gy * L.ey[m] + gz * L.ez[m];
    }
    return true;
}

// Laser initial condition f_init_cos (src/init.cpp:6-210), evaluated in FP64
// and stored in the field precision
template <typename real>
__global__ void __launch_bounds__(kBX * kBY) k_laser_cos(Grid g, Fields<real> f, LaserFrame L)
{
    const int k = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= g.nz || j >= g.ny) return;
    const int id = i * g.sx + j * g.sy + k;
    real *E[3] = {f.ex, f.ey, f.ez}, *B[3] = {f.bx, f.by, f.bz};
    double ev[3], bv[3];
    if (pulse_at(L, i * g.dx, j * g.dy, k * g.dz, 0.0, ev, bv)) {
        for (int m = 0; m < 3; ++m) {
            if (L.append) {
                E[m][id] = (real)((double)E[m][id] + ev[m]);
                B[m][id] = (real)((double)B[m][id] + bv[m]);
            } else {
                E[m][id] = (real)ev[m];
                B[m][id] = (real)bv[m];
            }
        }
    } else if (!L.append) {
        for (int m = 0; m < 3; ++m) { E[m][id] = real(0); B[m][id] = real(0); }
    }
}

// Total-field / scattered-field laser source (qk_add_laser_source).  On the
// source's total side of its plane the grid holds the total field, on the
// other side the field scattered off the incident pulse; an update reading
// across the plane is corrected by the same update applied to the incident
// field on the far side: + L(inc on the scattered side) at total nodes,
// - L(inc on the total side) at scattered nodes, L being the solver's own
// update (k_maxwell_b / k_maxwell_e) with its boundary-plane sets.  Both
// passes run as one launch on a stack of two copies of the W planes next to
// the source's wall, separated by kInjGap empty planes: the copy whose
// outputs are read at scattered nodes sits at the wall's end of the stack,
// so its planes near the wall see the grid's own boundary-plane sets.  The
// incident field is sampled where each component lives on the staggered
// grid (E_a a half cell along a, B_a a half cell across a).

// the slab of the stack at plane l: 0 low, 1 high, -1 gap; sets the grid plane
__device__ inline int stack_slab(const InjectArgs &a, int l, int &i)
{
    if (l < a.W) { i = a.main_lo + l; return 0; }
    if (l >= a.W + kInjGap) { i = a.main_lo + l - a.W - kInjGap; return 1; }
    return -1;
}

__device__ inline bool total_side(const InjectArgs &a, int i) { return a.side == 0 ? i >= a.plane : i < a.plane; }

// fill the stack's input component (E for a B half step, B for an E step)
// with the incident field of every source, masked per slab, and clear the
// output components
template <typename real>
__global__ void __launch_bounds__(kBX * kBY) k_inject_prepare(Grid gs, Fields<real> st, InjectArgs a)
{
    const int k = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int l = blockIdx.z;
    if (k >= gs.nz || j >= gs.ny) return;
    const int id = l * gs.sx + j * gs.sy + k;
    real *in[3] = {a.b_field ? st.ex : st.bx, a.b_field ? st.ey : st.by, a.b_field ? st.ez : st.bz};
    real *out[3] = {a.b_field ? st.bx : st.ex, a.b_field ? st.by : st.ey, a.b_field ? st.bz : st.ez};
    int i = 0;
    const int slab = stack_slab(a, l, i);
    // the low slab holds the pass read at scattered nodes when the wall is
    // at x-min, the pass read at total nodes when it is at x-max
    const bool reads_at_total = (slab == 1) == (a.side == 0);
    // pass read at total nodes: incident field on scattered nodes, and back
    const bool fill = slab >= 0 && abs(i - a.plane) <= 4 && total_side(a, i) != reads_at_total;
    double v[3] = {0, 0, 0};
    if (fill) {
        for (int n = 0; n < a.nsrc; ++n)
            for (int c = 0; c < 3; ++c) {
                // staggered position of component c of the input field
                double h[3];
                for (int m = 0; m < 3; ++m) h[m] = (a.b_field ? m == c : m != c) ? 0.5 : 0.0;
                double ev[3], bv[3];
                if (pulse_at(a.src[n], (i + a.ishift[n] + h[0]) * gs.dx, (j + h[1]) * gs.dy, (k + h[2]) * gs.dz,
                             a.shift[n], ev, bv))
                    v[c] += a.b_field ? ev[c] : bv[c];
            }
    }
    for (int c = 0; c < 3; ++c) {
        in[c][id] = (real)v[c];
        out[c][id] = real(0);
    }
}

// add the stack's corrections to the grid's planes within 3 of the source plane
template <typename real>
__global__ void __launch_bounds__(kBX * kBY) k_inject_apply(Grid g, Fields<real> f, Grid gs, Fields<real> st,
                                                            InjectArgs a)
{
    const int k = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const int i = a.plane - 3 + (int)blockIdx.z;
    if (k >= g.nz || j >= g.ny || i < 0 || i >= g.nx) return;
    const bool tot = total_side(a, i);
    // the slab read here: at total nodes the high slab for an x-min wall
    const bool high = tot == (a.side == 0);
    const int l = (i - a.main_lo) + (high ? a.W + kInjGap : 0);
    const int id = i * g.sx + j * g.sy + k, ls = l * gs.sx + j * gs.sy + k;
    real *dst[3] = {a.b_field ? f.bx : f.ex, a.b_field ? f.by : f.ey, a.b_field ? f.bz : f.ez};
    const real *src[3] = {a.b_field ? st.bx : st.ex, a.b_field ? st.by : st.ey, a.b_field ? st.bz : st.ez};
    for (int c = 0; c < 3; ++c) {
        const real d = src[c][ls];
        dst[c][id] = tot ? dst[c][id] + d : dst[c][id] - d;
    }
}

template <typename real>
Coef<real> make_coef(const Grid &g, const SolverCoef &c)
{
    Coef<real> r;
    const double d[3] = {g.dx, g.dy, g.dz};
    for (int a = 0; a < 3; ++a) {
        r.a[a] = (real)c.a[a];
        r.b[a] = (real)c.b[a];
        r.dtd[a] = (real)(g.dt / d[a]);
        r.d[a] = (real)d[a];
    }
    return r;
}

__global__ void k_zero(double *p, size_t n)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

}  // namespace

hipError_t launch_zero(hipStream_t st, double *p, size_t n)
{
    if (n == 0) return hipSuccess;
    k_zero<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_maxwell_b(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, const Fields<real> &f,
                            double *energy)
{
    const Coef<real> cf = make_coef<real>(g, c);
    const dim3 grid = cell_grid(g), block(kBX, kBY);
    if (solver == 0) k_maxwell_b<0, real><<<grid, block, 0, st>>>(g, cf, f.ex, f.ey, f.ez, f.bx, f.by, f.bz, energy);
    else if (solver == 1) k_maxwell_b<1, real><<<grid, block, 0, st>>>(g, cf, f.ex, f.ey, f.ez, f.bx, f.by, f.bz, energy);
    else k_maxwell_b<2, real><<<grid, block, 0, st>>>(g, cf, f.ex, f.ey, f.ez, f.bx, f.by, f.bz, energy);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_maxwell_e(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, const Fields<real> &f,
                            const Current<real> &j, const Current<real> &jn)
{
    const Coef<real> cf = make_coef<real>(g, c);
    const dim3 grid = cell_grid(g), block(kBX, kBY);
#define QK_E_ARGS g, cf, f.ex, f.ey, f.ez, f.bx, f.by, f.bz, j.jx, j.jy, j.jz, jn.jx, jn.jy, jn.jz
    if (solver == 0) k_maxwell_e<0, real><<<grid, block, 0, st>>>(QK_E_ARGS);
    else if (solver == 1) k_maxwell_e<1, real><<<grid, block, 0, st>>>(QK_E_ARGS);
    else k_maxwell_e<2, real><<<grid, block, 0, st>>>(QK_E_ARGS);
#undef QK_E_ARGS
    return hipGetLastError();
}

template <typename real>
hipError_t launch_interpolate_be(hipStream_t st, const Grid &g, const Fields<real> &f)
{
    constexpr int R = BeRun<real>::n;
    const dim3 grid((g.nz + kBeK - 1) / kBeK, (g.ny + kBeJ - 1) / kBeJ, (g.nx + R - 1) / R);
    k_interpolate_be<real><<<grid, dim3(kBeK, kBeJ), 0, st>>>(g, f.bx, f.by, f.bz, f.bex, f.bey, f.bez);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_init_boundaries(hipStream_t st, int solver, const Grid &g, const Fields<real> &f)
{
    k_init_boundaries<real><<<cell_grid(g), dim3(kBX, kBY), 0, st>>>(g, solver, f);
    return hipGetLastError();
}

bool laser_frame(const Grid &g, const double *P, LaserFrame &L)
{
    // propagation frame (src/init.cpp:32-93)
    const double PI = 3.141592653589793;
    L.a0y = P[0];
    L.a0z = P[1];
    L.xs = P[2] * 2 * sqrt(2 * PI) / 3;
    L.ys = P[3] * 2 * sqrt(2 * PI) / 3;
    L.zs = P[4] * 2 * sqrt(2 * PI) / 3;
    L.x0 = P[5];
    L.b_sign = P[6] != 0;
    const double x1 = P[7];
    L.phase = P[8];
    L.y0 = P[9];
    L.z0 = P[10];
    L.append = P[11] != 0;
    const double phi = P[12];
    L.xt = P[13];
    L.yt = P[14];
    L.zt = P[15];
    L.y12 = 0.5 * g.ny * g.dy;
    L.z12 = 0.5 * g.nz * g.dz;
    const double rn = sqrt(x1 * x1 + L.y0 * L.y0 + L.z0 * L.z0);
    if (!(rn > 0)) return false;
    L.r0[0] = -x1 / rn;
    L.r0[1] = -L.y0 / rn;
    L.r0[2] = -L.z0 / rn;
    const double t = sqrt(L.r0[0] * L.r0[0] + L.r0[1] * L.r0[1]);
    if (t != 0) { L.ey[0] = -L.r0[1] / t; L.ey[1] = L.r0[0] / t; L.ey[2] = 0; }
    else { L.ey[0] = 0; L.ey[1] = 1; L.ey[2] = 0; }
    L.ez[0] = L.r0[1] * L.ey[2] - L.r0[2] * L.ey[1];
    L.ez[1] = L.r0[2] * L.ey[0] - L.r0[0] * L.ey[2];
    L.ez[2] = L.r0[0] * L.ey[1] - L.r0[1] * L.ey[0];
    if (phi != 0) {
        const double c = cos(phi), s = sin(phi);
        double y[3], z[3];
        for (int m = 0; m < 3; ++m) { y[m] = L.ey[m]; z[m] = L.ez[m]; }
        for (int m = 0; m < 3; ++m) { L.ey[m] = y[m] * c + z[m] * s; L.ez[m] = z[m] * c - y[m] * s; }
    }
    return true;
}

template <typename real>
hipError_t launch_laser_cos(hipStream_t st, const Grid &g, const Fields<real> &f, const double *P)
{
    LaserFrame L;
    if (!laser_frame(g, P, L)) return hipErrorInvalidValue;
    k_laser_cos<real><<<cell_grid(g), dim3(kBX, kBY), 0, st>>>(g, f, L);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_inject(hipStream_t st, int solver, const Grid &g, const SolverCoef &c, const Fields<real> &f,
                         const Grid &gs, const Fields<real> &stack, const Current<real> &j0, const Current<real> &j1,
                         const InjectArgs &a)
{
    k_inject_prepare<real><<<cell_grid(gs), dim3(kBX, kBY), 0, st>>>(gs, stack, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = a.b_field ? launch_maxwell_b<real>(st, solver, gs, c, stack, nullptr)
                  : launch_maxwell_e<real>(st, solver, gs, c, stack, j0, j1);
    if (e != hipSuccess) return e;
    const dim3 grid((g.nz + kBX - 1) / kBX, (g.ny + kBY - 1) / kBY, 7);   // planes plane-3 .. plane+3
    k_inject_apply<real><<<grid, dim3(kBX, kBY), 0, st>>>(g, f, gs, stack, a);
    return hipGetLastError();
}

#define QK_INST(real)                                                                                            \
    template hipError_t launch_maxwell_b<real>(hipStream_t, int, const Grid &, const SolverCoef &, const Fields<real> &, \
                                               double *);                                                       \
    template hipError_t launch_maxwell_e<real>(hipStream_t, int, const Grid &, const SolverCoef &, const Fields<real> &, \
                                               const Current<real> &, const Current<real> &);                   \
    template hipError_t launch_interpolate_be<real>(hipStream_t, const Grid &, const Fields<real> &);           \
    template hipError_t launch_init_boundaries<real>(hipStream_t, int, const Grid &, const Fields<real> &);     \
    template hipError_t launch_laser_cos<real>(hipStream_t, const Grid &, const Fields<real> &, const double *); \
    template hipError_t launch_inject<real>(hipStream_t, int, const Grid &, const SolverCoef &, const Fields<real> &, \
                                            const Grid &, const Fields<real> &, const Current<real> &,           \
                                            const Current<real> &, const InjectArgs &);
QK_INST(float)
QK_INST(double)
#undef QK_INST

}  // namespace qk
