// qk_window.hip -- the moving window (spatial_region::moving_window,
// src/advance.cpp:389-454), its field-energy probe (get_max_w,
// src/low-level.cpp:512-532) and the lattice particle fill of new columns
// (fill_cell_by_particles, src/init.cpp:402-478).
#include <algorithm>

#include "qk_kernels.h"

namespace qk {

namespace {

constexpr int kWBlock = 256;

int64_t wblocks(int64_t n) { return (n + kWBlock - 1) / kWBlock; }

// the nine field components that move with the window (E, B and Be; J is
// zeroed by the next particle phase and does not move)
template <typename real>
struct Shifted {
    real *c[9];
};

// One thread per (component, j, k) column walks i upward: F[i] = F[i + 1]
// for i < nx - 1, then F[nx - 1] = 0.  A column is owned by one thread and
// every plane is read before it is overwritten, so the shift is in place.
// Reads are issued kRun planes ahead of the writes.
constexpr int kRun = 8;

template <typename real>
__global__ void __launch_bounds__(kWBlock) k_shift_fields(Shifted<real> f, Grid g)
{
    const int64_t t = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    const int64_t ncol = (int64_t)g.ny * g.nz;
    if (t >= ncol) return;
    real *__restrict__ F = f.c[blockIdx.y] + t;
    const int64_t sx = g.sx;
    int i = 0;
    for (; i + kRun < g.nx; i += kRun) {
        real v[kRun];
#pragma unroll
        for (int u = 0; u < kRun; ++u) v[u] = F[(i + u + 1) * sx];
#pragma unroll
        for (int u = 0; u < kRun; ++u) F[(i + u) * sx] = v[u];
    }
    for (; i + 1 < g.nx; ++i) F[i * sx] = F[(i + 1) * sx];
    F[(int64_t)(g.nx - 1) * sx] = real(0);
}

// particles: those in the plane i = 0 are erased, the others move one cell
// toward -x (x -= 1 is the cell index minus one; the offset is unchanged)
__global__ void __launch_bounds__(kWBlock) k_shift_particles(int *cell, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    if (p >= n) return;
    const int c = cell[p];
    if (c < 0) return;
    cell[p] = cell_i(c) == 0 ? -1 : c - (1 << 20);
}

// max over i in [ilo, ihi) of E^2 + B^2 at a node (components taken at their
// own staggered positions, as the reference does).  The value is >= 0, so its
// IEEE bit pattern orders like an unsigned integer: one atomic max per
// workgroup on the bits.
template <typename real>
__global__ void __launch_bounds__(kWBlock) k_max_w(Fields<real> f, int64_t first, int64_t n, unsigned long long *out)
{
    __shared__ double part[kWBlock / 64];
    double m = 0;
    for (int64_t t = (int64_t)blockIdx.x * kWBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kWBlock) {
        const int64_t o = first + t;
        const double ex = f.ex[o], ey = f.ey[o], ez = f.ez[o], bx = f.bx[o], by = f.by[o], bz = f.bz[o];
        const double w = ex * ex + ey * ey + ez * ez + bx * bx + by * by + bz * bz;
        m = fmax(m, w);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = part[0];
        for (int w = 1; w < kWBlock / 64; ++w) b = fmax(b, part[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(b));
    }
}

// fill_cell_by_particles with T = 0 and no displacement: bi x bj x bk
// particles per listed cell at (0.5 / b + cell + index / b) on each axis,
// u = (ux0, 0, 0), charge q0 = density / (bi bj bk) (negative for cmr < 0).
// Thread t fills particle t % npt of listed cell t / npt into slot first + t.
template <typename real>
__global__ void __launch_bounds__(kWBlock) k_fill_cells(Species<real> s, int64_t first, int64_t total, const int *cells,
                                                        const double *density, const int *idbase, int bi, int bj, int bk,
                                                        double sign, double ux0)
{
    const int64_t t = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    if (t >= total) return;
    const int npt = bi * bj * bk;
    const int64_t m = t / npt;
    const int r = (int)(t - m * npt);
    const int kp = r % bk, jp = (r / bk) % bj, ip = r / (bk * bj);
    const int i = cells[3 * m], j = cells[3 * m + 1], k = cells[3 * m + 2];
    // the reference's expressions, in its order of evaluation
    const double x = (0.5 / bi + double(i)) + double(ip) / double(bi);
    const double y = (0.5 / bj + double(j)) + double(jp) / double(bj);
    const double z = (0.5 / bk + double(k)) + double(kp) / double(bk);
    const double q0 = 1 / double(npt) * density[m];
    const int64_t o = first + t;
    s.cell[o] = pack_cell(i, j, k);
    real ox = (real)(x - i), oy = (real)(y - j), oz = (real)(z - k);
    // offsets stay below 1 after rounding to the storage type
    const real below1 = sizeof(real) == 4 ? (real)0.99999994f : (real)0.9999999999999999;
    ox = ox < real(1) ? ox : below1;
    oy = oy < real(1) ? oy : below1;
    oz = oz < real(1) ? oz : below1;
    s.ox[o] = ox;
    s.oy[o] = oy;
    s.oz[o] = oz;
    s.ux[o] = (real)ux0;
    s.uy[o] = real(0);
    s.uz[o] = real(0);
    s.q[o] = (real)(sign * q0);
    if (s.chi) s.chi[o] = real(0);
    if (s.id) s.id[o] = idbase[m] + r;
}

}  // namespace

template <typename real>
hipError_t launch_shift_fields(hipStream_t st, const Grid &g, const Fields<real> &f)
{
    Shifted<real> sh{{f.ex, f.ey, f.ez, f.bx, f.by, f.bz, f.bex, f.bey, f.bez}};
    const int64_t ncol = (int64_t)g.ny * g.nz;
    k_shift_fields<real><<<dim3((unsigned)wblocks(ncol), 9), kWBlock, 0, st>>>(sh, g);
    return hipGetLastError();
}

hipError_t launch_shift_particles(hipStream_t st, int *cell, int64_t n)
{
    if (n <= 0) return hipSuccess;
    k_shift_particles<<<wblocks(n), kWBlock, 0, st>>>(cell, n);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_max_w(hipStream_t st, const Grid &g, const Fields<real> &f, int ilo, int ihi,
                        unsigned long long *out)
{
    const int64_t first = (int64_t)ilo * g.sx, n = (int64_t)(ihi - ilo) * g.sx;
    if (n <= 0) return hipSuccess;
    const int64_t nb = std::min<int64_t>(wblocks(n), 2048);
    k_max_w<real><<<nb, kWBlock, 0, st>>>(f, first, n, out);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_fill_cells(hipStream_t st, const Species<real> &s, int64_t first, int64_t ncells, const int *cells,
                             const double *density, const int *idbase, const int npic[3], double sign, double ux0)
{
    const int64_t total = ncells * npic[0] * npic[1] * npic[2];
    if (total <= 0) return hipSuccess;
    k_fill_cells<real><<<wblocks(total), kWBlock, 0, st>>>(s, first, total, cells, density, idbase, npic[0], npic[1],
                                                           npic[2], sign, ux0);
    return hipGetLastError();
}

#define QK_INST(real)                                                                                              \
    template hipError_t launch_shift_fields<real>(hipStream_t, const Grid &, const Fields<real> &);                \
    template hipError_t launch_max_w<real>(hipStream_t, const Grid &, const Fields<real> &, int, int,              \
                                           unsigned long long *);                                                  \
    template hipError_t launch_fill_cells<real>(hipStream_t, const Species<real> &, int64_t, int64_t, const int *, \
                                                const double *, const int *, const int *, double, double);
QK_INST(float)
QK_INST(double)
#undef QK_INST

}  // namespace qk
