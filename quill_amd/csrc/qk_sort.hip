// qk_sort.hip -- cell-tile counting sort of one particle species on gfx950
// (replaces the per-cell linked lists and place() re-linking of
// src/plist.cpp:7-40 and the page allocator of src/low-level.cpp:186-286).
//
//   1. count live particles per cell tile
//   2. exclusive scan of the counts -> first slot of every tile
//   3. per tile: the number of kChunk-particle chunks and the spread
//      multiplier of its storage order
//   4. scatter every live particle to the next position of its tile in the
//      other buffer, stored at that tile's spread slot (free slots are
//      dropped: this is also the compaction)
//   5. chunk counts scanned into the work list of the tiled push kernel
//
// Particles arrive in the order of the previous sort, so the wavefronts of
// the count and scatter passes see long runs of one tile and need about one
// atomic per run.
#include "qk_kernels.h"

namespace qk {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ void k_fill_int(int *p, size_t n, int v)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// exclusive scan of kScanTile items per workgroup; block totals to sums
__global__ void __launch_bounds__(kScanBlock) k_scan_tile(const int *__restrict__ in, int *__restrict__ out, int n,
                                                          int *__restrict__ sums)
{
    __shared__ int wsum[kScanBlock / 64];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int local = 0;
#pragma unroll
    for (int m = 0; m < kScanItems; ++m) {
        v[m] = (base + m < n) ? in[base + m] : 0;
        local += v[m];
    }
    // inclusive scan of the thread totals inside the wavefront
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int wprefix = 0;
    for (int w = 0; w < wave; ++w) wprefix += wsum[w];
    int run = wprefix + x - local;   // exclusive prefix of this thread
#pragma unroll
    for (int m = 0; m < kScanItems; ++m) {
        if (base + m < n) out[base + m] = run;
        run += v[m];
    }
    if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = wprefix + x;
}

__global__ void k_scan_add(int *out, int n, const int *__restrict__ offs)
{
    const int t = blockIdx.x * kScanTile + threadIdx.x;
    const int add = offs[blockIdx.x];
    for (int m = 0; m < kScanItems; ++m) {
        const int id = t + m * kScanBlock;
        if (id < n) out[id] += add;
    }
}

// Runs of equal keys in consecutive lanes of a wavefront (particles arrive
// nearly sorted, so most of a wavefront shares a few keys): every lane learns
// the first lane of its run and the run length, and one atomic per run
// replaces one per particle.
__device__ inline void key_runs(int key, int lane, int &head, int &len)
{
    const int prev = __shfl_up(key, 1, 64);
    const bool is_head = lane == 0 || prev != key;
    const unsigned long long mask = __ballot(is_head);
    const unsigned long long at_or_below = mask & ((2ull << lane) - 1ull);   // lane 63: all bits
    head = 63 - __clzll(at_or_below);
    const unsigned long long above = lane == 63 ? 0ull : (mask >> (lane + 1));
    const int next = above ? lane + __ffsll(above) : 64;
    len = next - head;
}

__global__ void __launch_bounds__(256) k_count_keys(const int *__restrict__ cell, int64_t n, TileGeom tg, int *counts)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = p < n ? cell[p] : -1;
    const int key = c >= 0 ? sort_key(tg, c) : -1;
    const int lane = threadIdx.x & 63;
    int head, len;
    key_runs(key, lane, head, len);
    if (lane == head && key >= 0) atomicAdd(counts + key, len);
}

// The scatter reads its source in XCD-sized contiguous ranges (xcd_block):
// the particles of one tile, contiguous in the source, are then written by
// one XCD, whose L2 merges their scattered stores into whole lines instead
// of collecting partial lines in eight L2s.
template <typename real>
__global__ void __launch_bounds__(256) k_scatter(Species<real> src, Species<real> dst, int64_t n, TileGeom tg,
                                                 int *cursor, const int *__restrict__ tile_start,
                                                 const int *__restrict__ tile_mult)
{
    const int64_t p = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int c = p < n ? src.cell[p] : -1;
    const int key = c >= 0 ? sort_key(tg, c) : -1;
    const int lane = threadIdx.x & 63;
    int head, len;
    key_runs(key, lane, head, len);
    int base = 0;
    if (lane == head && key >= 0) base = atomicAdd(cursor + key, len);
    base = __shfl(base, head, 64);
    if (c < 0) return;
    // sorted position inside the tile -> its storage slot
    const int t = key;
    const int ts = tile_start[t], nt = tile_start[t + 1] - ts;
    const int o = ts + (int)((int64_t)(base + (lane - head) - ts) * tile_mult[t] % nt);
    dst.cell[o] = c;
    dst.ox[o] = src.ox[p];
    dst.oy[o] = src.oy[p];
    dst.oz[o] = src.oz[p];
    dst.ux[o] = src.ux[p];
    dst.uy[o] = src.uy[p];
    dst.uz[o] = src.uz[p];
    dst.q[o] = src.q[p];
    if (src.chi) dst.chi[o] = src.chi[p];
    if (src.id) dst.id[o] = src.id[p];
}

// Storage order inside a tile.  The n particles of a tile, in arrival order,
// are stored so that slot s holds arrival position (s * a) mod n, with a the
// first of kSpread coprime to n.  A wavefront reading 64 consecutive slots
// then gets particles a arrival positions apart; particles that arrive
// grouped by cell (a loader writes a cell's particles together) land in
// distinct cells, and already spread ones stay spread.  The push kernel's
// LDS atomics then hit distinct addresses while its memory accesses stay
// contiguous.  The scatter applies the inverse map, position e -> slot
// e * a^-1 mod n.
constexpr int kSpread[8] = {37, 41, 43, 47, 53, 59, 61, 67};

// a^-1 mod n for gcd(a, n) = 1 (extended Euclid); 0 for n = 1
__device__ inline int mod_inverse(int a, int n)
{
    int64_t t = 0, nt = 1, r = n, nr = a % n;
    while (nr != 0) {
        const int64_t q = r / nr;
        const int64_t t2 = t - q * nt;
        t = nt;
        nt = t2;
        const int64_t r2 = r - q * nr;
        r = nr;
        nr = r2;
    }
    return (int)(t < 0 ? t + n : t);
}

__global__ void k_tile_chunks(const int *__restrict__ offsets, int ntiles, int *tile_start, int *chunk_n,
                              int *tile_mult)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    const int s = offsets[t];
    tile_start[t] = s;
    if (t < ntiles) {
        const int cnt = offsets[t + 1] - s;
        chunk_n[t] = (cnt + kChunk - 1) / kChunk;
        int m = 1;   // identity if every spread shares a factor with cnt
        for (int a : kSpread)
            if (cnt > 1 && cnt % a != 0) {
                m = mod_inverse(a, cnt);
                break;
            }
        tile_mult[t] = m;
    } else {
        chunk_n[t] = 0;
    }
}

__global__ void k_write_chunks(const int *__restrict__ tile_start, const int *__restrict__ chunk_off, int ntiles,
                               int4 *chunks, int max_chunks)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int s = tile_start[t], e = tile_start[t + 1];
    int o = chunk_off[t];
    for (int p = s; p < e && o < max_chunks; p += kChunk, ++o) chunks[o] = make_int4(t, p, min(kChunk, e - p), 0);
}

inline unsigned nblocks(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t launch_fill_int(hipStream_t st, int *p, size_t n, int v)
{
    if (n == 0) return hipSuccess;
    k_fill_int<<<nblocks(n, 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

size_t scan_scratch_ints(size_t n)
{
    // mirrors the recursion of launch_exclusive_scan: a level over m items
    // uses 2 (nb + 1) ints and recurses on nb items
    size_t total = 0, m = n + 1;
    for (;;) {
        const size_t nb = (m + kScanTile - 1) / kScanTile;
        total += 2 * (nb + 1);
        if (nb == 1) break;
        m = nb + 1;
    }
    return total + 16;
}

// out[0..n] = exclusive prefix sums of in[0..n), out[n] = total
hipError_t launch_exclusive_scan(hipStream_t st, const int *in, int *out, int n, int *scratch)
{
    // the total lands in out[n] by scanning n + 1 items with a trailing zero;
    // in[n] must be readable and zero (callers size their arrays n + 1)
    const int m = n + 1;
    const int nb = (m + kScanTile - 1) / kScanTile;
    int *sums = scratch, *sums_scan = scratch + nb + 1;
    k_scan_tile<<<nb, kScanBlock, 0, st>>>(in, out, m, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nb == 1) return e;
    // the block totals are scanned recursively; their array gets a zero tail
    k_fill_int<<<1, 1, 0, st>>>(sums + nb, 1, 0);
    e = launch_exclusive_scan(st, sums, sums_scan, nb, scratch + 2 * (nb + 1));
    if (e != hipSuccess) return e;
    k_scan_add<<<nb, kScanBlock, 0, st>>>(out, m, sums_scan);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_sort(hipStream_t st, const TileGeom &tg, const Species<real> &src, const Species<real> &dst,
                       int64_t n, const SortBuffers &b)
{
    hipError_t e = launch_fill_int(st, b.counts, (size_t)tg.nkeys + 1, 0);
    if (e != hipSuccess) return e;
    if (n > 0) {
        k_count_keys<<<nblocks((size_t)n, 256), 256, 0, st>>>(src.cell, n, tg, b.counts);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = launch_exclusive_scan(st, b.counts, b.offsets, tg.nkeys, b.scratch)) != hipSuccess) return e;
    k_tile_chunks<<<nblocks((size_t)tg.ntiles + 1, 256), 256, 0, st>>>(b.offsets, tg.ntiles, b.tile_start, b.chunk_n,
                                                                       b.tile_mult);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (n > 0) {
        k_scatter<real><<<nblocks((size_t)n, 256), 256, 0, st>>>(src, dst, n, tg, b.offsets, b.tile_start,
                                                                 b.tile_mult);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = launch_exclusive_scan(st, b.chunk_n, b.chunk_off, tg.ntiles, b.scratch)) != hipSuccess) return e;
    k_write_chunks<<<nblocks((size_t)tg.ntiles, 256), 256, 0, st>>>(b.tile_start, b.chunk_off, tg.ntiles, b.chunks,
                                                                    b.max_chunks);
    return hipGetLastError();
}

template hipError_t launch_sort<float>(hipStream_t, const TileGeom &, const Species<float> &, const Species<float> &,
                                       int64_t, const SortBuffers &);
template hipError_t launch_sort<double>(hipStream_t, const TileGeom &, const Species<double> &,
                                        const Species<double> &, int64_t, const SortBuffers &);

}  // namespace qk
