// qk_sort.hip -- cell-tile counting sort of one particle species on gfx950
// (replaces the per-cell linked lists and place() re-linking of
// src/plist.cpp:7-40 and the page allocator of src/low-level.cpp:186-286).
//
//   1. count live particles per (tile, cell) key
//   2. exclusive scan of the counts -> first slot of every key
//   3. per tile: the number of kChunk-particle chunks and the spread
//      multiplier of its storage order
//   4. scatter every live particle to the next position of its key in the
//      other buffer, stored at its tile's spread slot (free slots are
//      dropped: this is also the compaction)
//   5. chunk counts scanned into the work list of the tiled push kernel
//
// The first sort (and any sort whose source is not the previous sort's
// output) orders by (tile, cell) with one counter per cell and stores each
// tile in its spread order.  Later sorts keep that order: the source is then
// the previous sort's output, each tile's particles contiguous, and between
// sorts few change tile.  One workgroup per previous tile compacts the
// particles that stayed, in their order, to the front of the tile's new range
// (a block-wide prefix count, no atomics); those that changed tile, and those
// appended since the last sort, take the positions after the stayers of
// their new tile from one counter per tile.  Stayers keep the spread order,
// and the writes of one workgroup are one contiguous run.
//
// The one-pass scatter with a counter per cell, run on an already sorted
// source, returns 64 atomics from L2 per wavefront (the spread order puts 64
// distinct cells in a wavefront): 2.3 ms for 31.6 M particles with 3.2x its
// payload written as partial lines; ranking the stayers per cell in LDS and
// writing them at their spread slots was slower still (4.1 ms, 7.2x).

#include <algorithm>

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

__global__ void __launch_bounds__(256) k_count_keys(const int *__restrict__ cell, int64_t first, int64_t n, TileGeom tg,
                                                    int *counts)
{
    const int64_t p = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = p < n ? cell[p] : -1;
    const int key = c >= 0 ? sort_key(tg, c) : -1;
    const int lane = threadIdx.x & 63;
    int head, len;
    key_runs(key, lane, head, len);
    if (lane == head && key >= 0) atomicAdd(counts + key, len);
}

constexpr int kSortBlock = 512;   // per step 0.4 % faster than 256, 0.5 % than 1024

// one record in registers, loaded before its destination is known
template <typename real>
struct HeldRecord {
    real o[3], u[3], q, chi;
    int id;
    __device__ inline void load(const Species<real> &src, int64_t p)
    {
        o[0] = src.ox[p];
        o[1] = src.oy[p];
        o[2] = src.oz[p];
        u[0] = src.ux[p];
        u[1] = src.uy[p];
        u[2] = src.uz[p];
        q = src.q[p];
        chi = src.chi ? src.chi[p] : real(0);
        id = src.id ? src.id[p] : 0;
    }
    __device__ inline void store(const Species<real> &dst, int c, int at) const
    {
        dst.cell[at] = c;
        dst.ox[at] = o[0];
        dst.oy[at] = o[1];
        dst.oz[at] = o[2];
        dst.ux[at] = u[0];
        dst.uy[at] = u[1];
        dst.uz[at] = u[2];
        dst.q[at] = q;
        if (dst.chi) dst.chi[at] = chi;
        if (dst.id) dst.id[at] = id;
    }
};


__device__ inline int tile_of_cell(const TileGeom &tg, int c)
{
    return tile_of(tg, c >> 20, (c >> 10) & 1023, c & 1023);
}

// A particle that changed tile has nearly always moved to one of the 26
// neighbours: movers are binned by the neighbour in LDS and each bin takes
// one global atomic, instead of one per particle on a few hot counters.
// Returns the bin (0..26) of tile tn seen from t, or -1 when farther.
__device__ inline int neighbour_bin(const TileGeom &tg, int t, int tn)
{
    const int tz = t % tg.ntz, ty = (t / tg.ntz) % tg.nty, tx = t / (tg.ntz * tg.nty);
    const int uz = tn % tg.ntz, uy = (tn / tg.ntz) % tg.nty, ux = tn / (tg.ntz * tg.nty);
    const int dx = ux - tx, dy = uy - ty, dz = uz - tz;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
    return ((dx + 1) * 3 + (dy + 1)) * 3 + (dz + 1);
}

// the tile in bin b of t (the bin exists only for tiles inside the grid)
__device__ inline int bin_tile(const TileGeom &tg, int t, int b)
{
    return t + ((b / 9 - 1) * tg.nty + (b / 3 % 3 - 1)) * tg.ntz + (b % 3 - 1);
}

// re-sort, count: the particles of previous tile t = blockIdx.x that stay
// (stay[t]) and, per new tile, those that move there (imm[], global)
__global__ void __launch_bounds__(kSortBlock) k_count_keep(const int *__restrict__ cell,
                                                           const int *__restrict__ old_start, TileGeom tg,
                                                           int *__restrict__ stay, int *imm)
{
    __shared__ int wsum[kSortBlock / 64];
    __shared__ int bins[27];
    const int t = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x < 27) bins[threadIdx.x] = 0;
    __syncthreads();
    const int s = old_start[t], e = old_start[t + 1];
    int kept = 0;
    // kCountIt cells per thread and pass, all loaded before the first is
    // used, so that a thread waits for memory once per pass
    constexpr int kCountIt = 4;
    for (int p0 = s + threadIdx.x; p0 < e; p0 += kCountIt * kSortBlock) {
        int cs[kCountIt];
#pragma unroll
        for (int m = 0; m < kCountIt; ++m) {
            const int p = p0 + m * kSortBlock;
            cs[m] = p < e ? cell[p] : -1;
        }
#pragma unroll
        for (int m = 0; m < kCountIt; ++m) {
            const int c = cs[m];
            if (c < 0) continue;
            const int tn = tile_of_cell(tg, c);
            if (tn == t) {
                ++kept;
            } else {
                const int b = neighbour_bin(tg, t, tn);
                if (b >= 0) atomicAdd(&bins[b], 1);
                else atomicAdd(imm + tn, 1);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) kept += __shfl_xor(kept, o);
    if (lane == 0) wsum[w] = kept;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int k = 0; k < kSortBlock / 64; ++k) sum += wsum[k];
        stay[t] = sum;
    }
    if (threadIdx.x < 27 && bins[threadIdx.x] > 0) atomicAdd(imm + bin_tile(tg, t, threadIdx.x), bins[threadIdx.x]);
}

// re-sort, appended particles: imm[tile] += 1 per run of equal tiles
__global__ void __launch_bounds__(256) k_count_tail(const int *__restrict__ cell, int64_t first, int64_t n, TileGeom tg,
                                                    int *imm)
{
    const int64_t p = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = p < n ? cell[p] : -1;
    const int key = c >= 0 ? tile_of_cell(tg, c) : -1;
    const int lane = threadIdx.x & 63;
    int head, len;
    key_runs(key, lane, head, len);
    if (lane == head && key >= 0) atomicAdd(imm + key, len);
}

// counts = stay + imm (into imm), the input of the scan
__global__ void k_sum_counts(const int *__restrict__ stay, int *imm, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) imm[k] += stay[k];
}

// re-sort, layout: tile_start = scan of stay + imm (done by the caller);
// the movers' counters start after each tile's stayers; the chunk counts
__global__ void k_keep_layout(const int *__restrict__ tile_start, const int *__restrict__ stay, int ntiles, int *cur,
                              int *chunk_n, int *tile_mult)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    if (t == ntiles) {
        chunk_n[t] = 0;
        return;
    }
    cur[t] = tile_start[t] + stay[t];
    const int cnt = tile_start[t + 1] - tile_start[t];
    chunk_n[t] = (cnt + kChunk - 1) / kChunk;
    tile_mult[t] = 1;
}

// re-sort, scatter: previous tile t = blockIdx.x.  Stayers go in their order
// to tile_start[t] + (number of stayers before them); movers take the next
// position of their new tile's counter
template <typename real>
__global__ void __launch_bounds__(kSortBlock) k_scatter_keep(Species<real> src, Species<real> dst,
                                                             const int *__restrict__ old_start, TileGeom tg,
                                                             const int *__restrict__ tile_start, int *cur)
{
    __shared__ int wsum[kSortBlock / 64];
    __shared__ int bins[27], bin_base[27];
    const int t = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s = old_start[t], e = old_start[t + 1];
    int run = tile_start[t];   // next stayer position
    // every thread runs the same number of rounds (the barriers need it)
    for (int b0 = s; b0 < e; b0 += kSortBlock) {
        if (threadIdx.x < 27) bins[threadIdx.x] = 0;
        const int p = b0 + threadIdx.x;
        const int c = p < e ? src.cell[p] : -1;
        const int tn = c >= 0 ? tile_of_cell(tg, c) : -1;
        // the record's loads are in flight across the barriers below
        HeldRecord<real> rec;
        if (c >= 0) rec.load(src, p);
        const bool keep = tn == t;
        const unsigned long long m = __ballot(keep);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        // movers: rank inside their neighbour bin, then one atomic per bin
        const int b = c >= 0 && !keep ? neighbour_bin(tg, t, tn) : -1;
        const int rank = b >= 0 ? atomicAdd(&bins[b], 1) : 0;
        int wbase = 0, total = 0;
        for (int k = 0; k < kSortBlock / 64; ++k) {
            wbase += k < w ? wsum[k] : 0;
            total += wsum[k];
        }
        __syncthreads();
        if (threadIdx.x < 27 && bins[threadIdx.x] > 0)
            bin_base[threadIdx.x] = atomicAdd(cur + bin_tile(tg, t, threadIdx.x), bins[threadIdx.x]);
        __syncthreads();
        if (keep) rec.store(dst, c, run + wbase + before);
        else if (b >= 0) rec.store(dst, c, bin_base[b] + rank);
        else if (c >= 0) rec.store(dst, c, atomicAdd(cur + tn, 1));
        run += total;
        __syncthreads();   // wsum and the bins are rewritten next round
    }
}

// re-sort, appended particles: positions from the tile counters, one atomic
// per run of equal tiles
template <typename real>
__global__ void __launch_bounds__(256) k_scatter_tail(Species<real> src, Species<real> dst, int64_t first, int64_t n,
                                                      TileGeom tg, int *cur)
{
    const int64_t p = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = p < n ? src.cell[p] : -1;
    const int key = c >= 0 ? tile_of_cell(tg, c) : -1;
    HeldRecord<real> rec;   // in flight while the position is claimed
    if (c >= 0) rec.load(src, p);
    const int lane = threadIdx.x & 63;
    int head, len;
    key_runs(key, lane, head, len);
    int base = 0;
    if (lane == head && key >= 0) base = atomicAdd(cur + key, len);
    base = __shfl(base, head, 64);
    if (c < 0) return;
    rec.store(dst, c, base + (lane - head));
}

// The scatter reads its source in XCD-sized contiguous ranges (xcd_block):
// the particles of one tile, contiguous in the source, are then written by
// one XCD, whose L2 merges their scattered stores into whole lines instead
// of collecting partial lines in eight L2s.
template <typename real>
__global__ void __launch_bounds__(256) k_scatter(Species<real> src, Species<real> dst, int64_t first, int64_t n,
                                                 TileGeom tg, int *cursor, const int *__restrict__ tile_start,
                                                 const int *__restrict__ tile_mult)
{
    const int64_t p = first + xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int c = p < n ? src.cell[p] : -1;
    const int key = c >= 0 ? sort_key(tg, c) : -1;
    HeldRecord<real> rec;   // in flight while the position is claimed
    if (c >= 0) rec.load(src, p);
    const int lane = threadIdx.x & 63;
    int head, len;
    key_runs(key, lane, head, len);
    int base = 0;
    if (lane == head && key >= 0) base = atomicAdd(cursor + key, len);
    base = __shfl(base, head, 64);
    if (c < 0) return;
    // sorted position inside the tile -> its spread storage slot
    const int t = key / kTileCells;
    const int ts = tile_start[t], nt = tile_start[t + 1] - ts;
    rec.store(dst, c, ts + (int)((int64_t)(base + (lane - head) - ts) * tile_mult[t] % nt));
}

// Storage order inside a tile.  The n particles of a tile, sorted by cell,
// are stored so that slot s holds sorted position (s * a) mod n, with a the
// first of kSpread coprime to n.  A wavefront reading 64 consecutive slots
// then gets particles a positions apart in cell order, i.e. from distinct
// cells while a cell holds fewer than a particles: the push kernel's LDS
// atomics hit distinct addresses while its memory accesses stay contiguous.
// The scatter applies the inverse map, position e -> slot e * a^-1 mod n.
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
    const int s = offsets[t * kTileCells];
    tile_start[t] = s;
    if (t < ntiles) {
        const int cnt = offsets[(t + 1) * kTileCells] - s;
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
    const int nc = (e - s + kChunk - 1) / kChunk;
    if (nc == 0) return;
    // equal parts: a tile a little over kChunk gives two halves, not a full
    // chunk and a nearly empty one that pays the tile's staging again
    const int sz = (e - s + nc - 1) / nc;
    const int o = chunk_off[t];
    for (int c = 0; c < nc && o + c < max_chunks; ++c) {
        const int p = s + c * sz;
        chunks[o + c] = make_int4(t, p, max(0, min(sz, e - p)), 0);
    }
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
    const int64_t n_sorted = b.old_start ? std::min<int64_t>(b.n_sorted, n) : 0;
    hipError_t e;
    if (n_sorted > 0) {
        // re-sort: tiles of the previous sort keep their order
        const int64_t tail = n - n_sorted;
        if ((e = launch_fill_int(st, b.imm, (size_t)tg.ntiles + 1, 0)) != hipSuccess) return e;
        k_count_keep<<<tg.ntiles, kSortBlock, 0, st>>>(src.cell, b.old_start, tg, b.counts, b.imm);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (tail > 0) {
            k_count_tail<<<nblocks((size_t)tail, 256), 256, 0, st>>>(src.cell, n_sorted, n, tg, b.imm);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        k_sum_counts<<<nblocks((size_t)tg.ntiles, 256), 256, 0, st>>>(b.counts, b.imm, tg.ntiles);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = launch_exclusive_scan(st, b.imm, b.tile_start, tg.ntiles, b.scratch)) != hipSuccess) return e;
        k_keep_layout<<<nblocks((size_t)tg.ntiles + 1, 256), 256, 0, st>>>(b.tile_start, b.counts, tg.ntiles, b.imm,
                                                                           b.chunk_n, b.tile_mult);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        k_scatter_keep<real><<<tg.ntiles, kSortBlock, 0, st>>>(src, dst, b.old_start, tg, b.tile_start, b.imm);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (tail > 0) {
            k_scatter_tail<real><<<nblocks((size_t)tail, 256), 256, 0, st>>>(src, dst, n_sorted, n, tg, b.imm);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    } else {
        // first sort: (tile, cell) keys, spread storage order
        if ((e = launch_fill_int(st, b.counts, (size_t)tg.nkeys + 1, 0)) != hipSuccess) return e;
        if (n > 0) {
            k_count_keys<<<nblocks((size_t)n, 256), 256, 0, st>>>(src.cell, 0, n, tg, b.counts);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        if ((e = launch_exclusive_scan(st, b.counts, b.offsets, tg.nkeys, b.scratch)) != hipSuccess) return e;
        k_tile_chunks<<<nblocks((size_t)tg.ntiles + 1, 256), 256, 0, st>>>(b.offsets, tg.ntiles, b.tile_start,
                                                                           b.chunk_n, b.tile_mult);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (n > 0) {
            k_scatter<real><<<nblocks((size_t)n, 256), 256, 0, st>>>(src, dst, 0, n, tg, b.offsets, b.tile_start,
                                                                     b.tile_mult);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
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
