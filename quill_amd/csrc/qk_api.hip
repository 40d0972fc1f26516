// qk_api.hip -- the C ABI of include/quillk.h: context (the reference's
// spatial_region at n_sr = 1, include/main.h:17-170), device buffers, the
// step driver in the order of src/main.cpp:2319-2332, and copies in and out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "quillk.h"
#include "qk_kernels.h"

using namespace qk;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return fail(QK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define QK_TRY(expr)              \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != QK_OK) return rc_; \
    } while (0)

// Device scratch of qk_compute_rho_noalloc: the three FP64 node-charge
// arrays of a grid of up to kNoallocCells cells (48 MiB, held on each GPU the
// library runs on until the process ends).  In host code `scratch` is a
// host-side stand-in, not a device address: each device's copy is resolved
// once, by qk_create with that device current, into g_scratch_ptr[device],
// and only that pointer reaches kernels and copies.
constexpr int64_t kNoallocCells = int64_t(1) << 21;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[3 * kNoallocCells];

std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// this device's copy of `scratch`; the device must be current
int resolve_scratch(int dev)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return QK_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(QK_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)ptr;
    return QK_OK;
}

double *scratch_ptr(int dev)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[dev];
}

// the arrays of one particle buffer (structure of arrays)
struct SoA {
    void *mem = nullptr;   // one allocation holding every array
    int *cell = nullptr;
    void *a[8] = {};       // ox oy oz ux uy uz q chi (chi only with QED)
    int *id = nullptr;     // only with track_ids
};

struct SpeciesBuf {
    bool active = false;
    double cmr = 0;
    int kind = 0;           // 0 charged, 1 photon
    int energy_mode = 0;    // 0 (g-1) q, 1 g q, 2 (g-1) q / cmr
    int64_t cap = 0;
    int64_t n = 0;          // slots in use (live or free)
    int64_t n_sorted = 0;   // slots [0, n_sorted) are in tile order with a valid work list
    SoA buf[2];             // current and sort target
    int cur = 0;
    int4 *chunks = nullptr; // tiled-kernel work list
    int *d_nchunks = nullptr;
    int max_chunks = 0;     // capacity of chunks
    int nchunks = 0;        // host copy after the last sort
    int steps_since_sort = 0;
    double fb_last = -1;    // window-fallback fraction of the last finished step (-1: unknown or re-sorted since)
    int ishift = 0;         // moving-window shifts since the last sort
    int *tile_start = nullptr;   // ntiles + 1: the tiles of slots [0, n_sorted) (the next sort's source)
};

// The recorded trajectory of qk_tape_begin: snapshots of the whole device
// state (the 15 field arrays and each species' record arrays) at steps 0,
// every, 2 every, ...  The adjoint recomputes one segment at a time from its
// snapshot, keeping that segment's per-step states, and runs the step
// adjoints backwards over them.
struct Tape {
    bool active = false;
    bool replay = false;          // recomputing a segment: no snapshots
    bool dirty = false;           // the state was changed outside qk_step
    int every = 0;
    int steps = 0;                // steps since qk_tape_begin
    int64_t n[QK_MAX_SPECIES] = {};
    size_t field_bytes = 0, bytes = 0;
    std::vector<void *> ckpt;     // snapshots at steps 0, every, ...
    std::vector<int> ckpt_jcur;
    std::vector<uint32_t> ckpt_step;   // step index at each snapshot (laser sources depend on it)
    void *seg = nullptr;          // every per-step states of one segment
    void *fin = nullptr;          // the end state, restored after the adjoint
    double *lam = nullptr;        // 12 field adjoints: E, B, Be, J
    double *lpart = nullptr;      // 7 record adjoints per species
    // tangent-linear run along the recorded steps (qk_tangent_begin)
    bool tangent = false;
    double *tfield = nullptr;     // 15 FP64 arrays: E, B, Be, J (two buffers)
    int tjcur = 0;
    double *tpart = nullptr;      // 7 record tangents per species (x y z ux uy uz q)
};

// the laser sources of one side of the box (qk_add_laser_source) and the
// stack their corrections run on
struct InjSide {
    int nsrc = 0;
    int plane = 0;
    int W = 0;                    // planes per slab
    LaserFrame src[kMaxSources];
    uint32_t t0[kMaxSources] = {};   // step index at which each source was added
    int shifts0[kMaxSources] = {};   // window shifts at that time
    void *mem = nullptr;          // 6 field + 6 current arrays of the stack
    Grid gs{};
};

}  // namespace

struct qk_ctx {
    qk_params p;
    Grid g;
    TileGeom tg;
    SolverCoef coef;
    hipStream_t stream = nullptr;
    size_t rsize = 4;               // bytes per real
    void *fmem = nullptr;           // all field arrays
    void *F[15] = {};               // ex ey ez bx by bz bex bey bez j0x j0y j0z j1x j1y j1z
    int jcur = 0;                   // current buffer deposited this step
    bool have_step = false;
    SpeciesBuf sp[QK_MAX_SPECIES];
    // shared sort buffers (allocated at the first sort)
    int *s_counts = nullptr, *s_imm = nullptr, *s_offsets = nullptr, *s_tile_start = nullptr, *s_tile_mult = nullptr,
        *s_chunk_n = nullptr, *s_chunk_off = nullptr, *s_scratch = nullptr;
    int *h_sortinfo = nullptr;      // pinned: [live count, chunk count]
    double *d_diag = nullptr;       // [QK_MAX_SPECIES][kDiagPerSpecies] + field energy
    double *h_diag = nullptr;       // pinned per-step copies of the accumulators
    size_t h_diag_cap = 0;          // steps that fit in h_diag
    std::vector<hipEvent_t> ev;     // timing events
    double ms_total = 0, ms_part = 0, ms_field = 0, fallback_frac = 0;
    // QED: device slot counters of every species plus an overflow flag, and
    // their pinned host mirror; the global step index keys the random streams
    int *d_count = nullptr;
    int *h_count = nullptr;
    uint32_t step_index = 0;
    // merging: device accumulators, their pinned mirror, and the deletion
    // fractions of a merge due at the start of the next step
    double *d_merge = nullptr;
    double *h_merge = nullptr;
    bool merge_due = false;
    double merge_ppd[QK_MAX_SPECIES] = {};
    // moving window: get_max_w result (device, pinned mirror) and the staging
    // buffer of qk_fill_cells' cell lists
    unsigned long long *d_maxw = nullptr, *h_maxw = nullptr;
    void *d_fill = nullptr;
    size_t fill_cap = 0;
    Tape tape;
    // laser sources at an x plane, per side (0: x-min, 1: x-max)
    InjSide inj[2];
    int window_shifts = 0;          // qk_moving_window calls so far
};

namespace {

// per-species accumulators, then the field-energy slots
constexpr size_t kDiagLen = QK_MAX_SPECIES * kDiagPerSpecies + kEnergySlots + 1;
// the last slot counts particles dropped for a non-finite momentum
constexpr size_t kNonfiniteSlot = kDiagLen - 1;

template <typename real>
Fields<real> fields_of(qk_ctx *c)
{
    Fields<real> f;
    f.ex = (real *)c->F[0]; f.ey = (real *)c->F[1]; f.ez = (real *)c->F[2];
    f.bx = (real *)c->F[3]; f.by = (real *)c->F[4]; f.bz = (real *)c->F[5];
    f.bex = (real *)c->F[6]; f.bey = (real *)c->F[7]; f.bez = (real *)c->F[8];
    return f;
}

template <typename real>
Current<real> current_of(qk_ctx *c, int which)
{
    Current<real> j;
    j.jx = (real *)c->F[9 + 3 * which];
    j.jy = (real *)c->F[10 + 3 * which];
    j.jz = (real *)c->F[11 + 3 * which];
    return j;
}

template <typename real>
Species<real> view(const SoA &b)
{
    Species<real> s;
    s.cell = b.cell;
    s.ox = (real *)b.a[0]; s.oy = (real *)b.a[1]; s.oz = (real *)b.a[2];
    s.ux = (real *)b.a[3]; s.uy = (real *)b.a[4]; s.uz = (real *)b.a[5];
    s.q = (real *)b.a[6];
    s.chi = (real *)b.a[7];
    s.id = b.id;
    return s;
}

double species_cmr(const qk_params &p, int s)
{
    if (s == QK_ELECTRON) return -1;
    if (s == QK_POSITRON) return 1;
    if (s == QK_PHOTON) return 0;
    return p.ion_cmr[s - QK_ION0];
}

bool species_valid(const qk_ctx *c, int s)
{
    if (s < 0 || s >= QK_MAX_SPECIES) return false;
    if (s >= QK_ION0 && s - QK_ION0 >= c->p.n_ion) return false;
    return true;
}

int alloc_soa(qk_ctx *c, SoA &b, int64_t cap)
{
    const size_t r = c->rsize;
    const bool chi = c->p.qed != 0, ids = c->p.track_ids != 0;
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t a_int = align((size_t)cap * 4), a_real = align((size_t)cap * r);
    const size_t bytes = a_int + 7 * a_real + (chi ? a_real : 0) + (ids ? a_int : 0);
    void *mem = nullptr;
    HIP_TRY(hipMalloc(&mem, bytes));
    char *m = (char *)mem;
    b.mem = mem;
    b.cell = (int *)m;
    m += a_int;
    for (int a = 0; a < 7; ++a, m += a_real) b.a[a] = m;
    b.a[7] = chi ? (void *)m : nullptr;
    m += chi ? a_real : 0;
    b.id = ids ? (int *)m : nullptr;
    return QK_OK;
}

// (re)allocate both buffers of one species to cap slots, keeping slots [0, n)
int reserve(qk_ctx *c, int s, int64_t cap)
{
    SpeciesBuf &b = c->sp[s];
    if (cap <= b.cap) return QK_OK;
    SoA nb[2];
    for (int k = 0; k < 2; ++k) {
        int rc = alloc_soa(c, nb[k], cap);
        if (rc) {
            for (int m = 0; m < k; ++m) (void)hipFree(nb[m].mem);
            return rc;
        }
    }
    const size_t r = c->rsize;
    SoA &old = b.buf[b.cur];
    if (old.mem && b.n > 0) {
        HIP_TRY(hipMemcpyAsync(nb[0].cell, old.cell, (size_t)b.n * 4, hipMemcpyDeviceToDevice, c->stream));
        for (int a = 0; a < 8; ++a)
            if (old.a[a] && nb[0].a[a])
                HIP_TRY(hipMemcpyAsync(nb[0].a[a], old.a[a], (size_t)b.n * r, hipMemcpyDeviceToDevice, c->stream));
        if (old.id && nb[0].id) HIP_TRY(hipMemcpyAsync(nb[0].id, old.id, (size_t)b.n * 4, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    for (int k = 0; k < 2; ++k)
        if (b.buf[k].mem) HIP_TRY(hipFree(b.buf[k].mem));
    b.buf[0] = nb[0];
    b.buf[1] = nb[1];
    b.cur = 0;
    b.cap = cap;
    // work list capacity: every tile may end in a partial chunk
    const int64_t mc = (int64_t)c->tg.ntiles + cap / kChunk + 1;
    if (b.chunks) HIP_TRY(hipFree(b.chunks));
    b.chunks = nullptr;
    HIP_TRY(hipMalloc(&b.chunks, (size_t)mc * sizeof(int4)));
    b.max_chunks = (int)mc;
    if (!b.d_nchunks) HIP_TRY(hipMalloc(&b.d_nchunks, sizeof(int)));
    b.n_sorted = 0;   // the work list refers to the old buffers
    b.nchunks = 0;
    return QK_OK;
}

int ensure_species(qk_ctx *c, int s, int64_t extra)
{
    SpeciesBuf &b = c->sp[s];
    if (!b.active) {
        b.active = true;
        b.cmr = species_cmr(c->p, s);
        b.kind = (s == QK_PHOTON) ? 1 : 0;
        b.energy_mode = (s == QK_PHOTON) ? 1 : (s >= QK_ION0 ? 2 : 0);
    }
    const int64_t need = b.n + extra;
    if (need <= b.cap) return QK_OK;
    // the push kernels address records with 32-bit byte offsets
    const int64_t limit = ((int64_t)1 << 32) / (int64_t)c->rsize;
    if (need >= limit)
        return fail(QK_ERR_CAPACITY, "more than " + std::to_string(limit - 1) + " particles in one species");
    int64_t cap = c->p.capacity > 0 ? c->p.capacity : std::min(std::max<int64_t>(need + need / 4, 1024), limit - 1);
    if (cap >= limit) return fail(QK_ERR_CAPACITY, "capacity " + std::to_string(cap) + " exceeds the 32-bit record offsets");
    if (cap < need) {
        if (c->p.capacity > 0)
            return fail(QK_ERR_CAPACITY, "particle capacity " + std::to_string(c->p.capacity) + " exceeded (need " +
                                             std::to_string(need) + ")");
        cap = need;
    }
    return reserve(c, s, cap);
}

int ensure_sort_buffers(qk_ctx *c)
{
    if (c->s_counts) return QK_OK;
    const size_t nk = (size_t)c->tg.nkeys + 1, nt = (size_t)c->tg.ntiles + 1;
    HIP_TRY(hipMalloc(&c->s_counts, nk * sizeof(int)));
    HIP_TRY(hipMalloc(&c->s_imm, nk * sizeof(int)));
    HIP_TRY(hipMalloc(&c->s_offsets, nk * sizeof(int)));
    HIP_TRY(hipMalloc(&c->s_tile_start, nt * sizeof(int)));
    HIP_TRY(hipMalloc(&c->s_tile_mult, nt * sizeof(int)));
    HIP_TRY(hipMalloc(&c->s_chunk_n, nt * sizeof(int)));
    HIP_TRY(hipMalloc(&c->s_chunk_off, nt * sizeof(int)));
    HIP_TRY(hipMalloc(&c->s_scratch, scan_scratch_ints(std::max(nk, nt)) * sizeof(int)));
    HIP_TRY(hipHostMalloc((void **)&c->h_sortinfo, 2 * sizeof(int)));
    return QK_OK;
}

// sort species s by (tile, cell), compacting free slots; builds its work list
template <typename real>
int sort_species(qk_ctx *c, int s)
{
    QK_TRY(ensure_sort_buffers(c));
    SpeciesBuf &b = c->sp[s];
    if (!b.tile_start) HIP_TRY(hipMalloc(&b.tile_start, ((size_t)c->tg.ntiles + 1) * sizeof(int)));
    SortBuffers sb;
    sb.counts = c->s_counts;
    sb.imm = c->s_imm;
    sb.old_start = b.n_sorted > 0 ? b.tile_start : nullptr;
    sb.n_sorted = b.n_sorted;
    sb.offsets = c->s_offsets;
    sb.tile_start = c->s_tile_start;
    sb.tile_mult = c->s_tile_mult;
    sb.chunk_n = c->s_chunk_n;
    sb.chunk_off = c->s_chunk_off;
    sb.chunks = b.chunks;
    sb.scratch = c->s_scratch;
    sb.max_chunks = b.max_chunks;
    HIP_TRY(launch_sort<real>(c->stream, c->tg, view<real>(b.buf[b.cur]), view<real>(b.buf[b.cur ^ 1]), b.n, sb));
    HIP_TRY(hipMemcpyAsync(b.d_nchunks, c->s_chunk_off + c->tg.ntiles, sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(b.tile_start, c->s_tile_start, ((size_t)c->tg.ntiles + 1) * sizeof(int),
                           hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_sortinfo, c->s_tile_start + c->tg.ntiles, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_sortinfo + 1, c->s_chunk_off + c->tg.ntiles, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    b.cur ^= 1;
    b.n = c->h_sortinfo[0];
    b.n_sorted = b.n;
    b.nchunks = c->h_sortinfo[1];
    b.steps_since_sort = 0;
    b.fb_last = -1;
    b.ishift = 0;
    if (b.nchunks > b.max_chunks) return fail(QK_ERR_STATE, "tile work list overflow");
    return QK_OK;
}

void unpack_diag(const qk_ctx *c, const double *h, qk_diag &d)
{
    std::memset(&d, 0, sizeof d);
    const double nn = c->p.norm_n, ef = c->p.norm_ef, ep = c->p.norm_ep;
    const double *e = h + QK_ELECTRON * kDiagPerSpecies, *po = h + QK_POSITRON * kDiagPerSpecies,
                 *ph = h + QK_PHOTON * kDiagPerSpecies;
    d.N_e = -e[0] * nn;   // electrons count with the opposite sign (fout.cpp:62)
    d.n_e = (int64_t)e[2];
    d.energy_e = -e[1] * ep;
    d.N_p = po[0] * nn;
    d.n_p = (int64_t)po[2];
    d.energy_p = po[1] * ep;
    d.N_ph = ph[0] * nn;
    d.n_ph = (int64_t)ph[2];
    d.energy_ph = ph[1] * ep;
    for (int s = QK_ION0; s < QK_MAX_SPECIES; ++s) {
        d.n_ion += (int64_t)h[s * kDiagPerSpecies + 2];
        d.energy_ions[s - QK_ION0] = h[s * kDiagPerSpecies + 1] * ep;
        d.energy_ion += d.energy_ions[s - QK_ION0];
    }
    double ef_sum = 0;
    for (int m = 0; m < kEnergySlots; ++m) ef_sum += h[QK_MAX_SPECIES * kDiagPerSpecies + m];
    d.energy_f = ef_sum * ef;
    d.n_nonfinite = (int64_t)h[kNonfiniteSlot];
}

int ensure_events(qk_ctx *c, size_t n)
{
    while (c->ev.size() < n) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        c->ev.push_back(e);
    }
    return QK_OK;
}

// pmerging (src/mc.cpp:202-462) on every species; see qk_pmerging
template <typename real>
int merge_impl(qk_ctx *c, int mode, const double *ppd)
{
    if (!c->d_merge) {
        HIP_TRY(hipMalloc(&c->d_merge, 4 * sizeof(double)));
        HIP_TRY(hipHostMalloc((void **)&c->h_merge, 4 * sizeof(double)));
    }
    auto reduce = [&](auto &&launch, double out[4]) -> int {
        HIP_TRY(launch_zero(c->stream, c->d_merge, 4));
        HIP_TRY(launch());
        HIP_TRY(hipMemcpyAsync(c->h_merge, c->d_merge, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int a = 0; a < 4; ++a) out[a] = c->h_merge[a];
        return QK_OK;
    };
    for (int s = 0; s < QK_MAX_SPECIES; ++s) {
        SpeciesBuf &b = c->sp[s];
        if (!b.active || b.n == 0) continue;
        const double pd = mode == 1 ? ppd[0] : ppd[s];
        if (!(pd > 0)) continue;
        const Species<real> v = view<real>(b.buf[b.cur]);
        double tot[4], got[4];
        QK_TRY(reduce([&] { return launch_charge_sum<real>(c->stream, v, b.n, c->d_merge); }, tot));
        if (tot[2] == 0) continue;
        const double av_q = tot[0] / tot[2];
        QK_TRY(reduce([&] {
            return launch_merge_delete<real>(c->stream, v, b.n, mode, pd, av_q, c->p.seed, c->step_index,
                                             0x400u + (uint32_t)s, c->d_merge);
        }, got));
        // "ti": Q / (Q - deleted); "nl": Q / (sum of the raised survivors)
        double a;
        if (mode == 1) a = tot[0] - got[0] != 0 ? tot[0] / (tot[0] - got[0]) : 1;
        else a = got[0] != 0 ? tot[0] / got[0] : 0;
        HIP_TRY(launch_scale_charge<real>(c->stream, v, b.n, a));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QK_OK;
}

// evaluate_merging_condition (src/main.cpp:1691-1748) on the counts of the
// step just finished (h: that step's accumulators)
void merge_condition(qk_ctx *c, const double *h)
{
    const double crnp = c->p.crnp;
    const int nsp = QK_ION0 + c->p.n_ion;
    double cnt[QK_MAX_SPECIES] = {}, total = 0;
    for (int s = 0; s < nsp; ++s) {
        cnt[s] = h[s * kDiagPerSpecies + 2];
        total += cnt[s];
    }
    c->merge_due = false;
    for (int s = 0; s < QK_MAX_SPECIES; ++s) c->merge_ppd[s] = 0;
    if (c->p.merging == 1) {
        if (total > nsp * crnp) {
            c->merge_due = true;
            c->merge_ppd[0] = (total - nsp * crnp) / total;
        }
    } else if (c->p.merging == 2) {
        for (int s = 0; s < nsp; ++s)
            if (cnt[s] > crnp) {
                c->merge_due = true;
                c->merge_ppd[s] = (cnt[s] - crnp) / cnt[s];
            }
    }
}

// QED slot counters: host -> device before kernels that append, device ->
// host (with a stream sync) after them
int qed_upload_counts(qk_ctx *c)
{
    for (int s = 0; s < QK_MAX_SPECIES; ++s) c->h_count[s] = (int)c->sp[s].n;
    c->h_count[QK_MAX_SPECIES] = 0;
    HIP_TRY(hipMemcpyAsync(c->d_count, c->h_count, (QK_MAX_SPECIES + 1) * sizeof(int), hipMemcpyHostToDevice,
                           c->stream));
    return QK_OK;
}

int qed_download_counts(qk_ctx *c)
{
    HIP_TRY(hipMemcpyAsync(c->h_count, c->d_count, (QK_MAX_SPECIES + 1) * sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->h_count[QK_MAX_SPECIES]) return fail(QK_ERR_CAPACITY, "QED appends exceeded the reserved capacity");
    for (int s = 0; s < QK_MAX_SPECIES; ++s)
        if (c->sp[s].active) c->sp[s].n = c->h_count[s];
    return QK_OK;
}

// The particle phase with QED on (src/main.cpp:2323-2324): birth_from_vacuum,
// then padvance with photon emission for leptons and pair creation for
// photons.  Capacity is reserved for the worst case (every lepton emits,
// every photon converts) so that no append can fail; the new counts are
// read back after the step.
template <typename real>
int qed_particle_phase(qk_ctx *c, const Fields<real> &f, const Current<real> &jc)
{
    SpeciesBuf &be = c->sp[QK_ELECTRON], &bp = c->sp[QK_POSITRON], &bph = c->sp[QK_PHOTON];
    const QedArgs qa{c->p.e_s, c->p.seed, c->step_index};
    auto target = [&](int s) {
        SpeciesBuf &b = c->sp[s];
        AppendTarget<real> t;
        t.s = view<real>(b.buf[b.cur]);
        t.count = c->d_count + s;
        t.cap = (int)b.cap;
        t.overflow = c->d_count + QK_MAX_SPECIES;
        t.diag = c->d_diag + s * kDiagPerSpecies;
        return t;
    };
    const int64_t ninner = std::max<int64_t>(0, (int64_t)(c->p.nx - 4) * (c->p.ny - 4) * (c->p.nz - 4));
    if (ninner > 0) {
        QK_TRY(ensure_species(c, QK_ELECTRON, ninner));
        QK_TRY(ensure_species(c, QK_POSITRON, ninner));
        QK_TRY(qed_upload_counts(c));
        HIP_TRY(launch_vacuum_pairs<real>(c->stream, c->g, f, c->p.pusher, c->p.q_vac, qa, target(QK_ELECTRON),
                                          target(QK_POSITRON)));
        QK_TRY(qed_download_counts(c));
    }
    QK_TRY(ensure_species(c, QK_PHOTON, be.n + bp.n));
    QK_TRY(ensure_species(c, QK_ELECTRON, bph.n));
    QK_TRY(ensure_species(c, QK_POSITRON, bph.n));
    QK_TRY(qed_upload_counts(c));
    const int64_t n_ph0 = bph.n;   // photons emitted in this step are not visited
    for (int s = 0; s < QK_MAX_SPECIES; ++s) {
        SpeciesBuf &b = c->sp[s];
        if (!b.active || b.n == 0 || s == QK_PHOTON) continue;
        PushArgs a{};
        a.pusher = c->p.pusher;
        a.kind = b.kind;
        a.cmr = b.cmr;
        a.external_bz = c->p.external_bz;
        a.diag = c->d_diag + s * kDiagPerSpecies;
        a.energy_mode = b.energy_mode;
        a.nonfinite = c->d_diag + kNonfiniteSlot;
        a.first = 0;
        a.n = b.n;
        const Species<real> v = view<real>(b.buf[b.cur]);
        if (s == QK_ELECTRON || s == QK_POSITRON)
            HIP_TRY(launch_push_qed_lepton<real>(c->stream, c->g, f, jc, v, a, qa, target(QK_PHOTON)));
        else
            HIP_TRY(launch_push_global<real>(c->stream, c->g, f, jc, v, a));
        ++b.steps_since_sort;
    }
    if (n_ph0 > 0) {
        PushArgs a{};
        a.pusher = c->p.pusher;
        a.kind = 1;
        a.cmr = 0;
        a.external_bz = c->p.external_bz;
        a.diag = c->d_diag + QK_PHOTON * kDiagPerSpecies;
        a.energy_mode = 1;
        a.nonfinite = c->d_diag + kNonfiniteSlot;
        a.first = 0;
        a.n = n_ph0;
        HIP_TRY(launch_photon_qed<real>(c->stream, c->g, f, jc, view<real>(bph.buf[bph.cur]), a, qa,
                                        target(QK_ELECTRON), target(QK_POSITRON)));
        ++bph.steps_since_sort;
    }
    return QK_OK;
}

// ---- adjoint tape -----------------------------------------------------------

// the record arrays a snapshot keeps per species: cell, ox oy oz ux uy uz q
size_t species_state_bytes(const qk_ctx *c, int64_t n) { return (size_t)n * (4 + 7 * c->rsize); }

int tape_save(qk_ctx *c, void *dst)
{
    char *d = (char *)dst;
    HIP_TRY(hipMemcpyAsync(d, c->fmem, c->tape.field_bytes, hipMemcpyDeviceToDevice, c->stream));
    d += c->tape.field_bytes;
    for (int s = 0; s < QK_MAX_SPECIES; ++s) {
        const int64_t n = c->tape.n[s];
        if (n == 0) continue;
        const SoA &b = c->sp[s].buf[c->sp[s].cur];
        HIP_TRY(hipMemcpyAsync(d, b.cell, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
        d += (size_t)n * 4;
        for (int a = 0; a < 7; ++a) {
            HIP_TRY(hipMemcpyAsync(d, b.a[a], (size_t)n * c->rsize, hipMemcpyDeviceToDevice, c->stream));
            d += (size_t)n * c->rsize;
        }
    }
    return QK_OK;
}

int tape_load(qk_ctx *c, const void *src)
{
    const char *d = (const char *)src;
    HIP_TRY(hipMemcpyAsync(c->fmem, d, c->tape.field_bytes, hipMemcpyDeviceToDevice, c->stream));
    d += c->tape.field_bytes;
    for (int s = 0; s < QK_MAX_SPECIES; ++s) {
        const int64_t n = c->tape.n[s];
        if (n == 0) continue;
        const SoA &b = c->sp[s].buf[c->sp[s].cur];
        HIP_TRY(hipMemcpyAsync(b.cell, d, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
        d += (size_t)n * 4;
        for (int a = 0; a < 7; ++a) {
            HIP_TRY(hipMemcpyAsync(b.a[a], d, (size_t)n * c->rsize, hipMemcpyDeviceToDevice, c->stream));
            d += (size_t)n * c->rsize;
        }
    }
    return QK_OK;
}

// views of a saved state: the fields (E, Be) and one species' records
template <typename real>
Fields<real> saved_fields(const qk_ctx *c, const void *state)
{
    const size_t fb = c->tape.field_bytes / 15;
    const char *m = (const char *)state;
    Fields<real> f;
    real *F[9];
    for (int a = 0; a < 9; ++a) F[a] = (real *)(m + a * fb);
    f.ex = F[0]; f.ey = F[1]; f.ez = F[2];
    f.bx = F[3]; f.by = F[4]; f.bz = F[5];
    f.bex = F[6]; f.bey = F[7]; f.bez = F[8];
    return f;
}

template <typename real>
Species<real> saved_species(const qk_ctx *c, const void *state, int s)
{
    const char *d = (const char *)state + c->tape.field_bytes;
    for (int t = 0; t < s; ++t) d += species_state_bytes(c, c->tape.n[t]);
    const int64_t n = c->tape.n[s];
    Species<real> v;
    v.cell = (int *)d;
    d += (size_t)n * 4;
    real *a[7];
    for (int k = 0; k < 7; ++k) {
        a[k] = (real *)d;
        d += (size_t)n * c->rsize;
    }
    v.ox = a[0]; v.oy = a[1]; v.oz = a[2];
    v.ux = a[3]; v.uy = a[4]; v.uz = a[5];
    v.q = a[6];
    v.chi = nullptr;
    v.id = nullptr;
    return v;
}

int tape_snapshot(qk_ctx *c)
{
    void *m = nullptr;
    HIP_TRY(hipMalloc(&m, c->tape.bytes));
    c->tape.ckpt.push_back(m);
    c->tape.ckpt_jcur.push_back(c->jcur);
    c->tape.ckpt_step.push_back(c->step_index);
    return tape_save(c, m);
}

void tape_free(qk_ctx *c)
{
    Tape &t = c->tape;
    for (void *m : t.ckpt) (void)hipFree(m);
    t.ckpt.clear();
    t.ckpt_jcur.clear();
    t.ckpt_step.clear();
    if (t.seg) (void)hipFree(t.seg);
    if (t.fin) (void)hipFree(t.fin);
    if (t.lam) (void)hipFree(t.lam);
    if (t.lpart) (void)hipFree(t.lpart);
    if (t.tfield) (void)hipFree(t.tfield);
    if (t.tpart) (void)hipFree(t.tpart);
    t.seg = t.fin = nullptr;
    t.lam = t.lpart = nullptr;
    t.tfield = t.tpart = nullptr;
    t.tangent = false;
    t.active = t.replay = t.dirty = false;
    t.steps = 0;
}

template <typename real>
int step_impl(qk_ctx *c, int nsteps, qk_diag *diag);

// the adjoint sweep of qk_adjoint (see there)
template <typename real>
int adjoint_impl(qk_ctx *c, const double *lam_E, const double *lam_B, const double *const *lam_part, double *grad_E,
                 double *grad_B, double *const *grad_part)
{
    Tape &t = c->tape;
    const int64_t nc = c->g.ncell;
    const int T = t.steps;
    // adjoint buffers
    int64_t np_total = 0;
    for (int s = 0; s < QK_MAX_SPECIES; ++s) np_total += t.n[s];
    if (!t.lam) HIP_TRY(hipMalloc(&t.lam, (size_t)nc * 12 * sizeof(double)));
    if (!t.lpart && np_total > 0) HIP_TRY(hipMalloc(&t.lpart, (size_t)np_total * 7 * sizeof(double)));
    if (!t.fin) HIP_TRY(hipMalloc(&t.fin, t.bytes));
    if (!t.seg) HIP_TRY(hipMalloc(&t.seg, t.bytes * (size_t)t.every));
    double *lam[12];
    for (int a = 0; a < 12; ++a) lam[a] = t.lam + (size_t)a * nc;
    double *lpart[QK_MAX_SPECIES][7] = {};
    {
        double *m = t.lpart;
        for (int s = 0; s < QK_MAX_SPECIES; ++s)
            for (int a = 0; a < 7; ++a) {
                lpart[s][a] = m;
                m += t.n[s];
            }
    }
    // the end state, and its bookkeeping, to come back to
    const int jcur_end = c->jcur;
    const uint32_t step_end = c->step_index;
    QK_TRY(tape_save(c, t.fin));
    // seeds: fields AoS -> SoA, records AoS (x y z ux uy uz q) -> SoA
    {
        std::vector<double> h((size_t)nc * 12, 0.0);
        for (int64_t m = 0; m < nc; ++m)
            for (int a = 0; a < 3; ++a) {
                if (lam_E) h[(size_t)a * nc + m] = lam_E[3 * m + a];
                if (lam_B) h[(size_t)(3 + a) * nc + m] = lam_B[3 * m + a];
            }
        HIP_TRY(hipMemcpyAsync(t.lam, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        for (int s = 0; s < QK_MAX_SPECIES; ++s) {
            const int64_t n = t.n[s];
            if (n == 0) continue;
            std::vector<double> hp((size_t)n * 7, 0.0);
            if (lam_part && lam_part[s])
                for (int64_t p = 0; p < n; ++p)
                    for (int a = 0; a < 7; ++a) hp[(size_t)a * n + p] = lam_part[s][7 * p + a];
            HIP_TRY(hipMemcpyAsync(lpart[s][0], hp.data(), hp.size() * sizeof(double), hipMemcpyHostToDevice,
                                   c->stream));
        }
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    // segments, last first
    t.replay = true;
    const int nseg = (T + t.every - 1) / t.every;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int first = seg * t.every, last = std::min(T, first + t.every);
        QK_TRY(tape_load(c, t.ckpt[seg]));
        c->jcur = t.ckpt_jcur[seg];
        c->step_index = t.ckpt_step[seg];
        for (int n = first; n < last; ++n) {
            QK_TRY(tape_save(c, (char *)t.seg + (size_t)(n - first) * t.bytes));
            if (n + 1 < last) QK_TRY(step_impl<real>(c, 1, nullptr));
        }
        for (int n = last - 1; n >= first; --n) {
            const void *st = (const char *)t.seg + (size_t)(n - first) * t.bytes;
            HIP_TRY(launch_adj_fields(c->stream, c->p.solver, c->g, c->coef, lam));
            const Fields<real> f = saved_fields<real>(c, st);
            for (int s = 0; s < QK_MAX_SPECIES; ++s) {
                if (t.n[s] == 0) continue;
                PushArgs a{};
                a.pusher = c->p.pusher;
                a.kind = 0;
                a.cmr = c->sp[s].cmr;
                a.external_bz = c->p.external_bz;
                HIP_TRY(launch_adj_particles<real>(c->stream, c->g, f, saved_species<real>(c, st, s), t.n[s], a, lam,
                                                   lpart[s]));
            }
        }
    }
    t.replay = false;
    // Be of step 0 came from interpolate_be, and init_boundaries zeroed the
    // boundary planes of the fields set before it: their adjoints are zero
    HIP_TRY(launch_adj_interp_be(c->stream, c->g, lam));
    {
        Fields<double> lf;
        lf.ex = lam[0]; lf.ey = lam[1]; lf.ez = lam[2];
        lf.bx = lam[3]; lf.by = lam[4]; lf.bz = lam[5];
        lf.bex = lam[6]; lf.bey = lam[7]; lf.bez = lam[8];
        HIP_TRY(launch_init_boundaries<double>(c->stream, c->p.solver, c->g, lf));
    }
    // results
    {
        std::vector<double> h((size_t)nc * 6);
        HIP_TRY(hipMemcpyAsync(h.data(), t.lam, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        std::vector<std::vector<double>> hp(QK_MAX_SPECIES);
        for (int s = 0; s < QK_MAX_SPECIES; ++s) {
            if (t.n[s] == 0 || !grad_part || !grad_part[s]) continue;
            hp[s].resize((size_t)t.n[s] * 7);
            HIP_TRY(hipMemcpyAsync(hp[s].data(), lpart[s][0], hp[s].size() * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
        }
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int64_t m = 0; m < nc; ++m)
            for (int a = 0; a < 3; ++a) {
                if (grad_E) grad_E[3 * m + a] = h[(size_t)a * nc + m];
                if (grad_B) grad_B[3 * m + a] = h[(size_t)(3 + a) * nc + m];
            }
        for (int s = 0; s < QK_MAX_SPECIES; ++s) {
            if (hp[s].empty()) continue;
            const int64_t n = t.n[s];
            for (int64_t p = 0; p < n; ++p)
                for (int a = 0; a < 7; ++a) grad_part[s][7 * p + a] = hp[s][(size_t)a * n + p];
        }
    }
    QK_TRY(tape_load(c, t.fin));
    c->jcur = jcur_end;
    c->step_index = step_end;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QK_OK;
}

// the tangent state: FP64 fields (E, B, Be) and the current buffer `which`
Fields<double> tangent_fields(const qk_ctx *c)
{
    const size_t n = (size_t)c->g.ncell;
    double *m = c->tape.tfield;
    Fields<double> f;
    f.ex = m; f.ey = m + n; f.ez = m + 2 * n;
    f.bx = m + 3 * n; f.by = m + 4 * n; f.bz = m + 5 * n;
    f.bex = m + 6 * n; f.bey = m + 7 * n; f.bez = m + 8 * n;
    return f;
}

Current<double> tangent_current(const qk_ctx *c, int which)
{
    const size_t n = (size_t)c->g.ncell;
    double *m = c->tape.tfield + (9 + 3 * which) * n;
    return Current<double>{m, m + n, m + 2 * n};
}

// the 7 record tangents of species s (slots of the tape)
void tangent_records(const qk_ctx *c, int s, double *out[7])
{
    double *m = c->tape.tpart;
    for (int t = 0; t < s; ++t) m += 7 * c->tape.n[t];
    for (int a = 0; a < 7; ++a) out[a] = m + (size_t)a * c->tape.n[s];
}

// views of a side's stack: fields (ex..bz; no Be) and two current sets
template <typename real>
void stack_views(const InjSide &side, Fields<real> &f, Current<real> &j0, Current<real> &j1)
{
    const size_t n = ((size_t)side.gs.ncell * sizeof(real) + 255) & ~(size_t)255;
    char *m = (char *)side.mem;
    real *a[12];
    for (int k = 0; k < 12; ++k) a[k] = (real *)(m + k * n);
    f.ex = a[0]; f.ey = a[1]; f.ez = a[2];
    f.bx = a[3]; f.by = a[4]; f.bz = a[5];
    f.bex = f.bey = f.bez = nullptr;
    j0.jx = a[6]; j0.jy = a[7]; j0.jz = a[8];
    j1.jx = a[9]; j1.jy = a[10]; j1.jz = a[11];
}

// Whether any pulse of the side can be nonzero on its stack planes after
// travelling shift[n]: the envelope is |x| < xs along the propagation axis,
// and x is linear in position, so the slab's corner points bound it.
bool side_active(const qk_ctx *c, const InjSide &side, const InjectArgs &a)
{
    const int lo = side.plane - 6 > 0 ? side.plane - 6 : 0;
    const int hi = side.plane + 6 < c->g.nx ? side.plane + 6 : c->g.nx;
    for (int n = 0; n < side.nsrc; ++n) {
        const LaserFrame &L = side.src[n];
        double xmin = 1e300, xmax = -1e300;
        for (int corner = 0; corner < 8; ++corner) {
            const double px = ((corner & 1) ? hi : lo) * c->g.dx + a.ishift[n] * c->g.dx;
            const double py = ((corner & 2) ? c->g.ny : 0) * c->g.dy, pz = ((corner & 4) ? c->g.nz : 0) * c->g.dz;
            double x = L.r0[0] * (px - L.x0 - L.xt) + L.r0[1] * (py - L.y12 - L.y0 - L.yt) +
                       L.r0[2] * (pz - L.z12 - L.z0 - L.zt);
            x -= L.b_sign ? a.shift[n] : -a.shift[n];
            xmin = std::min(xmin, x);
            xmax = std::max(xmax, x);
        }
        if (xmax > -L.xs && xmin < L.xs) return true;
    }
    return false;
}

// the laser sources' corrections of one sub-step of fadvance: b_field = 1
// after a B half step at time (step + frac) dt, 0 after the E step
template <typename real>
int inject_impl(qk_ctx *c, int b_field, double frac)
{
    for (int sd = 0; sd < 2; ++sd) {
        InjSide &side = c->inj[sd];
        if (side.nsrc == 0) continue;
        InjectArgs a{};
        a.nsrc = side.nsrc;
        for (int n = 0; n < side.nsrc; ++n) {
            a.src[n] = side.src[n];
            a.shift[n] = ((double)(c->step_index - side.t0[n]) + frac) * c->g.dt;
            a.ishift[n] = c->window_shifts - side.shifts0[n];
        }
        a.side = sd;
        a.plane = side.plane;
        a.W = side.W;
        a.main_lo = sd == 0 ? 0 : c->g.nx - side.W;
        a.b_field = b_field;
        if (!side_active(c, side, a)) continue;
        Fields<real> st;
        Current<real> j0, j1;
        stack_views<real>(side, st, j0, j1);
        HIP_TRY(launch_inject<real>(c->stream, c->p.solver, c->g, c->coef, fields_of<real>(c), side.gs, st, j0, j1, a));
    }
    return QK_OK;
}

template <typename real>
int step_impl(qk_ctx *c, int nsteps, qk_diag *diag)
{
    const Fields<real> f = fields_of<real>(c);
    // three events per step: before the particle phase, between phases, after the field phase
    QK_TRY(ensure_events(c, 3 * (size_t)nsteps));
    if (c->h_diag_cap < (size_t)nsteps) {
        if (c->h_diag) HIP_TRY(hipHostFree(c->h_diag));
        c->h_diag = nullptr;
        HIP_TRY(hipHostMalloc((void **)&c->h_diag, (size_t)nsteps * kDiagLen * sizeof(double)));
        c->h_diag_cap = (size_t)nsteps;
    }
    double *hd = c->h_diag;
    const bool qed = c->p.qed != 0;
    // the QED kernels deposit through global atomics
    const bool tiled = c->p.deposit == QK_DEPOSIT_TILED && !qed;
    for (int l = 0; l < nsteps; ++l) {
        Tape &tp = c->tape;
        if (tp.active && !tp.replay) {
            if (tp.steps > 0 && tp.steps % tp.every == 0) QK_TRY(tape_snapshot(c));
            ++tp.steps;
        }
        // The window fallback of step l - 2: once step l - 1 has begun its
        // particle phase, step l - 2 and its diagnostics copy are done, and
        // the GPU still has step l - 1 queued while the host decides.  The
        // first two steps of a call use the last step of the previous call
        // (kept per species), so the rule holds for callers that step one
        // step per call.
        bool fb_high[QK_MAX_SPECIES] = {};
        if (tiled && c->p.sort_fallback > 0 && l >= 2) {
            HIP_TRY(hipEventSynchronize(c->ev[3 * (l - 1)]));
            const double *h = hd + (size_t)(l - 2) * kDiagLen;
            for (int s = 0; s < QK_MAX_SPECIES; ++s) {
                const double live = h[s * kDiagPerSpecies + 2], fb = h[s * kDiagPerSpecies + 3];
                fb_high[s] = live > 0 && fb > c->p.sort_fallback * live;
            }
        } else if (tiled && c->p.sort_fallback > 0) {
            for (int s = 0; s < QK_MAX_SPECIES; ++s) fb_high[s] = c->sp[s].fb_last > c->p.sort_fallback;
        }
        // cell-tile sort: the cadence, a fallback above sort_fallback (seen
        // on a layout at least two steps old), or an unsorted tail (appended
        // particles, pushed by the global kernel meanwhile) above 1/16 of the
        // sorted part (photons only with QED, where pair creation leaves free
        // slots)
        for (int s = 0; s < QK_MAX_SPECIES; ++s) {
            SpeciesBuf &b = c->sp[s];
            if (!b.active || b.n == 0 || (b.kind == 1 && !qed)) continue;
            const bool due = (c->p.sort_every > 0 && b.steps_since_sort >= c->p.sort_every) ||
                             (fb_high[s] && b.steps_since_sort >= 2);
            const bool tail = tiled && (b.n_sorted == 0 || (b.n - b.n_sorted) * 16 > b.n_sorted);
            if (tail || due) QK_TRY(sort_species<real>(c, s));
        }
        // a merge decided after the previous step runs first (pmerging before
        // birth_from_vacuum, src/main.cpp:2321-2322)
        if (c->merge_due) {
            QK_TRY(merge_impl<real>(c, c->p.merging, c->merge_ppd));
            c->merge_due = false;
        }
        const Current<real> jc = current_of<real>(c, c->jcur), jn = current_of<real>(c, c->jcur ^ 1);
        HIP_TRY(launch_zero(c->stream, c->d_diag, kDiagLen));
        // padvance (src/advance.cpp:90-387); the push kernels erase the
        // particles p_boundary (src/plist.cpp:73-102) would: those outside
        // is_inside, which only a moving-window shift leaves behind
        HIP_TRY(hipEventRecord(c->ev[3 * l], c->stream));
        if (qed) QK_TRY(qed_particle_phase<real>(c, f, jc));
        for (int s = 0; s < QK_MAX_SPECIES && !qed; ++s) {
            SpeciesBuf &b = c->sp[s];
            if (!b.active || b.n == 0) continue;
            PushArgs a{};
            a.pusher = c->p.pusher;
            a.kind = b.kind;
            a.cmr = b.cmr;
            a.external_bz = c->p.external_bz;
            a.diag = c->d_diag + s * kDiagPerSpecies;
            a.energy_mode = b.energy_mode;
            a.nonfinite = c->d_diag + kNonfiniteSlot;
            const Species<real> v = view<real>(b.buf[b.cur]);
            if (tp.tangent && !tp.replay && b.kind == 0 && tp.n[s] > 0) {
                // the tangent of this push, from the state before it
                double *tr[7];
                tangent_records(c, s, tr);
                const Current<double> tj = tangent_current(c, tp.tjcur);
                double *dj[3] = {tj.jx, tj.jy, tj.jz};
                HIP_TRY(launch_tlm_particles<real>(c->stream, c->g, f, tangent_fields(c), v, tp.n[s], a, dj, tr));
            }
            if (tiled && b.kind == 0) {
                a.chunks = b.chunks;
                a.nchunks = b.d_nchunks;
                a.max_chunks = b.nchunks;
                a.tg = c->tg;
                a.tg.ishift = b.ishift;
                HIP_TRY(launch_push_tiled<real>(c->stream, c->g, f, jc, v, a));
                a.first = b.n_sorted;   // the unsorted tail, if any
                a.n = b.n - b.n_sorted;
            } else {
                a.first = 0;
                a.n = b.n;
            }
            HIP_TRY(launch_push_global<real>(c->stream, c->g, f, jc, v, a));
            ++b.steps_since_sort;
        }
        HIP_TRY(hipEventRecord(c->ev[3 * l + 1], c->stream));
        // fadvance (src/advance.cpp:13-18): B half, E, B half, interpolate_be.  The
        // first half step also sums compute_energy's field term, which sees the
        // fields before fadvance (src/main.cpp:2327)
        HIP_TRY(launch_maxwell_b<real>(c->stream, c->p.solver, c->g, c->coef, f,
                                       c->d_diag + QK_MAX_SPECIES * kDiagPerSpecies));
        QK_TRY(inject_impl<real>(c, 1, 0.0));
        HIP_TRY(launch_maxwell_e<real>(c->stream, c->p.solver, c->g, c->coef, f, jc, jn));
        QK_TRY(inject_impl<real>(c, 0, 0.5));
        HIP_TRY(launch_maxwell_b<real>(c->stream, c->p.solver, c->g, c->coef, f, nullptr));
        QK_TRY(inject_impl<real>(c, 1, 1.0));
        HIP_TRY(launch_interpolate_be<real>(c->stream, c->g, f));
        if (tp.tangent && !tp.replay) {
            // the tangent fields advance with the same (linear) field step
            const Fields<double> tf = tangent_fields(c);
            const Current<double> tjc = tangent_current(c, tp.tjcur), tjn = tangent_current(c, tp.tjcur ^ 1);
            HIP_TRY(launch_maxwell_b<double>(c->stream, c->p.solver, c->g, c->coef, tf, nullptr));
            HIP_TRY(launch_maxwell_e<double>(c->stream, c->p.solver, c->g, c->coef, tf, tjc, tjn));
            HIP_TRY(launch_maxwell_b<double>(c->stream, c->p.solver, c->g, c->coef, tf, nullptr));
            HIP_TRY(launch_interpolate_be<double>(c->stream, c->g, tf));
            tp.tjcur ^= 1;
        }
        HIP_TRY(hipEventRecord(c->ev[3 * l + 2], c->stream));
        c->jcur ^= 1;
        HIP_TRY(hipMemcpyAsync(hd + (size_t)l * kDiagLen, c->d_diag, kDiagLen * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        if (qed) QK_TRY(qed_download_counts(c));
        if (c->p.merging) {
            HIP_TRY(hipStreamSynchronize(c->stream));
            merge_condition(c, hd + (size_t)l * kDiagLen);
        }
        ++c->step_index;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    float t_part = 0, t_field = 0, t_tot = 0;
    for (int l = 0; l < nsteps; ++l) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, c->ev[3 * l], c->ev[3 * l + 1]));
        HIP_TRY(hipEventElapsedTime(&b, c->ev[3 * l + 1], c->ev[3 * l + 2]));
        t_part += a;
        t_field += b;
    }
    HIP_TRY(hipEventElapsedTime(&t_tot, c->ev[0], c->ev[3 * (nsteps - 1) + 2]));
    double fb = 0, live = 0;
    for (int l = 0; l < nsteps; ++l) {
        const double *h = hd + (size_t)l * kDiagLen;
        for (int s = 0; s < QK_MAX_SPECIES; ++s) {
            live += h[s * kDiagPerSpecies + 2];
            fb += h[s * kDiagPerSpecies + 3];
        }
        if (diag) unpack_diag(c, h, diag[l]);
    }
    c->fallback_frac = live > 0 ? fb / live : 0;
    {
        const double *h = hd + (size_t)(nsteps - 1) * kDiagLen;
        for (int s = 0; s < QK_MAX_SPECIES; ++s) {
            SpeciesBuf &b = c->sp[s];
            const double n_s = h[s * kDiagPerSpecies + 2];
            // a sort after the last push leaves the measurement stale
            b.fb_last = (n_s > 0 && b.steps_since_sort > 0) ? h[s * kDiagPerSpecies + 3] / n_s : -1;
        }
    }
    c->ms_total = t_tot;
    c->ms_part = t_part;
    c->ms_field = t_field;
    c->have_step = true;
    return QK_OK;
}

template <typename T>
void aos_to_soa(const double *aos, int64_t n, T *a, T *b, T *c)
{
    for (int64_t m = 0; m < n; ++m) {
        a[m] = (T)aos[3 * m];
        b[m] = (T)aos[3 * m + 1];
        c[m] = (T)aos[3 * m + 2];
    }
}

int upload3(qk_ctx *c, const double *aos, void *d0, void *d1, void *d2)
{
    const int64_t n = c->g.ncell;
    std::vector<char> buf((size_t)n * c->rsize * 3);
    if (c->rsize == 4) {
        float *h = (float *)buf.data();
        aos_to_soa<float>(aos, n, h, h + n, h + 2 * n);
    } else {
        double *h = (double *)buf.data();
        aos_to_soa<double>(aos, n, h, h + n, h + 2 * n);
    }
    const size_t bytes = (size_t)n * c->rsize;
    HIP_TRY(hipMemcpy(d0, buf.data(), bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d1, buf.data() + bytes, bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d2, buf.data() + 2 * bytes, bytes, hipMemcpyHostToDevice));
    return QK_OK;
}

int download3(qk_ctx *c, double *aos, const void *d0, const void *d1, const void *d2)
{
    const int64_t n = c->g.ncell;
    const size_t bytes = (size_t)n * c->rsize;
    std::vector<char> buf(bytes * 3);
    HIP_TRY(hipMemcpy(buf.data(), d0, bytes, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(buf.data() + bytes, d1, bytes, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(buf.data() + 2 * bytes, d2, bytes, hipMemcpyDeviceToHost));
    for (int64_t m = 0; m < n; ++m)
        for (int a = 0; a < 3; ++a)
            aos[3 * m + a] = c->rsize == 4 ? (double)((const float *)buf.data())[a * n + m]
                                           : ((const double *)buf.data())[a * n + m];
    return QK_OK;
}

}  // namespace

extern "C" {

const char *qk_last_error(void) { return g_err.c_str(); }

int qk_create(const qk_params *params, qk_ctx **out)
{
    if (!params || !out) return fail(QK_ERR_INVALID, "null argument");
    *out = nullptr;
    const qk_params &p = *params;
    if (p.nx < 5 || p.ny < 5 || p.nz < 5) return fail(QK_ERR_INVALID, "grid must be at least 5 cells per axis");
    if (p.nx > 2048 || p.ny > 1024 || p.nz > 1024)
        return fail(QK_ERR_INVALID, "grid exceeds 2048 x 1024 x 1024 (packed cell index)");
    {
        // cell and sort-key indices are 32-bit: the cell count and the tile
        // keys (whole 8x8x8 tiles) must stay below 2^31
        const int64_t ncell = (int64_t)p.nx * p.ny * p.nz;
        const int64_t nkeys = (int64_t)((p.nx + kTx - 1) / kTx) * ((p.ny + kTy - 1) / kTy) *
                              ((p.nz + kTz - 1) / kTz) * kTileCells;
        if (ncell >= ((int64_t)1 << 31) || nkeys >= ((int64_t)1 << 31))
            return fail(QK_ERR_INVALID, "grid too large: cell and tile-key indices must stay below 2^31");
    }
    if (!(p.dx > 0 && p.dy > 0 && p.dz > 0 && p.dt > 0)) return fail(QK_ERR_INVALID, "dx, dy, dz, dt must be > 0");
    if (p.solver < 0 || p.solver > 2) return fail(QK_ERR_INVALID, "unknown solver");
    if (p.pusher < 0 || p.pusher > 1) return fail(QK_ERR_INVALID, "unknown pusher");
    if (p.precision != QK_FP32 && p.precision != QK_FP64) return fail(QK_ERR_INVALID, "unknown precision");
    if (p.deposit != QK_DEPOSIT_TILED && p.deposit != QK_DEPOSIT_GLOBAL) return fail(QK_ERR_INVALID, "unknown deposit");
    if (p.n_ion < 0 || p.n_ion > QK_MAX_ION) return fail(QK_ERR_INVALID, "n_ion out of range");
    if (p.sort_every < 0) return fail(QK_ERR_INVALID, "sort_every must be >= 0");
    if (!(p.sort_fallback >= 0 && p.sort_fallback <= 1)) return fail(QK_ERR_INVALID, "sort_fallback must be in [0, 1]");
    if (p.qed && !(p.e_s > 0)) return fail(QK_ERR_INVALID, "QED needs e_s > 0");
    if (p.merging < 0 || p.merging > 2) return fail(QK_ERR_INVALID, "merging must be 0, 1 (ti) or 2 (nl)");
    if (p.merging && !(p.crnp > 0)) return fail(QK_ERR_INVALID, "merging needs crnp > 0");
    // the reference's deposit splits a path at no more than one face per axis
    if (p.dt >= p.dx || p.dt >= p.dy || p.dt >= p.dz)
        return fail(QK_ERR_INVALID, "dt must be below dx, dy and dz (one face crossing per axis and step)");

    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(QK_ERR_HIP, "no HIP device available");
    if (p.device < 0 || p.device >= ndev || p.device >= kMaxDevices)
        return fail(QK_ERR_INVALID, "device ordinal out of range");
    e = hipSetDevice(p.device);
    if (e != hipSuccess) return fail(QK_ERR_HIP, hipGetErrorString(e));
    if (int rc = resolve_scratch(p.device)) return rc;

    qk_ctx *c = new qk_ctx;
    c->p = p;
    if (c->p.norm_n == 0) c->p.norm_n = 1;
    if (c->p.norm_ef == 0) c->p.norm_ef = 1;
    if (c->p.norm_ep == 0) c->p.norm_ep = 1;
    c->g.nx = p.nx;
    c->g.ny = p.ny;
    c->g.nz = p.nz;
    c->g.sy = p.nz;
    c->g.sx = p.ny * p.nz;
    c->g.ncell = p.nx * p.ny * p.nz;
    c->g.dx = p.dx;
    c->g.dy = p.dy;
    c->g.dz = p.dz;
    c->g.dt = p.dt;
    c->tg.ntx = (p.nx + kTx - 1) / kTx;
    c->tg.nty = (p.ny + kTy - 1) / kTy;
    c->tg.ntz = (p.nz + kTz - 1) / kTz;
    c->tg.ntiles = c->tg.ntx * c->tg.nty * c->tg.ntz;
    c->tg.nkeys = c->tg.ntiles * kTileCells;
    c->tg.ishift = 0;
    c->rsize = p.precision == QK_FP64 ? 8 : 4;
    // solver coefficients (src/maxwell.cpp:24-33, 238-249)
    for (int a = 0; a < 3; ++a) {
        c->coef.a[a] = 0;
        c->coef.b[a] = 1;
    }
    const double d[3] = {p.dx, p.dy, p.dz};
    if (p.solver == QK_SOLVER_NDFX) {
        const double kappa = 1.25 * (std::sqrt(3.0) - 1) * 0.5 / std::sqrt(3.0);
        for (int a = 0; a < 3; ++a) {
            c->coef.b[a] = 1 - kappa * p.dx * p.dx / (d[a] * d[a]);
            c->coef.a[a] = 0.5 * (1 - c->coef.b[a]);
        }
        c->coef.b[0] = 1 - kappa;
        c->coef.a[0] = 0.5 * (1 - c->coef.b[0]);
    } else if (p.solver == QK_SOLVER_FIVE_POINT) {
        const double PI = 3.141592653589793;
        for (int a = 0; a < 3; ++a) {
            c->coef.a[a] = 0.25 * (1 - d[a] / p.dt * std::sin(0.5 * PI * std::fmin(1.0, p.dt / d[a])));
            c->coef.b[a] = 1 - 2 * c->coef.a[a];
        }
    }
    auto bail = [&](int code) {
        qk_destroy(c);
        return code;
    };
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(fail(QK_ERR_HIP, hipGetErrorString(e)));
    const size_t fbytes = ((size_t)c->g.ncell * c->rsize + 255) & ~(size_t)255;
    if ((e = hipMalloc(&c->fmem, fbytes * 15)) != hipSuccess)
        return bail(fail(QK_ERR_HIP, std::string("field allocation: ") + hipGetErrorString(e)));
    if ((e = hipMemset(c->fmem, 0, fbytes * 15)) != hipSuccess) return bail(fail(QK_ERR_HIP, hipGetErrorString(e)));
    for (int a = 0; a < 15; ++a) c->F[a] = (char *)c->fmem + a * fbytes;
    if ((e = hipMalloc(&c->d_diag, kDiagLen * sizeof(double))) != hipSuccess)
        return bail(fail(QK_ERR_HIP, hipGetErrorString(e)));
    if (p.qed) {
        if ((e = hipMalloc(&c->d_count, (QK_MAX_SPECIES + 1) * sizeof(int))) != hipSuccess ||
            (e = hipHostMalloc((void **)&c->h_count, (QK_MAX_SPECIES + 1) * sizeof(int))) != hipSuccess)
            return bail(fail(QK_ERR_HIP, hipGetErrorString(e)));
        // the species that QED appends to exist from the start
        for (int s : {QK_ELECTRON, QK_POSITRON, QK_PHOTON})
            if (int rc = ensure_species(c, s, 0)) return bail(rc);
    }
    *out = c;
    return QK_OK;
}

int qk_destroy(qk_ctx *c)
{
    if (!c) return QK_OK;
    // release everything even if one call fails; report the first failure
    hipError_t first = hipSuccess;
    auto note = [&](hipError_t e) {
        if (first == hipSuccess && e != hipSuccess) first = e;
    };
    if (c->stream) note(hipStreamSynchronize(c->stream));
    for (auto &b : c->sp) {
        for (auto &k : b.buf)
            if (k.mem) note(hipFree(k.mem));
        if (b.chunks) note(hipFree(b.chunks));
        if (b.d_nchunks) note(hipFree(b.d_nchunks));
        if (b.tile_start) note(hipFree(b.tile_start));
    }
    int *sb[8] = {c->s_counts, c->s_imm, c->s_offsets, c->s_tile_start, c->s_tile_mult, c->s_chunk_n, c->s_chunk_off,
                  c->s_scratch};
    for (int *p : sb)
        if (p) note(hipFree(p));
    if (c->h_sortinfo) note(hipHostFree(c->h_sortinfo));
    if (c->fmem) note(hipFree(c->fmem));
    if (c->d_diag) note(hipFree(c->d_diag));
    if (c->h_diag) note(hipHostFree(c->h_diag));
    if (c->d_count) note(hipFree(c->d_count));
    if (c->h_count) note(hipHostFree(c->h_count));
    if (c->d_merge) note(hipFree(c->d_merge));
    if (c->h_merge) note(hipHostFree(c->h_merge));
    for (void *m : c->tape.ckpt) note(hipFree(m));
    if (c->tape.seg) note(hipFree(c->tape.seg));
    if (c->tape.fin) note(hipFree(c->tape.fin));
    if (c->tape.lam) note(hipFree(c->tape.lam));
    if (c->tape.lpart) note(hipFree(c->tape.lpart));
    if (c->d_maxw) note(hipFree(c->d_maxw));
    if (c->h_maxw) note(hipHostFree(c->h_maxw));
    if (c->d_fill) note(hipFree(c->d_fill));
    for (auto &side : c->inj)
        if (side.mem) note(hipFree(side.mem));
    for (auto e : c->ev) note(hipEventDestroy(e));
    if (c->stream) note(hipStreamDestroy(c->stream));
    delete c;
    if (first != hipSuccess) return fail(QK_ERR_HIP, std::string("qk_destroy: ") + hipGetErrorString(first));
    return QK_OK;
}

int qk_set_fields(qk_ctx *c, const double *E, const double *B)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    c->tape.dirty = c->tape.active;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (E) QK_TRY(upload3(c, E, c->F[0], c->F[1], c->F[2]));
    if (B) QK_TRY(upload3(c, B, c->F[3], c->F[4], c->F[5]));
    return QK_OK;
}

int qk_get_fields(qk_ctx *c, double *E, double *B, double *J, double *Be)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (E) QK_TRY(download3(c, E, c->F[0], c->F[1], c->F[2]));
    if (B) QK_TRY(download3(c, B, c->F[3], c->F[4], c->F[5]));
    if (Be) QK_TRY(download3(c, Be, c->F[6], c->F[7], c->F[8]));
    if (J) {
        const int w = c->have_step ? (c->jcur ^ 1) : c->jcur;   // buffer of the last step
        QK_TRY(download3(c, J, c->F[9 + 3 * w], c->F[10 + 3 * w], c->F[11 + 3 * w]));
    }
    return QK_OK;
}

int qk_add_particles(qk_ctx *c, int s, int64_t n, const double *x, const double *y, const double *z, const double *ux,
                     const double *uy, const double *uz, const double *q, const int32_t *ids, int64_t *kept)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    if (!species_valid(c, s)) return fail(QK_ERR_INVALID, "invalid species");
    if (n < 0 || (n > 0 && (!x || !y || !z || !ux || !uy || !uz || !q))) return fail(QK_ERR_INVALID, "bad particle arrays");
    c->tape.dirty = c->tape.active;
    // place(): keep particles in is_inside cells (src/plist.cpp:22-31)
    std::vector<int64_t> keep;
    keep.reserve((size_t)n);
    for (int64_t m = 0; m < n; ++m) {
        if (!(x[m] >= 0 && y[m] >= 0 && z[m] >= 0 && x[m] < c->p.nx && y[m] < c->p.ny && z[m] < c->p.nz)) continue;
        const int i = (int)x[m], j = (int)y[m], k = (int)z[m];
        if (i > 0 && i < c->p.nx - 2 && j > 0 && j < c->p.ny - 2 && k > 0 && k < c->p.nz - 2) keep.push_back(m);
    }
    const int64_t nk = (int64_t)keep.size();
    QK_TRY(ensure_species(c, s, nk));
    SpeciesBuf &b = c->sp[s];
    SoA &dst = b.buf[b.cur];
    const size_t r = c->rsize;
    std::vector<int> cell((size_t)nk), id((size_t)nk);
    std::vector<char> buf((size_t)nk * r * 8);
    for (int64_t t = 0; t < nk; ++t) {
        const int64_t m = keep[(size_t)t];
        const int i = (int)x[m], j = (int)y[m], k = (int)z[m];
        cell[(size_t)t] = pack_cell(i, j, k);
        id[(size_t)t] = ids ? ids[m] : 0;
        const double v[8] = {x[m] - i, y[m] - j, z[m] - k, ux[m], uy[m], uz[m], q[m], 0.0};
        for (int a = 0; a < 8; ++a) {
            if (r == 4) {
                float fv = (float)v[a];
                if (a < 3 && fv >= 1.0f) fv = 0.99999994f;   // offsets stay below 1
                ((float *)buf.data())[(size_t)a * nk + t] = fv;
            } else {
                ((double *)buf.data())[(size_t)a * nk + t] = v[a];
            }
        }
    }
    if (nk > 0) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(dst.cell + b.n, cell.data(), (size_t)nk * 4, hipMemcpyHostToDevice));
        for (int a = 0; a < 8; ++a)
            if (dst.a[a])
                HIP_TRY(hipMemcpy((char *)dst.a[a] + (size_t)b.n * r, buf.data() + (size_t)a * nk * r, (size_t)nk * r,
                                  hipMemcpyHostToDevice));
        if (dst.id) HIP_TRY(hipMemcpy(dst.id + b.n, id.data(), (size_t)nk * 4, hipMemcpyHostToDevice));
        b.n += nk;
    }
    if (kept) *kept = nk;
    return QK_OK;
}

int qk_load_uniform_plasma(qk_ctx *c, int s, int ppc, double density, double sigma, uint64_t seed, int ilo, int ihi,
                           int64_t *added)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    c->tape.dirty = c->tape.active;
    if (!species_valid(c, s) || s == QK_PHOTON) return fail(QK_ERR_INVALID, "invalid species for a plasma");
    if (ppc <= 0) return fail(QK_ERR_INVALID, "ppc must be positive");
    if (ilo < 1) ilo = 1;
    if (ihi > c->p.nx - 2) ihi = c->p.nx - 2;
    const int64_t total = ihi > ilo ? (int64_t)(ihi - ilo) * (c->p.ny - 3) * (c->p.nz - 3) * ppc : 0;
    QK_TRY(ensure_species(c, s, total));
    SpeciesBuf &b = c->sp[s];
    const double q = (b.cmr < 0 ? -1.0 : 1.0) * density / ppc;   // init.cpp:413-414
    if (total > 0) {
        if (c->rsize == 4)
            HIP_TRY(launch_load_uniform<float>(c->stream, c->g, view<float>(b.buf[b.cur]), b.n, ppc, q, sigma, seed, ilo, ihi));
        else
            HIP_TRY(launch_load_uniform<double>(c->stream, c->g, view<double>(b.buf[b.cur]), b.n, ppc, q, sigma, seed, ilo, ihi));
        HIP_TRY(hipStreamSynchronize(c->stream));
        b.n += total;
    }
    if (added) *added = total;
    return QK_OK;
}

int qk_init_laser_cos(qk_ctx *c, const double L[16])
{
    if (!c || !L) return fail(QK_ERR_INVALID, "null argument");
    c->tape.dirty = c->tape.active;
    hipError_t e = c->rsize == 4 ? launch_laser_cos<float>(c->stream, c->g, fields_of<float>(c), L)
                                 : launch_laser_cos<double>(c->stream, c->g, fields_of<double>(c), L);
    if (e == hipErrorInvalidValue) return fail(QK_ERR_INVALID, "laser direction undefined (x1 = y0 = z0 = 0)");
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QK_OK;
}

int qk_init_boundaries(qk_ctx *c)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    c->tape.dirty = c->tape.active;
    if (c->rsize == 4) {
        HIP_TRY(launch_init_boundaries<float>(c->stream, c->p.solver, c->g, fields_of<float>(c)));
        HIP_TRY(launch_interpolate_be<float>(c->stream, c->g, fields_of<float>(c)));
    } else {
        HIP_TRY(launch_init_boundaries<double>(c->stream, c->p.solver, c->g, fields_of<double>(c)));
        HIP_TRY(launch_interpolate_be<double>(c->stream, c->g, fields_of<double>(c)));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QK_OK;
}

int qk_step(qk_ctx *c, int nsteps, qk_diag *diag)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    if (nsteps < 0) return fail(QK_ERR_INVALID, "negative step count");
    if (nsteps == 0) return QK_OK;
    return c->rsize == 4 ? step_impl<float>(c, nsteps, diag) : step_impl<double>(c, nsteps, diag);
}

int qk_sync(qk_ctx *c)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QK_OK;
}

int qk_particle_count(qk_ctx *c, int s, int64_t *n)
{
    if (!c || !n) return fail(QK_ERR_INVALID, "null argument");
    if (!species_valid(c, s)) return fail(QK_ERR_INVALID, "invalid species");
    SpeciesBuf &b = c->sp[s];
    *n = 0;
    if (b.n == 0) return QK_OK;
    std::vector<int> cell((size_t)b.n);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(cell.data(), b.buf[b.cur].cell, (size_t)b.n * 4, hipMemcpyDeviceToHost));
    int64_t live = 0;
    for (int v : cell) live += v >= 0;
    *n = live;
    return QK_OK;
}

int qk_get_particles(qk_ctx *c, int s, int64_t cap, double *rec, int32_t *ids, int64_t *n_out)
{
    if (!c || !n_out) return fail(QK_ERR_INVALID, "null argument");
    if (!species_valid(c, s)) return fail(QK_ERR_INVALID, "invalid species");
    SpeciesBuf &b = c->sp[s];
    *n_out = 0;
    if (b.n == 0) return QK_OK;
    if (!rec) return fail(QK_ERR_INVALID, "null output buffer");
    const SoA &src = b.buf[b.cur];
    const size_t r = c->rsize;
    const int64_t n = b.n;
    std::vector<int> cell((size_t)n), id((size_t)n);
    std::vector<char> buf((size_t)n * r * 8, 0);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(cell.data(), src.cell, (size_t)n * 4, hipMemcpyDeviceToHost));
    for (int a = 0; a < 8; ++a)
        if (src.a[a]) HIP_TRY(hipMemcpy(buf.data() + (size_t)a * n * r, src.a[a], (size_t)n * r, hipMemcpyDeviceToHost));
    if (src.id) HIP_TRY(hipMemcpy(id.data(), src.id, (size_t)n * 4, hipMemcpyDeviceToHost));
    auto get = [&](int a, int64_t m) {
        return r == 4 ? (double)((const float *)buf.data())[(size_t)a * n + m] : ((const double *)buf.data())[(size_t)a * n + m];
    };
    int64_t o = 0;
    for (int64_t m = 0; m < n; ++m) {
        const int cc = cell[(size_t)m];
        if (cc < 0) continue;
        if (o >= cap) return fail(QK_ERR_CAPACITY, "output buffer too small");
        double *w = rec + 10 * o;
        w[0] = cell_i(cc) + get(0, m);
        w[1] = cell_j(cc) + get(1, m);
        w[2] = cell_k(cc) + get(2, m);
        w[3] = get(3, m);
        w[4] = get(4, m);
        w[5] = get(5, m);
        const double u2 = w[3] * w[3] + w[4] * w[4] + w[5] * w[5];
        w[6] = b.kind == 1 ? std::sqrt(u2) : std::sqrt(1 + u2);
        w[7] = get(6, m);
        w[8] = b.cmr;
        w[9] = get(7, m);
        if (ids) ids[o] = src.id ? id[(size_t)m] : 0;
        ++o;
    }
    *n_out = o;
    return QK_OK;
}

int qk_compute_rho(qk_ctx *c, double *rho)
{
    if (!c || !rho) return fail(QK_ERR_INVALID, "null argument");
    const size_t bytes = (size_t)c->g.ncell * 3 * sizeof(double);
    double *d = nullptr;
    HIP_TRY(hipMalloc(&d, bytes));
    hipError_t e = launch_zero(c->stream, d, (size_t)c->g.ncell * 3);
    const int slots[3] = {QK_ELECTRON, QK_POSITRON, QK_PHOTON};
    for (int t = 0; t < 3 && e == hipSuccess; ++t) {
        SpeciesBuf &b = c->sp[slots[t]];
        if (!b.active || b.n == 0) continue;
        e = c->rsize == 4 ? launch_compute_rho<float>(c->stream, c->g, view<float>(b.buf[b.cur]), b.n, d, t)
                          : launch_compute_rho<double>(c->stream, c->g, view<double>(b.buf[b.cur]), b.n, d, t);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(rho, d, bytes, hipMemcpyDeviceToHost);
    const hipError_t ef = hipFree(d);
    HIP_TRY(e);
    HIP_TRY(ef);
    return QK_OK;
}

// qk_compute_rho without allocating: the node charges accumulate in this
// device's `scratch` (resolved by qk_create), cleared by a fill kernel first.
// Calls that overlap on one GPU (two contexts on two threads) share it and race.
int qk_compute_rho_noalloc(qk_ctx *c, double *rho)
{
    if (!c || !rho) return fail(QK_ERR_INVALID, "null argument");
    if ((int64_t)c->g.ncell > kNoallocCells)
        return fail(QK_ERR_CAPACITY, "qk_compute_rho_noalloc: the grid has " + std::to_string(c->g.ncell) +
                                         " cells, above its limit of " + std::to_string(kNoallocCells) +
                                         " (qk_compute_rho has none)");
    double *d = scratch_ptr(c->p.device);
    if (!d) return fail(QK_ERR_STATE, "qk_compute_rho_noalloc: no scratch resolved for this device");
    const size_t n = (size_t)c->g.ncell * 3;
    HIP_TRY(launch_zero(c->stream, d, n));
    const int slots[3] = {QK_ELECTRON, QK_POSITRON, QK_PHOTON};
    for (int t = 0; t < 3; ++t) {
        SpeciesBuf &b = c->sp[slots[t]];
        if (!b.active || b.n == 0) continue;
        HIP_TRY(c->rsize == 4 ? launch_compute_rho<float>(c->stream, c->g, view<float>(b.buf[b.cur]), b.n, d, t)
                              : launch_compute_rho<double>(c->stream, c->g, view<double>(b.buf[b.cur]), b.n, d, t));
    }
    HIP_TRY(hipMemcpyAsync(rho, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QK_OK;
}

int qk_add_laser_source(qk_ctx *c, const double L[16], int side, int plane)
{
    if (!c || !L) return fail(QK_ERR_INVALID, "null argument");
    const int nx = c->p.nx;
    if (side == 0 ? !(plane >= 2 && plane <= nx - 6) : !(side == 1 && plane >= 6 && plane <= nx - 2))
        return fail(QK_ERR_INVALID, "laser source: side 0 needs 2 <= plane <= nx - 6, side 1 6 <= plane <= nx - 2");
    InjSide &sd = c->inj[side];
    if (sd.nsrc == kMaxSources) return fail(QK_ERR_CAPACITY, "at most 4 laser sources per side");
    if (sd.nsrc > 0 && sd.plane != plane) return fail(QK_ERR_INVALID, "the sources of one side share one plane");
    LaserFrame F;
    if (!laser_frame(c->g, L, F)) return fail(QK_ERR_INVALID, "laser source: x1, y0, z0 all zero");
    if (!sd.mem) {
        sd.W = side == 0 ? plane + 6 : nx - plane + 6;
        sd.gs = c->g;
        sd.gs.nx = 2 * sd.W + kInjGap;
        sd.gs.ncell = sd.gs.nx * c->g.sx;
        const size_t n = ((size_t)sd.gs.ncell * c->rsize + 255) & ~(size_t)255;
        HIP_TRY(hipMalloc(&sd.mem, 12 * n));
        HIP_TRY(hipMemset(sd.mem, 0, 12 * n));   // the current sets stay zero
        sd.plane = plane;
    }
    sd.src[sd.nsrc] = F;
    sd.t0[sd.nsrc] = c->step_index;
    sd.shifts0[sd.nsrc] = c->window_shifts;
    ++sd.nsrc;
    return QK_OK;
}

int qk_pmerging(qk_ctx *c, int mode, const double *ppd)
{
    if (!c || !ppd) return fail(QK_ERR_INVALID, "null argument");
    c->tape.dirty = c->tape.active;
    if (mode != 1 && mode != 2) return fail(QK_ERR_INVALID, "merging mode must be 1 (ti) or 2 (nl)");
    for (int s = 0; s < QK_ION0 + c->p.n_ion; ++s)
        if (!(ppd[s] >= 0 && ppd[s] < 1)) return fail(QK_ERR_INVALID, "ppd must lie in [0, 1)");
    return c->rsize == 4 ? merge_impl<float>(c, mode, ppd) : merge_impl<double>(c, mode, ppd);
}

int qk_moving_window(qk_ctx *c)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    c->tape.dirty = c->tape.active;
    if (c->rsize == 4) HIP_TRY(launch_shift_fields<float>(c->stream, c->g, fields_of<float>(c)));
    else HIP_TRY(launch_shift_fields<double>(c->stream, c->g, fields_of<double>(c)));
    for (int s = 0; s < QK_MAX_SPECIES; ++s) {
        SpeciesBuf &b = c->sp[s];
        if (!b.active) continue;
        HIP_TRY(launch_shift_particles(c->stream, b.buf[b.cur].cell, b.n));
        ++b.ishift;   // the tiles of the last sort moved with their particles
    }
    ++c->window_shifts;   // grid plane i now lies at (i + shifts) dx in the frame of the laser sources
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QK_OK;
}

int qk_max_w(qk_ctx *c, int ilo, int ihi, double *max_w)
{
    if (!c || !max_w) return fail(QK_ERR_INVALID, "null argument");
    if (ilo < 0 || ihi > c->p.nx) return fail(QK_ERR_INVALID, "planes outside [0, nx)");
    if (!c->d_maxw) {
        HIP_TRY(hipMalloc(&c->d_maxw, sizeof(unsigned long long)));
        HIP_TRY(hipHostMalloc((void **)&c->h_maxw, sizeof(unsigned long long)));
    }
    HIP_TRY(launch_fill_int(c->stream, (int *)c->d_maxw, 2, 0));
    if (c->rsize == 4) HIP_TRY(launch_max_w<float>(c->stream, c->g, fields_of<float>(c), ilo, ihi, c->d_maxw));
    else HIP_TRY(launch_max_w<double>(c->stream, c->g, fields_of<double>(c), ilo, ihi, c->d_maxw));
    HIP_TRY(hipMemcpyAsync(c->h_maxw, c->d_maxw, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(max_w, c->h_maxw, sizeof(double));
    return QK_OK;
}

int qk_fill_cells(qk_ctx *c, int s, int64_t ncells, const int32_t *cells, const double *density,
                  const int32_t npic[3], double ux0, int32_t first_id, int64_t *added)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    c->tape.dirty = c->tape.active;
    if (!species_valid(c, s)) return fail(QK_ERR_INVALID, "invalid species");
    if (ncells < 0 || (ncells > 0 && (!cells || !density)) || !npic) return fail(QK_ERR_INVALID, "bad cell list");
    if (npic[0] <= 0 || npic[1] <= 0 || npic[2] <= 0 || (int64_t)npic[0] * npic[1] * npic[2] > 4096)
        return fail(QK_ERR_INVALID, "npic must be positive with at most 4096 particles per cell");
    if (added) *added = 0;
    // place() keeps the cells inside is_inside (src/plist.cpp:22-31); the ids
    // of a listed cell are fixed by its index in the list
    const int npt = npic[0] * npic[1] * npic[2];
    std::vector<int> kc, kid;
    std::vector<double> kd;
    for (int64_t m = 0; m < ncells; ++m) {
        const int i = cells[3 * m], j = cells[3 * m + 1], k = cells[3 * m + 2];
        if (!(i > 0 && i < c->p.nx - 2 && j > 0 && j < c->p.ny - 2 && k > 0 && k < c->p.nz - 2)) continue;
        kc.insert(kc.end(), {i, j, k});
        kd.push_back(density[m]);
        kid.push_back(first_id + (int32_t)(m * npt));
    }
    const int64_t nk = (int64_t)kd.size();
    if (nk == 0) return QK_OK;
    const int64_t total = nk * npt;
    QK_TRY(ensure_species(c, s, total));
    const size_t need = (size_t)nk * (3 * sizeof(int) + sizeof(double) + sizeof(int));
    if (need > c->fill_cap) {
        if (c->d_fill) HIP_TRY(hipFree(c->d_fill));
        c->d_fill = nullptr;
        c->fill_cap = 0;
        HIP_TRY(hipMalloc(&c->d_fill, need));
        c->fill_cap = need;
    }
    double *dd = (double *)c->d_fill;
    int *dc = (int *)(dd + nk), *di = dc + 3 * nk;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(dd, kd.data(), (size_t)nk * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dc, kc.data(), (size_t)nk * 3 * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di, kid.data(), (size_t)nk * sizeof(int), hipMemcpyHostToDevice));
    SpeciesBuf &b = c->sp[s];
    const double sign = b.cmr < 0 ? -1.0 : 1.0;
    if (c->rsize == 4)
        HIP_TRY(launch_fill_cells<float>(c->stream, view<float>(b.buf[b.cur]), b.n, nk, dc, dd, di, npic, sign, ux0));
    else
        HIP_TRY(launch_fill_cells<double>(c->stream, view<double>(b.buf[b.cur]), b.n, nk, dc, dd, di, npic, sign, ux0));
    HIP_TRY(hipStreamSynchronize(c->stream));
    b.n += total;
    if (added) *added = total;
    return QK_OK;
}

int qk_tape_begin(qk_ctx *c, int every)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    if (every <= 0) return fail(QK_ERR_INVALID, "checkpoint interval must be positive");
    const qk_params &p = c->p;
    if (p.qed || p.merging || p.sort_every != 0 || p.sort_fallback != 0)
        return fail(QK_ERR_INVALID, "the adjoint needs qed = 0, merging = 0, sort_every = 0, sort_fallback = 0");
    if (c->sp[QK_PHOTON].active && c->sp[QK_PHOTON].n > 0) return fail(QK_ERR_INVALID, "the adjoint covers charged species only");
    // With the tiled deposit the particles are sorted into tiles once, here;
    // no later sort runs (sort_every = 0, nothing is appended), so every
    // particle keeps its slot from snapshot to snapshot.  The tiles' windows
    // then age, and particles that leave them go through the kernel's global
    // fallback.
    if (p.deposit == QK_DEPOSIT_TILED)
        for (int s = 0; s < QK_MAX_SPECIES; ++s)
            if (c->sp[s].active && c->sp[s].n > 0 && c->sp[s].kind == 0)
                QK_TRY(c->rsize == 4 ? sort_species<float>(c, s) : sort_species<double>(c, s));
    tape_free(c);
    Tape &t = c->tape;
    t.every = every;
    t.field_bytes = (((size_t)c->g.ncell * c->rsize + 255) & ~(size_t)255) * 15;
    t.bytes = t.field_bytes;
    for (int s = 0; s < QK_MAX_SPECIES; ++s) {
        t.n[s] = c->sp[s].active ? c->sp[s].n : 0;
        t.bytes += species_state_bytes(c, t.n[s]);
    }
    t.active = true;
    QK_TRY(tape_snapshot(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QK_OK;
}

int qk_tangent_begin(qk_ctx *c, const double *dE, const double *dB, const double *const *dpart)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    Tape &t = c->tape;
    if (!t.active || t.steps != 0) return fail(QK_ERR_STATE, "qk_tangent_begin goes right after qk_tape_begin");
    const int64_t nc = c->g.ncell;
    int64_t np_total = 0;
    for (int s = 0; s < QK_MAX_SPECIES; ++s) np_total += t.n[s];
    if (!t.tfield) HIP_TRY(hipMalloc(&t.tfield, (size_t)nc * 15 * sizeof(double)));
    if (!t.tpart && np_total > 0) HIP_TRY(hipMalloc(&t.tpart, (size_t)np_total * 7 * sizeof(double)));
    // seeds: fields AoS -> SoA (then the boundary zeroing and interpolate_be
    // the set-up applied to the fields), records AoS (x y z ux uy uz q) -> SoA
    {
        std::vector<double> h((size_t)nc * 15, 0.0);
        for (int64_t m = 0; m < nc; ++m)
            for (int a = 0; a < 3; ++a) {
                if (dE) h[(size_t)a * nc + m] = dE[3 * m + a];
                if (dB) h[(size_t)(3 + a) * nc + m] = dB[3 * m + a];
            }
        HIP_TRY(hipMemcpy(t.tfield, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        for (int s = 0; s < QK_MAX_SPECIES; ++s) {
            const int64_t n = t.n[s];
            if (n == 0) continue;
            std::vector<double> hp((size_t)n * 7, 0.0);
            if (dpart && dpart[s])
                for (int64_t p = 0; p < n; ++p)
                    for (int a = 0; a < 7; ++a) hp[(size_t)a * n + p] = dpart[s][7 * p + a];
            double *tr[7];
            tangent_records(c, s, tr);
            HIP_TRY(hipMemcpy(tr[0], hp.data(), hp.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    const Fields<double> tf = tangent_fields(c);
    HIP_TRY(launch_init_boundaries<double>(c->stream, c->p.solver, c->g, tf));
    HIP_TRY(launch_interpolate_be<double>(c->stream, c->g, tf));
    HIP_TRY(hipStreamSynchronize(c->stream));
    t.tjcur = 0;
    t.tangent = true;
    return QK_OK;
}

int qk_tangent_get(qk_ctx *c, double *dE, double *dB, double *const *dpart)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    Tape &t = c->tape;
    if (!t.tangent) return fail(QK_ERR_STATE, "no tangent: call qk_tangent_begin after qk_tape_begin");
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int64_t nc = c->g.ncell;
    std::vector<double> h((size_t)nc * 6);
    HIP_TRY(hipMemcpy(h.data(), t.tfield, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t m = 0; m < nc; ++m)
        for (int a = 0; a < 3; ++a) {
            if (dE) dE[3 * m + a] = h[(size_t)a * nc + m];
            if (dB) dB[3 * m + a] = h[(size_t)(3 + a) * nc + m];
        }
    for (int s = 0; s < QK_MAX_SPECIES; ++s) {
        const int64_t n = t.n[s];
        if (n == 0 || !dpart || !dpart[s]) continue;
        std::vector<double> hp((size_t)n * 7);
        double *tr[7];
        tangent_records(c, s, tr);
        HIP_TRY(hipMemcpy(hp.data(), tr[0], hp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < n; ++p)
            for (int a = 0; a < 7; ++a) dpart[s][7 * p + a] = hp[(size_t)a * n + p];
    }
    return QK_OK;
}

int qk_adjoint(qk_ctx *c, const double *lam_E, const double *lam_B, const double *const *lam_part, double *grad_E,
               double *grad_B, double *const *grad_part)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    if (!c->tape.active) return fail(QK_ERR_STATE, "no tape: call qk_tape_begin before the steps");
    if (c->tape.dirty) return fail(QK_ERR_STATE, "fields or particles changed outside qk_step since qk_tape_begin");
    return c->rsize == 4 ? adjoint_impl<float>(c, lam_E, lam_B, lam_part, grad_E, grad_B, grad_part)
                         : adjoint_impl<double>(c, lam_E, lam_B, lam_part, grad_E, grad_B, grad_part);
}

int qk_last_timing(qk_ctx *c, double *ms_total, double *ms_particles, double *ms_fields)
{
    if (!c) return fail(QK_ERR_INVALID, "null context");
    if (ms_total) *ms_total = c->ms_total;
    if (ms_particles) *ms_particles = c->ms_part;
    if (ms_fields) *ms_fields = c->ms_field;
    return QK_OK;
}

int qk_last_fallback(qk_ctx *c, double *fraction)
{
    if (!c || !fraction) return fail(QK_ERR_INVALID, "null argument");
    *fraction = c->fallback_frac;
    return QK_OK;
}

int qk_get_stream(qk_ctx *c, void **stream)
{
    if (!c || !stream) return fail(QK_ERR_INVALID, "null argument");
    *stream = (void *)c->stream;
    return QK_OK;
}

}  // extern "C"
