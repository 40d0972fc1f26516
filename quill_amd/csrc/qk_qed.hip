// qk_qed.hip -- the particle phase with QED on (src/advance.cpp:90-387 with
// qed_enabled) and birth_from_vacuum (src/mc.cpp:142-196) on gfx950.
//
//   k_push_qed_lepton  electrons / positrons: push, chi at n+1/2, photon
//                      emission with recoil (rewind half a step, re-gather,
//                      re-push), deposit, re-bin
//   k_photon_qed       photons: chi, Breit-Wheeler pair creation (both
//                      leptons pushed and deposited), straight flight
//   k_vacuum_pairs     Schwinger pairs, one draw per cell
//
// One thread per particle; current contributions are global atomics.  New
// particles are appended to the end of their species buffer, one slot
// counter atomic per wavefront (the lanes that append take consecutive
// slots in lane order), and are not visited again in the same step, as
// particles born during padvance are not.  Random numbers come from counter-based
// Philox streams keyed by (seed, step, slot, stream tag): statistics match
// the reference, individual events do not (get_rand is a sequential
// lagged-Fibonacci generator).
#include "qk_push.h"
#include "qk_qed.h"

namespace qk {

namespace {

constexpr uint32_t kTagEmit = 0x100;     // + species slot
constexpr uint32_t kTagPair = 0x200;
constexpr uint32_t kTagVacuum = 0x300;

template <typename real>
__device__ inline void gather_global(const Fields<real> &f, const Grid &g, int i, int j, int k, const real pos[3],
                                     real e[3], real b[3])
{
    Corners<real> c;
    load_corners(f, i * g.sx + j * g.sy + k, g.sx, g.sy, c);
    interp_corners(c, pos[0], pos[1], pos[2], e, b);
}

// A slot of t's buffer for each active lane of the wavefront: the lowest
// active lane adds the number of active lanes to the counter once, and
// every lane takes the slot after the active lanes below it (prefix count of
// the wavefront's ballot).  Called by exactly the lanes that append.
template <typename real>
__device__ inline int wave_slot(const AppendTarget<real> &t)
{
    const unsigned long long active = __ballot(1);
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)active) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(t.count, __popcll(active));
    base = __shfl(base, leader, 64);
    return base + __popcll(active & ((1ull << lane) - 1ull));
}

// Append one particle (cell (i, j, k), offsets o) to t; returns false when
// the buffer is full (the particle is dropped and the overflow flag set)
template <typename real>
__device__ inline bool append(const AppendTarget<real> &t, int i, int j, int k, const real o[3], const real u[3],
                              real q, real chi)
{
    const int slot = wave_slot(t);
    if (slot >= t.cap) {
        atomicOr(t.overflow, 1);
        return false;
    }
    t.s.cell[slot] = pack_cell(i, j, k);
    t.s.ox[slot] = o[0];
    t.s.oy[slot] = o[1];
    t.s.oz[slot] = o[2];
    t.s.ux[slot] = u[0];
    t.s.uy[slot] = u[1];
    t.s.uz[slot] = u[2];
    t.s.q[slot] = q;
    if (t.s.chi) t.s.chi[slot] = chi;
    if (t.s.id) t.s.id[slot] = -1;
    return true;
}

// deposit the path d from pos in cell (i, j, k) with global atomics, move
// pos to its end and re-bin; returns whether the end cell is_inside
template <typename real>
__device__ inline bool deposit_move(const Grid &g, const Current<real> &cur, int &i, int &j, int &k, real pos[3],
                                    const real d[3], real q,
                                    double *nonfinite)
{
    if (!(fabs(d[0]) < real(1) && fabs(d[1]) < real(1) && fabs(d[2]) < real(1))) {
        count_nonfinite(nonfinite);
        return false;
    }
    const int base = i * g.sx + j * g.sy + k;
    GlobalAcc<real> acc{cur.jx, cur.jy, cur.jz, g.sx, g.sy, base, base};
    deposit_path<real>(acc, q, pos, d);
    rebin(i, j, k, pos);
    return is_inside(g, i, j, k);
}

template <typename real, int PUSHER>
__global__ void __launch_bounds__(kBlock) k_push_qed_lepton(Grid g, Fields<real> f, Current<real> cur,
                                                            Species<real> s, int64_t first, int64_t n,
                                                            StepConst<real> sc, real e_s, QedArgs qa, uint32_t tag,
                                                            AppendTarget<real> ph, double *diag)
{
    const int64_t p = first + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    DiagAcc da, dph;
    int cell = p < first + n ? s.cell[p] : -1;
    if (cell >= 0 && !is_inside(g, cell_i(cell), cell_j(cell), cell_k(cell))) {
        s.cell[p] = -1;   // p_boundary (src/plist.cpp:73-102)
        cell = -1;
    }
    if (cell >= 0) {
        int i = cell_i(cell), j = cell_j(cell), k = cell_k(cell);
        real pos[3] = {s.ox[p], s.oy[p], s.oz[p]};
        real u[3] = {s.ux[p], s.uy[p], s.uz[p]};
        const real q = s.q[p];
        const real chi_prev = s.chi[p];
        real e[3], b[3];
        gather_global(f, g, i, j, k, pos, e, b);
        b[2] += sc.ext_bz;
        const real u_prev[3] = {u[0], u[1], u[2]};
        const real g_prev = qk_sqrt(real(1) + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        real gam = g_prev;
        push<PUSHER>(sc.cmr_dt, e, b, u, gam);   // p_n -> p_{n+1}
        const real ui[3] = {real(0.5) * (u[0] + u_prev[0]), real(0.5) * (u[1] + u_prev[1]),
                            real(0.5) * (u[2] + u_prev[2])};
        real chi = qed_chi(real(1) / e_s, e, b, ui, real(0.5) * (gam + g_prev));   // chi_{n+1/2}
        const QedDraw dr = qed_draw(qa.seed, qa.step, (uint32_t)p, tag);
        real r = dr.r;
        const real w = qed_rate_w(e_s, g_prev, r, real(0.5) * (chi + chi_prev));
        const bool emit = dr.acc < (double)(w * g_prev * sc.dt);
        bool alive;
        if (!emit) {
            const real a = sc.dt * qk_rcp(gam);
            const real d[3] = {u[0] * a * sc.inv_d[0], u[1] * a * sc.inv_d[1], u[2] * a * sc.inv_d[2]};
            alive = deposit_move(g, cur, i, j, k, pos, d, q, sc.nonfinite);
        } else {
            // photon emission (src/advance.cpp:291-357)
            const real a = real(0.5) * sc.dt / g_prev;
            const real up = sqrt(u_prev[0] * u_prev[0] + u_prev[1] * u_prev[1] + u_prev[2] * u_prev[2]);
            const real g_ph = r * g_prev;
            const real u_ph[3] = {g_ph * u_prev[0] / up, g_ph * u_prev[1] / up, g_ph * u_prev[2] / up};
            const real a3 = real(3) * a * g_prev / g_ph;
            real pph[3];
            for (int m = 0; m < 3; ++m) pph[m] = pos[m] - a * u_prev[m] * sc.inv_d[m] + a3 * u_ph[m] * sc.inv_d[m];
            int ip = i, jp = j, kp = k;
            rebin(ip, jp, kp, pph);
            const real qph = fabs(q);
            if (is_inside(g, ip, jp, kp) && append(ph, ip, jp, kp, pph, u_ph, qph, r * chi)) {
                dph.q += (double)qph;
                dph.n += 1;
                dph.e += (double)g_ph * (double)qph;
            }
            // recoil: the lepton keeps (1 - r) of its energy along u_prev
            chi = (real(1) - r) * chi;
            gam = (real(1) - r) * g_prev;
            const real sc_u = sqrt((gam * gam - real(1)) / (up * up));
            for (int m = 0; m < 3; ++m) u[m] = u_prev[m] * sc_u;
            const real a2 = a * g_prev / gam;
            for (int m = 0; m < 3; ++m) pos[m] += (-a * u_prev[m] + a2 * u[m]) * sc.inv_d[m];
            rebin(i, j, k, pos);
            alive = is_inside(g, i, j, k);
            if (alive) {
                gather_global(f, g, i, j, k, pos, e, b);
                b[2] += sc.ext_bz;
                push<PUSHER>(sc.cmr_dt, e, b, u, gam);
                const real a4 = sc.dt * qk_rcp(gam);
                const real d[3] = {u[0] * a4 * sc.inv_d[0], u[1] * a4 * sc.inv_d[1], u[2] * a4 * sc.inv_d[2]};
                alive = deposit_move(g, cur, i, j, k, pos, d, q, sc.nonfinite);
            }
        }
        if (alive) {
            s.cell[p] = pack_cell(i, j, k);
            s.ox[p] = pos[0];
            s.oy[p] = pos[1];
            s.oz[p] = pos[2];
            s.ux[p] = u[0];
            s.uy[p] = u[1];
            s.uz[p] = u[2];
            s.chi[p] = chi;
            da.q += (double)q;
            da.n += 1;
            da.e += kinetic(u, gam) * (double)q;
        } else {
            s.cell[p] = -1;
        }
    }
    block_accumulate(diag, da.q, da.e, da.n, 0.0);
    block_accumulate(ph.diag, dph.q, dph.e, dph.n, 0.0);
}

// One lepton of a Breit-Wheeler pair (bear_particle, src/mc.cpp:118-139,
// then push, chi, deposit, place as src/advance.cpp:190-240 do)
template <typename real, int PUSHER>
__device__ inline void pair_lepton(const Grid &g, const Current<real> &cur, const StepConst<real> &sc, real cmr_dt,
                                   real e_s, const real e[3], const real b[3], int i, int j, int k, const real pos0[3],
                                   const real dir[3], real dir2, real gb, real qb, const AppendTarget<real> &t,
                                   DiagAcc &da)
{
    // the chi given at birth is replaced by chi at n+1/2 after the push
    const real sc_u = sqrt((gb * gb - real(1)) / dir2);
    real u[3] = {sc_u * dir[0], sc_u * dir[1], sc_u * dir[2]};
    const real u_prev[3] = {u[0], u[1], u[2]};
    real gam = gb;
    push<PUSHER>(cmr_dt, e, b, u, gam);
    const real ui[3] = {real(0.5) * (u[0] + u_prev[0]), real(0.5) * (u[1] + u_prev[1]),
                        real(0.5) * (u[2] + u_prev[2])};
    const real chi = qed_chi(real(1) / e_s, e, b, ui, real(0.5) * (gam + gb));
    real pos[3] = {pos0[0], pos0[1], pos0[2]};
    const real a = sc.dt * qk_rcp(gam);
    const real d[3] = {u[0] * a * sc.inv_d[0], u[1] * a * sc.inv_d[1], u[2] * a * sc.inv_d[2]};
    if (deposit_move(g, cur, i, j, k, pos, d, qb, sc.nonfinite) && append(t, i, j, k, pos, u, qb, chi)) {
        da.q += (double)qb;
        da.n += 1;
        da.e += kinetic(u, gam) * (double)qb;
    }
}

template <typename real, int PUSHER>
__global__ void __launch_bounds__(kBlock) k_photon_qed(Grid g, Fields<real> f, Current<real> cur, Species<real> s,
                                                       int64_t first, int64_t n, StepConst<real> sc, real e_s,
                                                       QedArgs qa, AppendTarget<real> el, AppendTarget<real> po,
                                                       double *diag)
{
    const int64_t p = first + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    DiagAcc da, de, dp;
    int cell = p < first + n ? s.cell[p] : -1;
    if (cell >= 0 && !is_inside(g, cell_i(cell), cell_j(cell), cell_k(cell))) {
        s.cell[p] = -1;   // p_boundary (src/plist.cpp:73-102)
        cell = -1;
    }
    if (cell >= 0) {
        int i = cell_i(cell), j = cell_j(cell), k = cell_k(cell);
        real pos[3] = {s.ox[p], s.oy[p], s.oz[p]};
        const real u[3] = {s.ux[p], s.uy[p], s.uz[p]};
        const real q = s.q[p];
        const real chi_prev = s.chi[p];
        real e[3], b[3];
        gather_global(f, g, i, j, k, pos, e, b);
        b[2] += sc.ext_bz;
        const real u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const real gam = sqrt(u2);
        const real chi = qed_chi(real(1) / e_s, e, b, u, gam);   // chi_{n+1/2}
        const QedDraw dr = qed_draw(qa.seed, qa.step, (uint32_t)p, kTagPair);
        const real r = dr.r;
        const bool pair =
            dr.acc < (double)(qed_rate_tilde_w(e_s, gam, r, real(0.5) * (chi + chi_prev)) * gam * sc.dt);
        if (!pair) {
            // straight flight (src/advance.cpp:173-183)
            const real a = sc.dt / gam;
            for (int m = 0; m < 3; ++m) pos[m] += u[m] * a * sc.inv_d[m];
            const bool finite = fabs(u[0] * a) < real(1e30) && fabs(u[1] * a) < real(1e30) && fabs(u[2] * a) < real(1e30);
            if (!finite) count_nonfinite(sc.nonfinite);
            rebin(i, j, k, pos);
            if (finite && is_inside(g, i, j, k)) {
                s.cell[p] = pack_cell(i, j, k);
                s.ox[p] = pos[0];
                s.oy[p] = pos[1];
                s.oz[p] = pos[2];
                s.chi[p] = chi;
                da.q += (double)q;
                da.n += 1;
                da.e += (double)gam * (double)q;
            } else {
                s.cell[p] = -1;
            }
        } else {
            // pair production at the photon's position; the photon is destroyed
            pair_lepton<real, PUSHER>(g, cur, sc, -sc.dt, e_s, e, b, i, j, k, pos, u, u2, r * gam, -q, el, de);
            pair_lepton<real, PUSHER>(g, cur, sc, sc.dt, e_s, e, b, i, j, k, pos, u, u2, (real(1) - r) * gam, q, po,
                                      dp);
            s.cell[p] = -1;
        }
    }
    block_accumulate(diag, da.q, da.e, da.n, 0.0);
    block_accumulate(el.diag, de.q, de.e, de.n, 0.0);
    block_accumulate(po.diag, dp.q, dp.e, dp.n, 0.0);
}

template <typename real, int PUSHER>
__global__ void __launch_bounds__(kBlock) k_vacuum_pairs(Grid g, Fields<real> f, double q_vac, QedArgs qa,
                                                         AppendTarget<real> el, AppendTarget<real> po, int64_t ncells)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= ncells) return;
    const int mk = g.nz - 4, mj = g.ny - 4;
    const int k = 2 + (int)(t % mk);
    const int j = 2 + (int)((t / mk) % mj);
    const int i = 2 + (int)(t / ((int64_t)mk * mj));
    const real half[3] = {real(0.5), real(0.5), real(0.5)};
    real e[3], b[3];
    gather_global(f, g, i, j, k, half, e, b);
    const double ed[3] = {e[0], e[1], e[2]}, bd[3] = {b[0], b[1], b[2]};
    const double rate = qed_vacuum_rate(qa.e_s, ed, bd) * g.dt * g.dx * g.dy * g.dz;
    const uint32_t cell = (uint32_t)((i * g.ny + j) * g.nz + k);
    if (!(qed_draw(qa.seed, qa.step, cell, kTagVacuum).acc < rate)) return;
    // both born at rest (g = 1), pushed back by dt / 2 (src/mc.cpp:170-176)
    for (int sp = 0; sp < 2; ++sp) {
        const real cmr = sp == 0 ? real(-1) : real(1);
        real u[3] = {0, 0, 0};
        real gam = real(1);
        push<PUSHER>(cmr * real(-0.5 * g.dt), e, b, u, gam);
        append(sp == 0 ? el : po, i, j, k, half, u, (real)(sp == 0 ? -q_vac : q_vac), real(0));
    }
}

}  // namespace

template <typename real>
hipError_t launch_push_qed_lepton(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                                  const Species<real> &s, const PushArgs &a, const QedArgs &q,
                                  const AppendTarget<real> &ph)
{
    if (a.n <= 0) return hipSuccess;
    const StepConst<real> sc = step_const<real>(g, a);
    const uint32_t tag = kTagEmit + (a.cmr < 0 ? 0u : 1u);
    if (a.pusher == 1)
        k_push_qed_lepton<real, 1><<<blocks_for(a.n), kBlock, 0, st>>>(g, f, j, s, a.first, a.n, sc, (real)q.e_s, q,
                                                                      tag, ph, a.diag);
    else
        k_push_qed_lepton<real, 0><<<blocks_for(a.n), kBlock, 0, st>>>(g, f, j, s, a.first, a.n, sc, (real)q.e_s, q,
                                                                      tag, ph, a.diag);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_photon_qed(hipStream_t st, const Grid &g, const Fields<real> &f, const Current<real> &j,
                             const Species<real> &s, const PushArgs &a, const QedArgs &q,
                             const AppendTarget<real> &el, const AppendTarget<real> &po)
{
    if (a.n <= 0) return hipSuccess;
    const StepConst<real> sc = step_const<real>(g, a);
    if (a.pusher == 1)
        k_photon_qed<real, 1><<<blocks_for(a.n), kBlock, 0, st>>>(g, f, j, s, a.first, a.n, sc, (real)q.e_s, q, el,
                                                                 po, a.diag);
    else
        k_photon_qed<real, 0><<<blocks_for(a.n), kBlock, 0, st>>>(g, f, j, s, a.first, a.n, sc, (real)q.e_s, q, el,
                                                                 po, a.diag);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_vacuum_pairs(hipStream_t st, const Grid &g, const Fields<real> &f, int pusher, double q_vac,
                               const QedArgs &q, const AppendTarget<real> &el, const AppendTarget<real> &po)
{
    const int64_t ncells = (int64_t)(g.nx - 4) * (g.ny - 4) * (g.nz - 4);
    if (ncells <= 0) return hipSuccess;
    if (pusher == 1)
        k_vacuum_pairs<real, 1><<<blocks_for(ncells), kBlock, 0, st>>>(g, f, q_vac, q, el, po, ncells);
    else
        k_vacuum_pairs<real, 0><<<blocks_for(ncells), kBlock, 0, st>>>(g, f, q_vac, q, el, po, ncells);
    return hipGetLastError();
}

#define QK_INST(real)                                                                                          \
    template hipError_t launch_push_qed_lepton<real>(hipStream_t, const Grid &, const Fields<real> &,          \
                                                     const Current<real> &, const Species<real> &,             \
                                                     const PushArgs &, const QedArgs &, const AppendTarget<real> &); \
    template hipError_t launch_photon_qed<real>(hipStream_t, const Grid &, const Fields<real> &,               \
                                                const Current<real> &, const Species<real> &, const PushArgs &, \
                                                const QedArgs &, const AppendTarget<real> &,                   \
                                                const AppendTarget<real> &);                                   \
    template hipError_t launch_vacuum_pairs<real>(hipStream_t, const Grid &, const Fields<real> &, int, double, \
                                                  const QedArgs &, const AppendTarget<real> &,                 \
                                                  const AppendTarget<real> &);
QK_INST(float)
QK_INST(double)
#undef QK_INST

}  // namespace qk
