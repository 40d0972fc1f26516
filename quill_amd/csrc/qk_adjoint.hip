// qk_adjoint.hip -- adjoint of the deterministic step (QED off): the
// transposes of fadvance (the three solvers' B and E updates and
// interpolate_be, src/maxwell.cpp, src/advance.cpp:13-83) and the per-particle
// adjoint of padvance (gather, push, displacement, charge-conserving deposit,
// src/advance.cpp:90-387).
//
// The field adjoints are kept in FP64 and scattered with FP64 atomics.  The
// Maxwell transposes enumerate, for every cell a forward update writes, the
// nodes it reads and their coefficients -- the forward operator's own
// update sets and stencils (one description serves all three solvers) --
// and add coefficient x adjoint to each node.
#include "qk_dual.h"
#include "qk_push.h"

namespace qk {

namespace {

constexpr int kABlock = 256;

inline unsigned ablocks(int64_t n) { return (unsigned)((n + kABlock - 1) / kABlock); }

struct AdjCoef {
    double a[3], b[3];     // solver coefficients per axis (FDTD: a = 0, b = 1)
    double dtd[3];         // dt / dx, dt / dy, dt / dz
    double d[3];           // dx, dy, dz
    int solver;
};

__device__ inline bool in_range(int v, int lo, int hi) { return v >= lo && v <= hi; }

// the cells whose component comp of B the half step updates
// (src/maxwell.cpp:445-484 FDTD, 251-290 NDFX, 35-85 five-point)
__device__ inline bool b_updates(const Grid &g, int solver, int comp, const int c[3])
{
    const int n[3] = {g.nx, g.ny, g.nz};
    if (solver == 2) {
        int top = 0;
        for (int a = 0; a < 3; ++a) {
            if (c[a] == n[a] - 1) ++top;
            else if (!in_range(c[a], 1, n[a] - 3)) return false;
        }
        return top <= 1;
    }
    const int lo = solver == 0 ? 0 : 1;
    int nlow = 0;
    for (int a = 0; a < 3; ++a) {
        if (a == comp) {
            if (!in_range(c[a], lo + 1, n[a] - 2)) return false;
        } else {
            if (!in_range(c[a], lo, n[a] - 2)) return false;
            if (c[a] == lo) ++nlow;
        }
    }
    return nlow <= 1;
}

// ... of E (src/maxwell.cpp:486-522 FDTD, 292-328 NDFX, 87-125 five-point)
__device__ inline bool e_updates(const Grid &g, int solver, int comp, const int c[3])
{
    const int n[3] = {g.nx, g.ny, g.nz};
    const int lo = solver == 0 ? 0 : 1;
    const int top = solver == 2 ? 3 : 2;
    for (int a = 0; a < 3; ++a) {
        const int first = a == comp ? lo : lo + 1;
        if (!in_range(c[a], first, n[a] - top)) return false;
    }
    return true;
}

// The difference of component comp of a field along axis a at cell c, as
// (node, coefficient) pairs given to emit: fwd = forward difference (the B
// update reads E), else backward (the E update reads B).
//   FDTD        F[hi] - F[lo]
//   NDFX        b_c (F[hi] - F[lo]) + a_c (the same pair shifted by +-1
//               along the component's own axis)
//   five-point  b_a (F[hi] - F[lo]) + a_a (F[hi+1] - F[hi] + F[lo] - F[lo-1]),
//               nodes outside the grid read as zero
template <class Emit>
__device__ inline void diff_terms(const Grid &g, const AdjCoef &cf, int comp, int a, const int c[3], bool fwd,
                                  double w, Emit &&emit)
{
    int lo[3] = {c[0], c[1], c[2]};
    if (!fwd) lo[a] -= 1;
    int hi[3] = {lo[0], lo[1], lo[2]};
    hi[a] += 1;
    auto node = [&](const int p[3], int shift_axis, int shift, double coef) {
        int q[3] = {p[0], p[1], p[2]};
        if (shift) q[shift_axis] += shift;
        if ((unsigned)q[0] < (unsigned)g.nx && (unsigned)q[1] < (unsigned)g.ny && (unsigned)q[2] < (unsigned)g.nz)
            emit(q[0] * g.sx + q[1] * g.sy + q[2], coef * w);
    };
    if (cf.solver == 0) {
        node(hi, 0, 0, 1.0);
        node(lo, 0, 0, -1.0);
    } else if (cf.solver == 1) {
        const double bc = cf.b[comp], ac = cf.a[comp];
        node(hi, 0, 0, bc);
        node(lo, 0, 0, -bc);
        node(hi, comp, +1, ac);
        node(lo, comp, +1, -ac);
        node(hi, comp, -1, ac);
        node(lo, comp, -1, -ac);
    } else {
        const double ba = cf.b[a], aa = cf.a[a];
        node(hi, 0, 0, ba - aa);
        node(lo, 0, 0, -ba + aa);
        node(hi, a, +1, aa);
        node(lo, a, -1, -aa);
    }
}

template <typename L>
struct LamField {
    L *c[3];
};

// transpose of the B half step: lE[node] += coefficient x lB[cell] for
// every updated B cell (the identity part leaves lB as it is)
__global__ void __launch_bounds__(kABlock) k_adj_b(Grid g, AdjCoef cf, LamField<const double> lB, LamField<double> lE)
{
    const int64_t t = (int64_t)blockIdx.x * kABlock + threadIdx.x;
    if (t >= g.ncell) return;
    const int c[3] = {(int)(t / g.sx), (int)((t / g.sy) % g.ny), (int)(t % g.nz)};
    // B_x += 0.5 (-dtdy dE_z/dy + dtdz dE_y/dz), cyclically (src/maxwell.cpp:445-484)
    for (int comp = 0; comp < 3; ++comp) {
        if (!b_updates(g, cf.solver, comp, c)) continue;
        const double l = lB.c[comp][t];
        if (l == 0.0) continue;
        const int a1 = (comp + 1) % 3, a2 = (comp + 2) % 3;
        // -dtd[a1] d(E_a2)/d(a1) + dtd[a2] d(E_a1)/d(a2)
        double *e2 = lE.c[a2], *e1 = lE.c[a1];
        diff_terms(g, cf, a2, a1, c, true, -0.5 * cf.dtd[a1] * l,
                   [&](int64_t n, double v) { atomicAdd(e2 + n, v); });
        diff_terms(g, cf, a1, a2, c, true, 0.5 * cf.dtd[a2] * l,
                   [&](int64_t n, double v) { atomicAdd(e1 + n, v); });
    }
}

// transpose of the E step: lB[node] += coefficient x lE[cell] for every
// updated E cell, and lJ = -d_comp lE there (0 elsewhere)
__global__ void __launch_bounds__(kABlock) k_adj_e(Grid g, AdjCoef cf, LamField<const double> lE, LamField<double> lB,
                                                   LamField<double> lJ)
{
    const int64_t t = (int64_t)blockIdx.x * kABlock + threadIdx.x;
    if (t >= g.ncell) return;
    const int c[3] = {(int)(t / g.sx), (int)((t / g.sy) % g.ny), (int)(t % g.nz)};
    // E_x += dtdy dB_z/dy - dtdz dB_y/dz - J_x dx, cyclically (src/maxwell.cpp:486-522)
    for (int comp = 0; comp < 3; ++comp) {
        if (!e_updates(g, cf.solver, comp, c)) {
            lJ.c[comp][t] = 0.0;
            continue;
        }
        const double l = lE.c[comp][t];
        lJ.c[comp][t] = -cf.d[comp] * l;
        if (l == 0.0) continue;
        const int a1 = (comp + 1) % 3, a2 = (comp + 2) % 3;
        double *b2 = lB.c[a2], *b1 = lB.c[a1];
        diff_terms(g, cf, a2, a1, c, false, cf.dtd[a1] * l, [&](int64_t n, double v) { atomicAdd(b2 + n, v); });
        diff_terms(g, cf, a1, a2, c, false, -cf.dtd[a2] * l, [&](int64_t n, double v) { atomicAdd(b1 + n, v); });
    }
}

// transpose of interpolate_be (src/advance.cpp:20-83): lB[node] += w lBe[cell]
// over the 8-point (4 on one boundary plane) mean, then lBe = 0.  Cells on
// two or more boundary planes keep Be in the forward step; no particle
// gathers there, so their adjoint is zero.
__global__ void __launch_bounds__(kABlock) k_adj_interp_be(Grid g, LamField<double> lBe, LamField<double> lB)
{
    const int64_t t = (int64_t)blockIdx.x * kABlock + threadIdx.x;
    if (t >= g.ncell) return;
    const int c[3] = {(int)(t / g.sx), (int)((t / g.sy) % g.ny), (int)(t % g.nz)};
    const int n[3] = {g.nx, g.ny, g.nz};
    const int nb = (c[0] == 0 || c[0] == n[0] - 1) + (c[1] == 0 || c[1] == n[1] - 1) + (c[2] == 0 || c[2] == n[2] - 1);
    const int off[3][3] = {{+1, -1, -1}, {-1, +1, -1}, {-1, -1, +1}};
    for (int comp = 0; comp < 3; ++comp) {
        const double l = lBe.c[comp][t];
        lBe.c[comp][t] = 0.0;
        if (nb >= 2 || l == 0.0) continue;
        int lo[3], hi[3], count = 1;
        for (int a = 0; a < 3; ++a) {
            const int o = c[a] + off[comp][a];
            if (o < 0 || o > n[a] - 1) {
                lo[a] = hi[a] = c[a];
            } else {
                lo[a] = o < c[a] ? o : c[a];
                hi[a] = o < c[a] ? c[a] : o;
                count *= 2;
            }
        }
        const double w = (count == 8 ? 0.125 : 0.25) * l;
        for (int i = lo[0]; i <= hi[0]; ++i)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) atomicAdd(lB.c[comp] + (int64_t)i * g.sx + j * g.sy + k, w);
    }
}

// The deposit as seen by the adjoint: each current term, a Dual, is
// contracted with the current's adjoint at its node.
template <typename D>
struct ContractAcc {
    const double *lj[3];
    int sx, sy;
    int base, cur;
    D s;
    __device__ inline void set_cell(int wi, int wj, int wk) { cur = base + wi * sx + wj * sy + wk; }
    __device__ inline void operator()(int comp, int di, int dj, int dk, const D &v)
    {
        s += v * D(lj[comp][cur + di * sx + dj * sy + dk]);
    }
};

// The scalar S of one particle's step: the adjoints of its outputs (the next
// position, momentum and charge) and of the current, contracted with those
// outputs.  Its gradient with respect to the step's inputs is the particle's
// adjoint; D seeds the inputs.
template <typename D, int PUSHER, typename real>
__device__ inline D step_scalar(const Grid &g, const StepConst<real> &sc, D e[3], D b[3], D u[3], D o[3], D q,
                                const double lx[3], const double lu[3], double lq, const double *const lj[3], int base)
{
    b[2] += D(sc.ext_bz);
    D gam = qk_sqrt(D(1) + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    push<PUSHER>(D(sc.cmr_dt), e, b, u, gam);
    const D a = D(sc.dt) * qk_rcp(gam);
    const D d[3] = {u[0] * a * D(sc.inv_d[0]), u[1] * a * D(sc.inv_d[1]), u[2] * a * D(sc.inv_d[2])};
    ContractAcc<D> acc{{lj[0], lj[1], lj[2]}, g.sx, g.sy, base, base, D(0)};
    D pos[3] = {o[0], o[1], o[2]};
    deposit_path<D>(acc, q, pos, d);
    D s = acc.s + D(lq) * q;
#pragma unroll
    for (int m = 0; m < 3; ++m) s += D(lx[m]) * pos[m] + D(lu[m]) * u[m];
    return s;
}

template <typename real>
struct AdjParticles {
    double *lx[3], *lu[3], *lq;
};

// Adjoint of one particle's push at step n.  The record (before the push) and
// the fields E, Be of step n come from the recorded state.  On entry lx, lu,
// lq hold the adjoints of the particle after the step; on return those of
// the particle before it.  lE and lBe receive the gather's node adjoints.
template <typename real, int PUSHER>
__global__ void __launch_bounds__(kABlock) k_adj_particles(Grid g, Fields<real> f, Species<real> s, int64_t n,
                                                          StepConst<real> sc, LamField<const double> lJ,
                                                          LamField<double> lE, LamField<double> lBe,
                                                          AdjParticles<real> lp)
{
    const int64_t p = (int64_t)blockIdx.x * kABlock + threadIdx.x;
    if (p >= n) return;
    const int c = s.cell[p];
    if (c < 0) return;   // not alive at step n: its adjoint stays as it is (zero)
    if (!is_inside(g, cell_i(c), cell_j(c), cell_k(c))) {
        // erased by p_boundary before the push: it influences nothing after
        lp.lx[0][p] = lp.lx[1][p] = lp.lx[2][p] = 0;
        lp.lu[0][p] = lp.lu[1][p] = lp.lu[2][p] = 0;
        lp.lq[p] = 0;
        return;
    }
    const int base = cell_base(g, c);
    Corners<real> nodes;
    load_corners(f, base, g.sx, g.sy, nodes);
    const real o0[3] = {s.ox[p], s.oy[p], s.oz[p]}, u0[3] = {s.ux[p], s.uy[p], s.uz[p]};
    const real q0 = s.q[p];
    const double lx[3] = {lp.lx[0][p], lp.lx[1][p], lp.lx[2][p]};
    const double lu[3] = {lp.lu[0][p], lp.lu[1][p], lp.lu[2][p]};
    const double lq = lp.lq[p];
    const double *const lj[3] = {lJ.c[0], lJ.c[1], lJ.c[2]};

    // pass 1: gradient with respect to the offsets, momentum and charge
    // (through the gather, the push and the deposit)
    {
        using D = Dual<7, real>;
        Corners<D> nd;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                nd.e[a][m] = D((double)nodes.e[a][m]);
                nd.b[a][m] = D((double)nodes.b[a][m]);
            }
        D o[3], u[3], e[3], b[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            o[m] = D::var(o0[m], m);
            u[m] = D::var(u0[m], 3 + m);
        }
        const D q = D::var(q0, 6);
        interp_corners(nd, o[0], o[1], o[2], e, b);
        const D sum = step_scalar<D, PUSHER>(g, sc, e, b, u, o, q, lx, lu, lq, lj, base);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            lp.lx[m][p] = (double)sum.d[m];
            lp.lu[m][p] = (double)sum.d[3 + m];
        }
        lp.lq[p] = (double)sum.d[6];
    }
    // pass 2: gradient with respect to the gathered E and B, spread to the
    // nodes with the gather's weights
    {
        using D = Dual<6, real>;
        real ev[3], bv[3];
        interp_corners(nodes, o0[0], o0[1], o0[2], ev, bv);
        D e[3], b[3], u[3], o[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            e[m] = D::var(ev[m], m);
            b[m] = D::var(bv[m], 3 + m);
            u[m] = D((double)u0[m]);
            o[m] = D((double)o0[m]);
        }
        const D sum = step_scalar<D, PUSHER>(g, sc, e, b, u, o, D((double)q0), lx, lu, lq, lj, base);
        const real xa = o0[0], ya = o0[1], za = o0[2];
        const real xb = real(1) - xa, yb = real(1) - ya, zb = real(1) - za;
        const real w[3][4] = {{yb * zb, ya * zb, yb * za, ya * za},
                              {xb * zb, xa * zb, xb * za, xa * za},
                              {xb * yb, xa * yb, xb * ya, xa * ya}};
        const int ox[4] = {0, g.sy, 1, g.sy + 1};
        const int oy[4] = {0, g.sx, 1, g.sx + 1};
        const int oz[4] = {0, g.sx, g.sy, g.sx + g.sy};
        const int *offs[3] = {ox, oy, oz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double ge = (double)sum.d[a], gb = (double)sum.d[3 + a];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int64_t node = base + offs[a][m];
                if (ge != 0.0) atomicAdd(lE.c[a] + node, (double)w[a][m] * ge);
                if (gb != 0.0) atomicAdd(lBe.c[a] + node, (double)w[a][m] * gb);
            }
        }
    }
}

// ---- tangent-linear step (forward mode) ----------------------------------

// the deposit of a dual current term: its derivative part is the tangent of
// the current, added to the tangent current (FP64)
template <typename D>
struct TangentAcc {
    double *dj[3];
    int sx, sy;
    int base, cur;
    __device__ inline void set_cell(int wi, int wj, int wk) { cur = base + wi * sx + wj * sy + wk; }
    __device__ inline void operator()(int comp, int di, int dj_, int dk, const D &v)
    {
        const double t = (double)v.d[0];
        if (t != 0.0) atomicAdd(dj[comp] + cur + di * sx + dj_ * sy + dk, t);
    }
};

template <typename real>
struct TangentParticles {
    double *dx[3], *du[3], *dq;
};

// a dual number with value v and tangent t
template <typename D, typename real>
__device__ inline D dual_of(real v, double t)
{
    D r((double)v);
    r.v = v;
    r.d[0] = (real)t;
    return r;
}

// Tangent of one particle's push at step n, from the state before it: the
// record s with its tangent tp (offsets, momentum, charge), the fields E, Be
// of step n with their tangents df.  The same templated forward code (gather,
// pusher, displacement, charge-conserving deposit) runs on Dual<1> numbers;
// branches follow the values, as in the forward step.  tp becomes the
// tangent after the step; dJ receives the tangent of the particle's current.
template <typename real, int PUSHER>
__global__ void __launch_bounds__(kABlock) k_tlm_particles(Grid g, Fields<real> f, Fields<double> df, Species<real> s,
                                                          int64_t n, StepConst<real> sc, LamField<double> dJ,
                                                          TangentParticles<real> tp)
{
    const int64_t p = (int64_t)blockIdx.x * kABlock + threadIdx.x;
    if (p >= n) return;
    const int c = s.cell[p];
    if (c < 0) return;
    if (!is_inside(g, cell_i(c), cell_j(c), cell_k(c))) {
        // erased by p_boundary before the push
        tp.dx[0][p] = tp.dx[1][p] = tp.dx[2][p] = 0;
        tp.du[0][p] = tp.du[1][p] = tp.du[2][p] = 0;
        tp.dq[p] = 0;
        return;
    }
    using D = Dual<1, real>;
    const int base = cell_base(g, c);
    Corners<real> nodes;
    Corners<double> dn;
    load_corners(f, base, g.sx, g.sy, nodes);
    load_corners(df, base, g.sx, g.sy, dn);
    Corners<D> nd;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            nd.e[a][m] = dual_of<D>(nodes.e[a][m], dn.e[a][m]);
            nd.b[a][m] = dual_of<D>(nodes.b[a][m], dn.b[a][m]);
        }
    D o[3] = {dual_of<D>(s.ox[p], tp.dx[0][p]), dual_of<D>(s.oy[p], tp.dx[1][p]), dual_of<D>(s.oz[p], tp.dx[2][p])};
    D u[3] = {dual_of<D>(s.ux[p], tp.du[0][p]), dual_of<D>(s.uy[p], tp.du[1][p]), dual_of<D>(s.uz[p], tp.du[2][p])};
    const D q = dual_of<D>(s.q[p], tp.dq[p]);
    D e[3], b[3];
    interp_corners(nd, o[0], o[1], o[2], e, b);
    b[2] += D(sc.ext_bz);
    D gam = qk_sqrt(D(1) + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    push<PUSHER>(D(sc.cmr_dt), e, b, u, gam);
    const D a = D(sc.dt) * qk_rcp(gam);
    const D d[3] = {u[0] * a * D(sc.inv_d[0]), u[1] * a * D(sc.inv_d[1]), u[2] * a * D(sc.inv_d[2])};
    TangentAcc<D> acc{{dJ.c[0], dJ.c[1], dJ.c[2]}, g.sx, g.sy, base, base};
    D pos[3] = {o[0], o[1], o[2]};
    deposit_path<D>(acc, q, pos, d);
    // the offset after re-binning differs from pos by a whole number of cells
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        tp.dx[m][p] = (double)pos[m].d[0];
        tp.du[m][p] = (double)u[m].d[0];
    }
}

}  // namespace

template <typename real>
hipError_t launch_tlm_particles(hipStream_t st, const Grid &g, const Fields<real> &f, const Fields<double> &df,
                                const Species<real> &s, int64_t n, const PushArgs &a, double *const dJ[3],
                                double *const tpart[7])
{
    if (n <= 0) return hipSuccess;
    const StepConst<real> sc = step_const<real>(g, a);
    LamField<double> j{{dJ[0], dJ[1], dJ[2]}};
    TangentParticles<real> tp{{tpart[0], tpart[1], tpart[2]}, {tpart[3], tpart[4], tpart[5]}, tpart[6]};
    if (a.pusher == 1)
        k_tlm_particles<real, 1><<<ablocks(n), kABlock, 0, st>>>(g, f, df, s, n, sc, j, tp);
    else
        k_tlm_particles<real, 0><<<ablocks(n), kABlock, 0, st>>>(g, f, df, s, n, sc, j, tp);
    return hipGetLastError();
}

template hipError_t launch_tlm_particles<float>(hipStream_t, const Grid &, const Fields<float> &,
                                                const Fields<double> &, const Species<float> &, int64_t,
                                                const PushArgs &, double *const[3], double *const[7]);
template hipError_t launch_tlm_particles<double>(hipStream_t, const Grid &, const Fields<double> &,
                                                 const Fields<double> &, const Species<double> &, int64_t,
                                                 const PushArgs &, double *const[3], double *const[7]);

hipError_t launch_adj_fields(hipStream_t st, int solver, const Grid &g, const SolverCoef &sc, double *const lam[12])
{
    // lam: lE[3], lB[3], lBe[3], lJ[3]
    AdjCoef cf;
    for (int a = 0; a < 3; ++a) {
        cf.a[a] = sc.a[a];
        cf.b[a] = sc.b[a];
    }
    cf.d[0] = g.dx;
    cf.d[1] = g.dy;
    cf.d[2] = g.dz;
    for (int a = 0; a < 3; ++a) cf.dtd[a] = g.dt / cf.d[a];
    cf.solver = solver;
    LamField<double> lE{{lam[0], lam[1], lam[2]}}, lB{{lam[3], lam[4], lam[5]}}, lBe{{lam[6], lam[7], lam[8]}},
        lJ{{lam[9], lam[10], lam[11]}};
    LamField<const double> cE{{lam[0], lam[1], lam[2]}}, cB{{lam[3], lam[4], lam[5]}};
    const unsigned nb = ablocks(g.ncell);
    // fadvance is B half, E, B half, interpolate_be: the transposes run in reverse
    k_adj_interp_be<<<nb, kABlock, 0, st>>>(g, lBe, lB);
    k_adj_b<<<nb, kABlock, 0, st>>>(g, cf, cB, lE);
    k_adj_e<<<nb, kABlock, 0, st>>>(g, cf, cE, lB, lJ);
    k_adj_b<<<nb, kABlock, 0, st>>>(g, cf, cB, lE);
    return hipGetLastError();
}

hipError_t launch_adj_interp_be(hipStream_t st, const Grid &g, double *const lam[12])
{
    LamField<double> lB{{lam[3], lam[4], lam[5]}}, lBe{{lam[6], lam[7], lam[8]}};
    k_adj_interp_be<<<ablocks(g.ncell), kABlock, 0, st>>>(g, lBe, lB);
    return hipGetLastError();
}

template <typename real>
hipError_t launch_adj_particles(hipStream_t st, const Grid &g, const Fields<real> &f, const Species<real> &s,
                                int64_t n, const PushArgs &a, double *const lam[12], double *const lpart[7])
{
    if (n <= 0) return hipSuccess;
    const StepConst<real> sc = step_const<real>(g, a);
    LamField<const double> lJ{{lam[9], lam[10], lam[11]}};
    LamField<double> lE{{lam[0], lam[1], lam[2]}}, lBe{{lam[6], lam[7], lam[8]}};
    AdjParticles<real> lp{{lpart[0], lpart[1], lpart[2]}, {lpart[3], lpart[4], lpart[5]}, lpart[6]};
    if (a.pusher == 1)
        k_adj_particles<real, 1><<<ablocks(n), kABlock, 0, st>>>(g, f, s, n, sc, lJ, lE, lBe, lp);
    else
        k_adj_particles<real, 0><<<ablocks(n), kABlock, 0, st>>>(g, f, s, n, sc, lJ, lE, lBe, lp);
    return hipGetLastError();
}

template hipError_t launch_adj_particles<float>(hipStream_t, const Grid &, const Fields<float> &,
                                                const Species<float> &, int64_t, const PushArgs &, double *const[12],
                                                double *const[7]);
template hipError_t launch_adj_particles<double>(hipStream_t, const Grid &, const Fields<double> &,
                                                 const Species<double> &, int64_t, const PushArgs &,
                                                 double *const[12], double *const[7]);

}  // namespace qk
