"""The deterministic step (QED off) as a torch.autograd.Function.

forward runs `nsteps` steps of one subdomain on the GPU (libquillk.so) from
initial fields and particles and returns the final ones; backward runs the
library's hand-written adjoint (qk_tape_begin / qk_adjoint): the
transposes of the three Maxwell solvers and of interpolate_be, and a
per-particle adjoint of gather, push and charge-conserving deposit, swept
backwards over a trajectory kept as snapshots every `every` steps and
recomputed one segment at a time.  The differentiated loop body is
src/main.cpp:2319-2332 (padvance, src/advance.cpp:90-387, and fadvance,
src/advance.cpp:13-83) with QED and merging off; with the tiled deposit
(the default here) the particles are sorted into tiles once, at
qk_tape_begin, and keep their slots after that.

Tensors cross the boundary on the host in FP64, as the C ABI takes them;
the arithmetic runs on the GPU in the context's precision.
"""
from dataclasses import dataclass, field

import numpy as np
import torch

from . import ION0, Region


@dataclass
class RunSpec:
    """What one differentiable run does.

    region: keyword arguments of Region (grid, steps, solver, pusher,
    precision, ion_cmr, deposit, ...); sort_every, sort_fallback and
    track_ids are set here, deposit defaults to "tiled".
    species: the species slots whose particles are passed, in order.
    """
    region: dict
    nsteps: int
    every: int = 25
    species: tuple = (0,)
    alive: dict = field(default_factory=dict)   # filled by forward: species -> bool mask of survivors


def simulate(spec, E0, B0, parts, tape=False):
    """Run spec.nsteps steps from (E0, B0, parts) (tensors as PICRun takes
    them).  Returns (region, E, B, [particles]) with FP64 tensors; the region
    is open, recorded for qk_adjoint when `tape`, and closed otherwise."""
    kw = dict(spec.region)
    kw.update(sort_every=0, sort_fallback=0.0, track_ids=True)
    kw.setdefault("deposit", "tiled")
    r = Region(**kw)
    try:
        r.set_fields(E0.detach().cpu().double().numpy(), B0.detach().cpu().double().numpy())
        for s, P in zip(spec.species, parts):
            a = P.detach().cpu().double().numpy()
            n = len(a)
            kept = r.add_particles(s, *[a[:, c] for c in range(7)], ids=np.arange(n, dtype=np.int32))
            if kept != n:
                raise ValueError(f"species {s}: {n - kept} particles outside the is_inside cells")
        r.init_boundaries()
        if tape:
            r.tape_begin(spec.every)
        r.step(spec.nsteps)
        E, B, _, _ = r.get_fields()
        outs = []
        for s, P in zip(spec.species, parts):
            rec, ids = r.particles(s)
            out = np.zeros((len(P), 7))
            alive = np.zeros(len(P), bool)
            out[ids] = rec[:, [0, 1, 2, 3, 4, 5, 7]]
            alive[ids] = True
            spec.alive[s] = alive
            outs.append(torch.from_numpy(out))
    except Exception:
        r.close()
        raise
    if not tape:
        r.close()
    return r, torch.from_numpy(E), torch.from_numpy(B), outs


class PICRun(torch.autograd.Function):
    """(E0, B0, *particles) -> (E, B, *particles) after spec.nsteps steps.

    Fields are (nx, ny, nz, 3) tensors; the particles of species
    spec.species[k] are an (n, 7) tensor of rows x y z ux uy uz q (positions
    in cells, u = p / m c).  Every particle must lie in an is_inside cell.
    Output rows keep the input order; particles deleted on the way (they
    left the inner region) come back as zero rows, marked False in
    spec.alive.
    """

    @staticmethod
    def forward(ctx, spec, E0, B0, *parts):
        r, E, B, outs = simulate(spec, E0, B0, parts, tape=True)
        ctx.region = r
        ctx.spec = spec
        ctx.shapes = [tuple(P.shape) for P in parts]
        return (E, B, *outs)

    @staticmethod
    def backward(ctx, gE, gB, *gparts):
        spec, r = ctx.spec, ctx.region
        # rows are ids (the order added); the library's adjoint works in the
        # slot order after tape_begin
        lam, order = {}, {s: r.tape_ids(s) for s in spec.species if s in r._tape_ids}
        for s, g in zip(spec.species, gparts):
            if g is not None and s in order:
                lam[s] = (g.detach().cpu().double().numpy() * spec.alive[s][:, None])[order[s]]
        try:
            gE0, gB0, gp = r.adjoint(gE.detach().cpu().double().numpy() if gE is not None else None,
                                     gB.detach().cpu().double().numpy() if gB is not None else None, lam)
        finally:
            r.close()
        grads = []
        for s, P in zip(spec.species, ctx.shapes):
            if s in gp:
                g = np.zeros(P)
                g[order[s]] = gp[s]
                grads.append(torch.from_numpy(g))
            else:
                grads.append(None)
        return (None, torch.from_numpy(gE0), torch.from_numpy(gB0), *grads)


def run(spec, E0, B0, *parts):
    """Differentiable run: PICRun.apply with the spec first."""
    return PICRun.apply(spec, E0, B0, *parts)


# ---- differentiable set-up of the laser and density parameters --------------

N_PHASE = 16   # transverse-phase coefficients: cos modes 0..3 in y times 0..3 in z


def laser_fields(theta, shape, d):
    """Initial E, B of a linearly polarized pulse travelling along +x, the
    cos-envelope field of f_init_cos (src/init.cpp:6-210, sscos = 0, the
    components at the nodes as there) with two additions of this builder:
    a linear chirp and a transverse phase.  The carrier phase is
    x + phase + chirp x^2 + sum_ab c_ab cos(a pi y / 2 ys) cos(b pi z / 2 zs).

    theta: dict of 0-d tensors a0, xsigma, ysigma, zsigma, x0 (centre from
    the left edge), phase, chirp, and c (N_PHASE); lengths in c/omega.
    Returns (E, B) as (nx, ny, nz, 3) FP64 tensors."""
    nx, ny, nz = shape
    dx, dy, dz = d
    pi = np.pi
    f = 2 * np.sqrt(2 * pi) / 3   # the envelope half-length per sigma (src/init.cpp:35-37)
    xs, ys, zs = theta["xsigma"] * f, theta["ysigma"] * f, theta["zsigma"] * f
    X = torch.arange(nx, dtype=torch.float64)[:, None, None] * dx - theta["x0"]
    Y = torch.arange(ny, dtype=torch.float64)[None, :, None] * dy - 0.5 * ny * dy
    Z = torch.arange(nz, dtype=torch.float64)[None, None, :] * dz - 0.5 * nz * dz
    inside = ((X.abs() < xs) & (Y.abs() < ys) & (Z.abs() < zs)).to(torch.float64)
    modes_y = torch.stack([torch.cos(a * pi * Y / (2 * ys)) for a in range(4)])
    modes_z = torch.stack([torch.cos(b * pi * Z / (2 * zs)) for b in range(4)])
    c = theta["c"].reshape(4, 4)
    phi_t = torch.einsum("ab,ajk,bjk->jk", c, modes_y.expand(4, 1, ny, 1).reshape(4, ny, 1),
                         modes_z.expand(4, 1, 1, nz).reshape(4, 1, nz))[None]
    xi = X + theta["phase"] + theta["chirp"] * X * X + phi_t
    cx, sx = torch.cos(pi * X / 2 / xs), torch.sin(pi * X / 2 / xs)
    cy, sy = torch.cos(pi * Y / 2 / ys), torch.sin(pi * Y / 2 / ys)
    cz, sz = torch.cos(pi * Z / 2 / zs), torch.sin(pi * Z / 2 / zs)
    a0 = theta["a0"]
    fy = a0 * cy * cy * cz * cz * (torch.cos(xi) * cx * cx - pi / xs * torch.sin(xi) * cx * sx)
    fx = a0 * torch.sin(xi) * cx * cx * cz * cz * pi / ys * cy * sy
    gx = a0 * torch.sin(xi) * cx * cx * cy * cy * pi / zs * cz * sz
    zero = torch.zeros_like(fy)
    E = torch.stack([fx, fy, zero], -1) * inside[..., None]
    B = torch.stack([gx, zero, fy], -1) * inside[..., None]
    return E, B


def ramp_charges(nodes, x, x_lo, x_hi, ppc):
    """Electron charges of a density ramp: n(x) piecewise linear over len(nodes)
    equally spaced nodes on [x_lo, x_hi] (n_cr units, constant beyond), q =
    -n(x) / ppc at each particle's x (cells)."""
    m = len(nodes)
    t = ((x - x_lo) / (x_hi - x_lo) * (m - 1)).clamp(0, m - 1)
    i0 = t.floor().long().clamp(max=m - 2)
    w = t - i0
    n = nodes[i0] * (1 - w) + nodes[i0 + 1] * w
    return -n / ppc


def softplus_energy(P, threshold, width):
    """sum over particles of |q| * width * softplus((gamma - 1 - threshold) / width):
    the kinetic energy above `threshold` (mc^2 units), smoothed over `width`
    so that it has a gradient everywhere; deleted particles (zero rows)
    contribute width * softplus(-threshold / width) * 0 = 0 through q = 0."""
    gam = torch.sqrt(1 + (P[:, 3:6] ** 2).sum(1))
    return (P[:, 6].abs() * width * torch.nn.functional.softplus((gam - 1 - threshold) / width)).sum()


class LaserPlasma:
    """A laser pulse in a plasma whose density varies along x, differentiable
    in the pulse's and the density profile's parameters.

    The pulse is laser_fields' cos-envelope pulse (a0, waist = ysigma =
    zsigma, duration xsigma, linear chirp, N_PHASE transverse-phase
    coefficients) placed at x0 inside the box at t = 0.  The electrons, ppc
    per is_inside cell at positions uniform in each cell from `seed`, start at
    rest with charges from a piecewise-linear density of `nodes` equally
    spaced over the box (ramp_charges); the ions are an immobile background.
    Electrons are loaded `margin` cells or more away from the layers where
    particles are deleted.
    The objective is the final electrons' kinetic energy above `threshold`
    mc^2 (softplus_energy).  Lengths in lambda, densities in n_cr."""

    def __init__(self, shape=(128, 64, 64), ppc=8, steps=300, every=25, dx_lambda=0.05, dt_lambda=0.025,
                 solver="ndfx", pusher="vay", density=0.005, n_nodes=32, a0=2.0, waist=0.8, duration=1.0,
                 x0=1.8, chirp=0.02, threshold=5.0, width=1.0, seed=7, margin=0):
        self.shape, self.ppc, self.steps, self.every = tuple(shape), ppc, steps, every
        self.d = dx_lambda * 2 * np.pi
        self.dt = dt_lambda * 2 * np.pi
        self.solver, self.pusher = solver, pusher
        self.threshold, self.width = threshold, width
        nx, ny, nz = self.shape
        rng = np.random.default_rng(seed)
        # every is_inside cell (1 <= i < nx - 2, likewise j, k) at least
        # `margin` cells from the deleting layers, ppc positions each
        m = 1 + margin
        i, j, k = np.meshgrid(np.arange(m, nx - 1 - m), np.arange(m, ny - 1 - m), np.arange(m, nz - 1 - m),
                              indexing="ij")
        cells = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).repeat(ppc, 0).astype(np.float64)
        self.pos = torch.from_numpy(cells + rng.uniform(0, 1, cells.shape))
        lam = 2 * np.pi
        ramp = density * np.linspace(0.5, 1.5, n_nodes)   # a linear ramp about `density`
        self.theta0 = {
            "a0": a0, "ysigma": waist * lam, "zsigma": waist * lam, "xsigma": duration * lam, "x0": x0 * lam,
            "phase": 0.0, "chirp": chirp / lam, "c": 0.1 * rng.standard_normal(N_PHASE), "nodes": ramp}
        self.theta0 = {k: torch.tensor(v, dtype=torch.float64) for k, v in self.theta0.items()}

    # the parameters that are differentiated; the rest of theta0 stays fixed
    FREE = ("a0", "ysigma", "chirp", "c", "nodes")

    def spec(self, precision):
        nx, ny, nz = self.shape
        return RunSpec(region=dict(nx=nx, ny=ny, nz=nz, dx=self.d, dy=self.d, dz=self.d, dt=self.dt,
                                   solver=self.solver, pusher=self.pusher, precision=precision),
                       nsteps=self.steps, every=self.every, species=(0,))

    def inputs(self, theta):
        """(E0, B0, electrons (n, 7)) from the parameters (torch, differentiable)."""
        th = dict(theta)
        th["zsigma"] = th["ysigma"]
        E0, B0 = laser_fields(th, self.shape, (self.d,) * 3)
        x = self.pos[:, 0]
        q = ramp_charges(th["nodes"], x, 0.0, float(self.shape[0]), self.ppc)
        P = torch.cat([self.pos, torch.zeros(len(x), 3, dtype=torch.float64), q[:, None]], 1)
        return E0, B0, P

    def objective(self, P):
        return softplus_energy(P, self.threshold, self.width)

    def gradient(self, precision="fp32", theta=None):
        """(objective, {name: d objective / d name}) through the adjoint."""
        theta = {k: v.clone().requires_grad_(k in self.FREE) for k, v in (theta or self.theta0).items()}
        E0, B0, P = self.inputs(theta)
        _, _, Pf = run(self.spec(precision), E0, B0, P)
        L = self.objective(Pf)
        L.backward()
        return float(L), {k: theta[k].grad.detach().clone() for k in self.FREE}

    def value(self, precision="fp64", theta=None):
        """The objective from a forward run alone."""
        with torch.no_grad():
            E0, B0, P = self.inputs(theta or self.theta0)
            _, _, _, outs = simulate(self.spec(precision), E0, B0, [P])
            return float(self.objective(outs[0]))

    def direction(self, seed):
        """A random direction over FREE, each entry scaled to its parameter."""
        rng = np.random.default_rng(seed)
        scale = {"a0": 0.1, "ysigma": 0.3, "chirp": 0.002, "c": 0.1, "nodes": 0.001}
        return {k: torch.from_numpy(np.asarray(rng.standard_normal(self.theta0[k].shape) * scale[k])) for k in self.FREE}

    def central_difference(self, v, h, precision="fp64"):
        """(L(theta + h v) - L(theta - h v)) / 2h from forward runs."""
        plus = {k: t + h * v[k] if k in v else t for k, t in self.theta0.items()}
        minus = {k: t - h * v[k] if k in v else t for k, t in self.theta0.items()}
        return (self.value(precision, plus) - self.value(precision, minus)) / (2 * h)


__all__ = ["RunSpec", "PICRun", "run", "simulate", "ION0", "laser_fields", "ramp_charges", "N_PHASE",
           "softplus_energy", "LaserPlasma"]
