"""The deterministic step (QED off) as a torch.autograd.Function.

forward runs `nsteps` steps of one subdomain on the GPU (libquillk.so) from
initial fields and particles and returns the final ones; backward runs the
library's hand-written adjoint (qk_tape_begin / qk_adjoint): the
transposes of the three Maxwell solvers and of interpolate_be, and a
per-particle adjoint of gather, push and charge-conserving deposit, swept
backwards over a trajectory kept as snapshots every `every` steps and
recomputed one segment at a time.  The differentiated loop body is
src/main.cpp:2319-2332 (padvance, src/advance.cpp:90-387, and fadvance,
src/advance.cpp:13-83) with QED, merging and sorting off.

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
    precision, ion_cmr, ...); deposit, sort_every and track_ids are set here.
    species: the species slots whose particles are passed, in order.
    """
    region: dict
    nsteps: int
    every: int = 25
    species: tuple = (0,)
    alive: dict = field(default_factory=dict)   # filled by forward: species -> bool mask of survivors


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
        kw = dict(spec.region)
        kw.update(deposit="global", sort_every=0, track_ids=True)
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
        ctx.region = r
        ctx.spec = spec
        return (torch.from_numpy(E), torch.from_numpy(B), *outs)

    @staticmethod
    def backward(ctx, gE, gB, *gparts):
        spec, r = ctx.spec, ctx.region
        lam = {}
        for s, g in zip(spec.species, gparts):
            if g is not None:
                lam[s] = g.detach().cpu().double().numpy() * spec.alive[s][:, None]
        try:
            gE0, gB0, gp = r.adjoint(gE.detach().cpu().double().numpy() if gE is not None else None,
                                     gB.detach().cpu().double().numpy() if gB is not None else None, lam)
        finally:
            r.close()
        return (None, torch.from_numpy(gE0), torch.from_numpy(gB0),
                *[torch.from_numpy(gp[s]) if s in gp else None for s in spec.species])


def run(spec, E0, B0, *parts):
    """Differentiable run: PICRun.apply with the spec first."""
    return PICRun.apply(spec, E0, B0, *parts)


__all__ = ["RunSpec", "PICRun", "run", "ION0"]
