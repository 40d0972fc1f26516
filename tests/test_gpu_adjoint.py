"""The hand-written adjoint of the deterministic step on the GPU.

Gradients of a scalar of the final state -- a random linear functional of
the final fields and particle records plus the particles' kinetic energy --
with respect to the initial fields, positions, momenta and charges, from
qk_tape_begin / qk_adjoint (through the torch.autograd.Function PICRun),
against central finite differences of the FP64 oracle run on the same
inputs (the FP64 reference gradient):

  * fields only: the step is linear, so the adjoint's directional derivative
    must match the oracle's difference quotient to round-off, for all three
    solvers (the transposes of the B and E updates and of interpolate_be);
  * electrons, positrons and ions in random fields: FP64 to 1e-6 relative,
    FP32 to 1e-3 (the stated bar for FP32 gradients), with snapshots every 3 steps so the
    segment recomputation is exercised.
The step is piecewise smooth (face crossings switch the deposit's pieces);
the perturbation is small enough not to cross a kink in these cases, and no
particle gets near the deleted layers.
"""
import numpy as np
import pytest
import torch

import oracle
import quill_amd as qa
from quill_amd.adjoint import RunSpec, run

pytestmark = pytest.mark.gpu

D = 0.05 * 2 * np.pi
SOLVERS = {"fdtd": oracle.FDTD, "ndfx": oracle.NDFX, "five_point": oracle.FIVE_POINT}
PUSHERS = {"boris": oracle.BORIS, "vay": oracle.VAY}
ION_CMR = 1.0 / 1836.1526721
CMR = {qa.ELECTRON: -1.0, qa.POSITRON: 1.0, qa.ION0: ION_CMR}


def case(seed, nx=16, ny=12, nz=12, particles=True):
    rng = np.random.default_rng(seed)
    E = rng.uniform(-0.3, 0.3, (nx, ny, nz, 3))
    B = rng.uniform(-0.3, 0.3, (nx, ny, nz, 3))
    parts = {}
    if particles:
        for s, n, q in ((qa.ELECTRON, 300, -0.05), (qa.POSITRON, 60, 0.05), (qa.ION0, 60, 0.08)):
            P = np.zeros((n, 7))
            P[:, 0] = rng.uniform(5, nx - 5, n)
            P[:, 1] = rng.uniform(4, ny - 4, n)
            P[:, 2] = rng.uniform(4, nz - 4, n)
            P[:, 3:6] = rng.normal(0, 0.05, (n, 3))
            P[:, 6] = q * rng.uniform(0.5, 1.5, n)
            parts[s] = P
    return E, B, parts


def weights(seed, E, parts):
    rng = np.random.default_rng(seed + 1000)
    w = {"E": rng.normal(size=E.shape), "B": rng.normal(size=E.shape),
         "kin": 3.0, "p": {s: rng.normal(size=P.shape) for s, P in parts.items()}}
    return w


def objective(w, E, B, parts):
    """Linear functional of the final state plus sum of (gamma - 1) |q| (numpy or torch)."""
    lib = torch if isinstance(E, torch.Tensor) else np
    arr = (lambda a: torch.from_numpy(a)) if lib is torch else (lambda a: a)
    L = (arr(w["E"]) * E).sum() + (arr(w["B"]) * B).sum()
    for s, P in parts.items():
        L = L + (arr(w["p"][s]) * P).sum()
        if CMR[s] != 0:
            u2 = (P[:, 3:6] ** 2).sum(1)
            L = L + w["kin"] * ((lib.sqrt(1 + u2) - 1) * lib.abs(P[:, 6])).sum()
    return L


def oracle_objective(w, E0, B0, parts, solver, pusher, nsteps, nx, ny, nz):
    o = oracle.Oracle(nx, ny, nz, D, D, D, 0.5 * D, solver=SOLVERS[solver], pusher=PUSHERS[pusher])
    o.field(oracle.E)[:] = E0
    o.field(oracle.B)[:] = B0
    for s, P in parts.items():
        ids = np.arange(len(P), dtype=np.int32) + 100000 * s
        assert o.add_particles(np.c_[P, np.full(len(P), CMR[s])], ids) == len(P)
    o.init_boundaries()
    o.step(nsteps)
    rec, ids = o.particles()
    out = {}
    for s, P in parts.items():
        sel = (ids >= 100000 * s) & (ids < 100000 * s + len(P))
        assert sel.sum() == len(P), "a particle left the inner region"
        r = np.zeros((len(P), 7))
        r[ids[sel] - 100000 * s] = rec[sel][:, [0, 1, 2, 3, 4, 5, 7]]
        out[s] = r
    return objective(w, o.field(oracle.E), o.field(oracle.B), out)


def adjoint_gradient(w, E0, B0, parts, solver, pusher, precision, nsteps, every, nx, ny, nz):
    spec = RunSpec(region=dict(nx=nx, ny=ny, nz=nz, dx=D, dy=D, dz=D, dt=0.5 * D, solver=solver, pusher=pusher,
                               precision=precision, ion_cmr=[ION_CMR]),
                   nsteps=nsteps, every=every, species=tuple(parts))
    tE = torch.tensor(E0, requires_grad=True)
    tB = torch.tensor(B0, requires_grad=True)
    tP = [torch.tensor(P, requires_grad=True) for P in parts.values()]
    out = run(spec, tE, tB, *tP)
    for s in parts:
        assert spec.alive[s].all()
    L = objective(w, out[0], out[1], dict(zip(parts, out[2:])))
    L.backward()
    return float(L), tE.grad.numpy(), tB.grad.numpy(), {s: t.grad.numpy() for s, t in zip(parts, tP)}


def directional(seed, E0, B0, parts):
    rng = np.random.default_rng(seed + 2000)
    dE = rng.normal(size=E0.shape) * 0.3
    dB = rng.normal(size=B0.shape) * 0.3
    dP = {}
    for s, P in parts.items():
        d = np.zeros_like(P)
        d[:, 0:3] = rng.normal(size=(len(P), 3)) * 0.3
        d[:, 3:6] = rng.normal(size=(len(P), 3)) * 0.05
        d[:, 6] = rng.normal(size=len(P)) * 0.05
        dP[s] = d
    return dE, dB, dP


def check(seed, solver, pusher, precision, nsteps, every, particles, rtol, h=1e-6):
    nx, ny, nz = 16, 12, 12
    E0, B0, parts = case(seed, nx, ny, nz, particles)
    w = weights(seed, E0, parts)
    print(f"adjoint {solver}/{pusher} {precision} particles={particles} steps={nsteps} every={every}", flush=True)
    L, gE, gB, gP = adjoint_gradient(w, E0, B0, parts, solver, pusher, precision, nsteps, every, nx, ny, nz)
    dE, dB, dP = directional(seed, E0, B0, parts)
    ad = (gE * dE).sum() + (gB * dB).sum() + sum((gP[s] * dP[s]).sum() for s in parts)
    plus = {s: parts[s] + h * dP[s] for s in parts}
    minus = {s: parts[s] - h * dP[s] for s in parts}
    Lp = oracle_objective(w, E0 + h * dE, B0 + h * dB, plus, solver, pusher, nsteps, nx, ny, nz)
    Lm = oracle_objective(w, E0 - h * dE, B0 - h * dB, minus, solver, pusher, nsteps, nx, ny, nz)
    fd = (Lp - Lm) / (2 * h)
    L0 = oracle_objective(w, E0, B0, parts, solver, pusher, nsteps, nx, ny, nz)
    assert L == pytest.approx(L0, rel=1e-4 if precision == "fp32" else 1e-10)
    assert ad == pytest.approx(fd, rel=rtol), (ad, fd)
    return ad, fd


@pytest.mark.parametrize("solver", ["fdtd", "ndfx", "five_point"])
def test_fields_only_is_exact(solver):
    # linear: a large step keeps the quotient's round-off small
    check(11, solver, "boris", "fp64", 6, 4, False, 1e-9, h=1e-2)


@pytest.mark.parametrize("solver, pusher", [("fdtd", "boris"), ("ndfx", "vay"), ("five_point", "vay")])
def test_fp64_adjoint_matches_oracle_differences(solver, pusher):
    check(21, solver, pusher, "fp64", 8, 3, True, 1e-6)


@pytest.mark.parametrize("solver, pusher", [("fdtd", "boris"), ("ndfx", "vay")])
def test_fp32_adjoint_within_1e_3(solver, pusher):
    check(31, solver, pusher, "fp32", 8, 3, True, 1e-3)


def test_particle_gradients_per_block():
    """Each block of the gradient (fields, positions, momenta, charges)
    against its own oracle difference quotient, FP64."""
    nx, ny, nz = 16, 12, 12
    E0, B0, parts = case(41, nx, ny, nz, True)
    w = weights(41, E0, parts)
    print("adjoint per block", flush=True)
    _, gE, gB, gP = adjoint_gradient(w, E0, B0, parts, "fdtd", "vay", "fp64", 6, 4, nx, ny, nz)
    dE, dB, dP = directional(41, E0, B0, parts)
    h = 1e-6
    zeroP = {s: np.zeros_like(P) for s, P in parts.items()}
    zf = np.zeros_like(E0)
    blocks = {
        "E": (dE, zf, zeroP, (gE * dE).sum()),
        "B": (zf, dB, zeroP, (gB * dB).sum()),
    }
    for name, cols in (("x", slice(0, 3)), ("u", slice(3, 6)), ("q", slice(6, 7))):
        dp = {s: np.zeros_like(P) for s, P in parts.items()}
        for s in parts:
            dp[s][:, cols] = dP[s][:, cols]
        blocks[name] = (zf, zf, dp, sum((gP[s] * dp[s]).sum() for s in parts))
    for name, (a, b, dp, ad) in blocks.items():
        Lp = oracle_objective(w, E0 + h * a, B0 + h * b, {s: parts[s] + h * dp[s] for s in parts}, "fdtd", "vay", 6,
                              nx, ny, nz)
        Lm = oracle_objective(w, E0 - h * a, B0 - h * b, {s: parts[s] - h * dp[s] for s in parts}, "fdtd", "vay", 6,
                              nx, ny, nz)
        fd = (Lp - Lm) / (2 * h)
        assert ad == pytest.approx(fd, rel=1e-6, abs=1e-9), (name, ad, fd)
