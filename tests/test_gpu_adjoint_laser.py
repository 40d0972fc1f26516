"""Gradients with respect to laser and density-profile parameters.

LaserPlasma builds the initial fields from the pulse parameters (a0, waist,
linear chirp, 16 transverse-phase coefficients) and the electron charges from
a piecewise-linear density ramp, in torch; PICRun carries the gradient of the
final electrons' kinetic energy above a threshold back through the run with
the library's adjoint (snapshots every 10 of 30 steps).  Directional
derivatives along random parameter directions, and along the laser and the
density blocks alone, are checked against central differences of the FP64
oracle run on the same inputs: the FP64 adjoint to 1e-6 relative, the FP32
adjoint to 1e-3.

The discrete step is piecewise smooth with jumps: each field component is
interpolated bilinearly across its own axis only (src/advance.cpp, the
gather), so the force on a particle jumps when it crosses a cell face, and
particles entering the outer layers are deleted.  A difference quotient
whose interval holds such an event does not approximate the derivative, so
each quotient is taken at two step sizes, h and h / 10, which must agree to
1e-7 before it is compared: the interval is then free of those events.  The
electrons start 2 cells away from the deleting layers.
"""
import numpy as np
import pytest
import torch

import oracle
from quill_amd.adjoint import LaserPlasma

pytestmark = pytest.mark.gpu

LASER = ("a0", "ysigma", "chirp", "c")
SOLVERS = {"fdtd": oracle.FDTD, "ndfx": oracle.NDFX, "five_point": oracle.FIVE_POINT}
PUSHERS = {"boris": oracle.BORIS, "vay": oracle.VAY}


def small():
    # 32x16x16 cells, 2 per cell, a pulse of waist 0.2 lambda in the box
    return LaserPlasma(shape=(32, 16, 16), ppc=2, steps=30, every=10, waist=0.2, duration=0.3, x0=0.6,
                       threshold=0.5, width=0.5, margin=2)


def oracle_value(lp, theta):
    """The objective after lp.steps steps of the FP64 oracle."""
    with torch.no_grad():
        E0, B0, P = lp.inputs(theta)
    nx, ny, nz = lp.shape
    o = oracle.Oracle(nx, ny, nz, lp.d, lp.d, lp.d, lp.dt, solver=SOLVERS[lp.solver], pusher=PUSHERS[lp.pusher])
    o.field(oracle.E)[:] = E0.numpy()
    o.field(oracle.B)[:] = B0.numpy()
    Pn = P.numpy()
    ids = np.arange(len(Pn), dtype=np.int32)
    assert o.add_particles(np.c_[Pn, np.full(len(Pn), -1.0)], ids) == len(Pn)
    o.init_boundaries()
    o.step(lp.steps)
    rec, ids = o.particles()
    out = np.zeros((len(Pn), 7))
    out[ids] = rec[:, [0, 1, 2, 3, 4, 5, 7]]
    return float(lp.objective(torch.from_numpy(out)))


def central_difference(lp, v, h):
    plus = {k: t + h * v[k] if k in v else t for k, t in lp.theta0.items()}
    minus = {k: t - h * v[k] if k in v else t for k, t in lp.theta0.items()}
    return (oracle_value(lp, plus) - oracle_value(lp, minus)) / (2 * h)


def dot(g, v):
    return sum(float((g[k] * v[k]).sum()) for k in v)


@pytest.fixture(scope="module")
def case():
    lp = small()
    print("laser-plasma adjoint 32x16x16 ppc 2, 30 steps", flush=True)
    L64, g64 = lp.gradient("fp64")
    L32, g32 = lp.gradient("fp32")
    L0 = oracle_value(lp, lp.theta0)
    assert L64 == pytest.approx(L0, rel=1e-10)
    assert L32 == pytest.approx(L0, rel=1e-4)
    return lp, g64, g32


@pytest.mark.parametrize("block", ["all", "laser", "density"])
def test_directional_derivative(case, block):
    lp, g64, g32 = case
    v = lp.direction(5)
    if block == "laser":
        v = {k: t for k, t in v.items() if k in LASER}
    elif block == "density":
        v = {"nodes": v["nodes"]}
    fd = central_difference(lp, v, 1e-6)
    fd10 = central_difference(lp, v, 1e-7)
    assert fd10 == pytest.approx(fd, rel=1e-7), "a face crossing or a deletion inside the interval"
    assert dot(g64, v) == pytest.approx(fd, rel=1e-6), (dot(g64, v), fd)
    assert dot(g32, v) == pytest.approx(fd, rel=1e-3), (dot(g32, v), fd)


def test_fp32_gradient_vector_matches_fp64(case):
    """The whole FP32 gradient against the FP64 reference gradient."""
    _, g64, g32 = case
    a = torch.cat([g64[k].reshape(-1) for k in g64])
    b = torch.cat([g32[k].reshape(-1) for k in g32])
    assert float((a - b).norm() / a.norm()) < 1e-3
