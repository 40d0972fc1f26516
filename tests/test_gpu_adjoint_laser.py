"""Gradients with respect to laser and density-profile parameters.

LaserPlasma builds the initial fields from the pulse parameters (a0, waist,
linear chirp, 16 transverse-phase coefficients) and the electron charges from
a piecewise-linear density ramp, in torch; PICRun carries the gradient of the
final electrons' kinetic energy above a threshold back through the run with
the library's adjoint (snapshots every 10 of 40 steps).  Directional
derivatives along random parameter directions, and along the laser and the
density blocks alone, are checked against central differences of FP64
forward runs of the library (whose FP64 step matches the oracle's): the FP64
adjoint to 1e-6 relative, the FP32 adjoint to 1e-3.
"""
import numpy as np
import pytest
import torch

from quill_amd.adjoint import LaserPlasma

pytestmark = pytest.mark.gpu

LASER = ("a0", "ysigma", "chirp", "c")


def small():
    # 32x16x16 cells, 2 per cell, a pulse of waist 0.2 lambda in the box
    return LaserPlasma(shape=(32, 16, 16), ppc=2, steps=40, every=10, waist=0.2, duration=0.3, x0=0.6,
                       threshold=0.5, width=0.5)


def dot(g, v):
    return sum(float((g[k] * v[k]).sum()) for k in v)


@pytest.fixture(scope="module")
def case():
    lp = small()
    print("laser-plasma adjoint 32x16x16 ppc 2, 40 steps", flush=True)
    L64, g64 = lp.gradient("fp64")
    L32, g32 = lp.gradient("fp32")
    assert L64 > 0
    assert L32 == pytest.approx(L64, rel=1e-4)
    return lp, g64, g32


@pytest.mark.parametrize("block", ["all", "laser", "density"])
def test_directional_derivative(case, block):
    lp, g64, g32 = case
    v = lp.direction(5)
    if block == "laser":
        v = {k: t for k, t in v.items() if k in LASER}
    elif block == "density":
        v = {"nodes": v["nodes"]}
    fd = lp.central_difference(v, 1e-5)
    assert dot(g64, v) == pytest.approx(fd, rel=1e-6), (dot(g64, v), fd)
    assert dot(g32, v) == pytest.approx(fd, rel=1e-3), (dot(g32, v), fd)


def test_fp32_gradient_vector_matches_fp64(case):
    """The whole FP32 gradient against the FP64 reference gradient."""
    _, g64, g32 = case
    a = torch.cat([g64[k].reshape(-1) for k in g64])
    b = torch.cat([g32[k].reshape(-1) for k in g32])
    assert float((a - b).norm() / a.norm()) < 1e-3
