"""Laser sources at an x plane (qk_add_laser_source) on the GPU against the
oracle's (oq_add_laser_source), on identical inputs: two pulses entering
through x-min and x-max into a thermal plasma, each solver, FP64 to 1e-10
and FP32 to the 1e-4 field bar."""
import math

import numpy as np
import pytest

import oracle
import quill_amd as qa

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi
DX = 0.05 * TWO_PI
SOLVERS = {"fdtd": (oracle.FDTD, 0.5), "ndfx": (oracle.NDFX, 0.5), "five_point": (oracle.FIVE_POINT, 0.35)}


def pulses(nx):
    """(side, plane, L): a linearly polarised pulse entering through x-min and
    a circularly polarised one through x-max, both with their fronts at the
    walls when the run starts."""
    xsig = 0.3 * TWO_PI
    xs = xsig * 2 * math.sqrt(2 * math.pi) / 3
    wy = 0.3 * TWO_PI
    left = [1.0, 0.2, xsig, wy, wy, -xs, 1, -1.0, 0.3, 0.0, 0.0, 1, 0.0, 0.0, 0.0, 0.0]
    right = [0.7, 0.7, xsig, wy, 0.8 * wy, nx * DX + xs, 1, 1.0, 0.0, 0.1, -0.1, 1, 0.2, 0.0, 0.0, 0.0]
    return [(0, 3, np.array(left)), (1, nx - 4, np.array(right))]


def run_pair(solver, nsteps=40, shape=(40, 24, 24)):
    code, dtf = SOLVERS[solver]
    dt = dtf * DX
    nx, ny, nz = shape
    o = oracle.Oracle(*shape, DX, DX, DX, dt, solver=code, pusher=oracle.VAY)
    rec, ids = oracle.synth_plasma(*shape, 1, 0.02, 0.1, 17)
    o.add_particles(rec, ids)
    for side, plane, L in pulses(nx):
        o.laser_source(L, side, plane)
    o.init_boundaries()
    o.step(nsteps)
    out = {}
    for precision in ("fp64", "fp32"):
        print(f"laser sources {solver} {precision}", flush=True)
        with qa.Region(*shape, DX, DX, DX, dt, solver=solver, pusher="vay", precision=precision, deposit="tiled",
                       sort_every=5, track_ids=True) as r:
            r.add_records(rec, ids)
            for side, plane, L in pulses(nx):
                r.add_laser_source(side, plane, *L[:11], phi=L[12])
            r.init_boundaries()
            r.step(nsteps)
            out[precision] = r.get_fields()
    return o, out


@pytest.mark.parametrize("solver", sorted(SOLVERS))
def test_laser_sources_match_oracle(solver):
    o, out = run_pair(solver)
    ref = (o.field(oracle.E), o.field(oracle.B), o.field(oracle.J))
    assert np.abs(ref[0][..., 1]).max() > 0.1   # the pulses are in
    for name, a, b in zip("EBJ", out["fp64"][:3], ref):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"FP64 {name}: {err:.2e}", flush=True)
        assert err < 1e-10, name
    for name, a, b in zip("EB", out["fp32"][:2], ref[:2]):
        for c in range(3):
            err = np.linalg.norm(a[..., c] - b[..., c]) / np.linalg.norm(b[..., c])
            print(f"FP32 {name}{'xyz'[c]}: {err:.2e}", flush=True)
            assert err < 1e-4, (name, c)
