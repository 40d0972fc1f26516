"""The adjoint checked at full size without finite differences.

qk_tangent_begin / qk_tangent_get run the tangent-linear model J v of the
recorded steps (forward mode along the forward step's own branches);
qk_adjoint gives J^T lam on the same tape.  For any lam and v,
<lam, J v> = <J^T lam, v> holds to round-off when the adjoint is the exact
transpose of the derivative the step has; a wrong sign, a dropped term or a
transposed operand in either breaks it by O(1).

* jump-free small case: the tangent equals FP64 central differences of the
  forward runs of this library (jump-free: quotients at h and h / 10 agree);
* the differentiable laser-plasma case at its full size -- 128 x 64 x 64
  cells, 8 per cell at 0.005 n_cr (3.72 M electrons), a0 = 2 laser, NDFX +
  Vay, 300 steps, a snapshot every 25 -- the dot-product identity, FP64 to
  1e-9 and FP32 to 1e-3 (random lam over the final fields and the surviving
  particles, random v over the initial fields and particles).

Finite differences are ill-posed at that size: each field component is
interpolated across its own axis only, so the force jumps when a particle
crosses a cell face, and with 3.7 M particles over 300 steps every practical
step h contains such jumps.
"""
import math

import numpy as np
import pytest

import quill_amd as qa

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi
D = 0.05 * TWO_PI


def dot_fields(a, b):
    return float((a * b).sum())


def small_region(E0, B0, parts, precision, tape_every, seed_v=None, h=0.0):
    nx, ny, nz = E0.shape[:3]
    r = qa.Region(nx, ny, nz, D, D, D, 0.5 * D, solver="ndfx", pusher="vay", precision=precision, deposit="tiled",
                  sort_every=0, sort_fallback=0.0, track_ids=True)
    E, B, P = E0, B0, parts
    if seed_v is not None:
        E = E0 + h * seed_v["E"]
        B = B0 + h * seed_v["B"]
        P = parts + h * seed_v["p"]
    r.set_fields(E, B)
    ids = np.arange(len(P), dtype=np.int32)
    r.add_particles(qa.ELECTRON, *[P[:, c] for c in range(7)], ids=ids)
    r.init_boundaries()
    r.tape_begin(tape_every)
    return r


def final_functional(r, lam, order):
    """<lam, final state> with the particle part in the input order."""
    E, B, _, _ = r.get_fields()
    rec, ids = r.particles(qa.ELECTRON)
    P = np.zeros((len(order), 7))
    P[ids] = rec[:, [0, 1, 2, 3, 4, 5, 7]]
    return dot_fields(lam["E"], E) + dot_fields(lam["B"], B) + float((lam["p"] * P).sum())


def test_tangent_matches_central_differences():
    rng = np.random.default_rng(3)
    nx, ny, nz, n, steps = 16, 12, 12, 300, 10
    E0 = rng.uniform(-0.3, 0.3, (nx, ny, nz, 3))
    B0 = rng.uniform(-0.3, 0.3, (nx, ny, nz, 3))
    P = np.zeros((n, 7))
    P[:, 0] = rng.uniform(5, nx - 5, n)
    P[:, 1] = rng.uniform(4, ny - 4, n)
    P[:, 2] = rng.uniform(4, nz - 4, n)
    P[:, 3:6] = rng.normal(0, 0.05, (n, 3))
    P[:, 6] = -0.05 * rng.uniform(0.5, 1.5, n)
    v = {"E": rng.normal(size=E0.shape), "B": rng.normal(size=B0.shape), "p": rng.normal(size=P.shape) * 0.1}
    lam = {"E": rng.normal(size=E0.shape), "B": rng.normal(size=B0.shape), "p": rng.normal(size=P.shape)}
    order = np.arange(n)

    r = small_region(E0, B0, P, "fp64", 3)
    slot_ids = r.tape_ids(qa.ELECTRON)
    r.tangent_begin(v["E"], v["B"], {qa.ELECTRON: v["p"][slot_ids]})
    r.step(steps)
    tE, tB, tp = r.tangent_get()
    tP = np.zeros_like(P)
    tP[slot_ids] = tp[qa.ELECTRON]
    assert r.count(qa.ELECTRON) == n
    r.close()
    tangent = dot_fields(lam["E"], tE) + dot_fields(lam["B"], tB) + float((lam["p"] * tP).sum())

    def quotient(h):
        out = []
        for sign in (1, -1):
            rr = small_region(E0, B0, P, "fp64", 3, v, sign * h)
            rr.step(steps)
            out.append(final_functional(rr, lam, order))
            rr.close()
        return (out[0] - out[1]) / (2 * h)

    fd, fd10 = quotient(1e-6), quotient(1e-7)
    print(f"tangent {tangent:.12e}, central differences {fd:.12e} (h / 10: {fd10:.12e})", flush=True)
    assert abs(fd10 - fd) <= 1e-7 * abs(fd), "the interval holds a jump"
    assert abs(tangent - fd) <= 1e-6 * abs(fd)


def laser_plasma_region(precision):
    """The differentiable laser-plasma case at its full size (uniform density
    at 0.005 n_cr, a0 = 2 pulse at x-min)."""
    nx, ny, nz = 128, 64, 64
    r = qa.Region(nx, ny, nz, D, D, D, 0.5 * D, solver="ndfx", pusher="vay", precision=precision, deposit="tiled",
                  sort_every=0, sort_fallback=0.0, track_ids=True)
    r.load_uniform_plasma(qa.ELECTRON, 8, 0.005, 0.01, 7)
    xsigma = 1.5 * TWO_PI
    xs = xsigma * 2 * math.sqrt(2 * math.pi) / 3
    r.init_laser_cos(a0y=2.0, a0z=0.0, xsigma=xsigma, ysigma=1.5 * TWO_PI, zsigma=1.5 * TWO_PI, x0=xs, b_sign=1,
                     x1=xs - nx * D / 2)
    r.init_boundaries()
    return r


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision,bar", [("fp64", 1e-9), ("fp32", 1e-3)])
def test_dot_product_identity_full_size(precision, bar):
    steps, every = 300, 25
    print(f"dot-product check {precision}: 128x64x64, {steps} steps, snapshot every {every}", flush=True)
    r = laser_plasma_region(precision)
    rng = np.random.default_rng(11)
    shape = r.shape + (3,)
    r.tape_begin(every)
    n = r.count(qa.ELECTRON)
    assert n == 125 * 61 * 61 * 8
    v = {"E": rng.normal(size=shape) * 1e-2, "B": rng.normal(size=shape) * 1e-2,
         "p": rng.normal(size=(n, 7)) * np.array([1e-2, 1e-2, 1e-2, 1e-3, 1e-3, 1e-3, 1e-6])}
    r.tangent_begin(v["E"], v["B"], {qa.ELECTRON: v["p"]})
    r.step(steps)
    tE, tB, tp = r.tangent_get()
    # lam: random on the final fields and on the slots still alive
    slot_of = {int(i): s for s, i in enumerate(r.tape_ids(qa.ELECTRON))}
    _, live_ids = r.particles(qa.ELECTRON)
    alive = np.zeros(n, bool)
    alive[[slot_of[int(i)] for i in live_ids]] = True
    lam_E, lam_B = rng.normal(size=shape), rng.normal(size=shape)
    lam_p = rng.normal(size=(n, 7)) * alive[:, None]
    lhs = dot_fields(lam_E, tE) + dot_fields(lam_B, tB) + float((lam_p * tp[qa.ELECTRON]).sum())
    gE, gB, gp = r.adjoint(lam_E, lam_B, {qa.ELECTRON: lam_p})
    rhs = dot_fields(gE, v["E"]) + dot_fields(gB, v["B"]) + float((gp[qa.ELECTRON] * v["p"]).sum())
    r.close()
    rel = abs(lhs - rhs) / abs(rhs)
    print(f"<lam, J v> = {lhs:.15e}, <J^T lam, v> = {rhs:.15e}, relative difference {rel:.2e}; "
          f"{alive.sum()} of {n} electrons alive", flush=True)
    assert rel < bar
