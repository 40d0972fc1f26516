"""The QED statistics case shared by make_golden.py and the tests.

A cold electron beam (u_x = 200) in a uniform B_z = 2000 (m c omega / e):
chi ~ 1, so photons are emitted within a step and the hardest photons make
Breit-Wheeler pairs.  The reference runs it with its own random stream for
16 seeds (qed_stats.npz); the oracle and the GPU run it with their
counter-based streams, and the summaries are compared as distributions.
"""
import numpy as np

PI = np.pi
SEEDS = 16
BINS = np.geomspace(1.0, 400.0, 9)   # photon gamma histogram edges
E_S = 1e-4 / 2.4263086e-10           # src/main.cpp:2276 for lambda = 1 um


def inputs(b0=2000.0, n=4000):
    nx = ny = nz = 24
    dx = 0.05 * 2 * PI
    dt = 0.025 * 2 * PI
    E = np.zeros((nx, ny, nz, 3))
    B = np.zeros((nx, ny, nz, 3))
    B[..., 2] = b0
    rng = np.random.default_rng(2024)
    rec = np.zeros((n, 8))
    rec[:, 0:3] = rng.uniform(8, 16, (n, 3))
    rec[:, 3] = 200.0
    rec[:, 6] = -1e-3
    rec[:, 7] = -1.0
    ids = np.arange(n, dtype=np.int32)
    return dict(nx=nx, ny=ny, nz=nz, dx=dx, dt=dt, E=E, B=B, rec=rec, ids=ids, nsteps=10)


def summary(n_ph, n_p, energy_ph, photon_g):
    """Per-run numbers compared across implementations: final photon and
    positron counts, photon energy, photon-gamma histogram."""
    return np.r_[float(n_ph), float(n_p), float(energy_ph), np.histogram(photon_g, bins=BINS)[0].astype(float)]


NAMES = ["n_ph", "n_p", "energy_ph"] + [f"hist[{a:.3g},{b:.3g})" for a, b in zip(BINS[:-1], BINS[1:])]


def zscores(a, b):
    """Two-sample z of the means, per column (runs are rows)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    return (a.mean(0) - b.mean(0)) / np.where(se > 0, se, 1.0)
