"""The QED statistics case shared by make_golden.py and the tests.

A cold electron beam (u_x = 200) in a uniform B_z = 2000 (m c omega / e):
chi ~ 1, so photons are emitted within a step and the hardest photons make
Breit-Wheeler pairs.  The reference runs it with its own random stream for
16 seeds (qed_stats.npz); the oracle and the GPU run it with their
counter-based streams, and the summaries are compared as distributions.

The pair case (inputs_pairs, qed_pairs.npz): u_x = 300 in B_z = 3000 for 12
steps, chi ~ 2, about 1400 pairs per run, so the positron yield and
spectrum are measured to a few per cent per run.
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


def inputs_pairs():
    c = inputs(b0=3000.0, n=4000)
    c["rec"][:, 3] = 300.0
    c["nsteps"] = 12
    return c


PAIR_BINS = np.geomspace(1.0, 300.0, 8)   # positron gamma histogram edges


def inputs_cascade():
    """A reduced QED cascade of the colliding-pulse kind: 32^3 cells over a
    2 lambda cube (lambda = 1 um), five-point solver, Vay pusher, dt = dx / 4.
    Two counter-propagating circularly polarised pulses (f_init_cos, peak
    |a| = 1000, waist 1.5 lambda, duration 5 lambda/c) centred on the box as
    initial conditions, the reference's way of setting a laser; 1000
    electrons at rest, uniform in the central (lambda / 2)^3; 200 steps.
    About 1800 positrons and 3e4 photons per run.  Returns the set-up and the
    two pulses' f_init_cos argument vectors."""
    n = 32
    lam = 2 * PI
    d = 2 * lam / n
    dt = 0.25 * d
    a = 1000 / np.sqrt(2)
    lasers = []
    for x1 in (-1.0, 1.0):
        L = np.zeros(16)
        # a0y a0z xsigma ysigma zsigma x0 b_sign x1 phase y0 z0 append phi xt yt zt
        L[:12] = [a, a, 2.5 * lam, 1.5 * lam, 1.5 * lam, n * d / 2, 1, x1, 0.0, 0.0, 0.0, 1]
        lasers.append(L)
    nseed = 1000
    rng = np.random.default_rng(2025)
    rec = np.zeros((nseed, 8))
    rec[:, 0:3] = n / 2 + (rng.random((nseed, 3)) - 0.5) * (0.5 * lam / d)
    rec[:, 6] = -1e-3
    rec[:, 7] = -1.0
    ids = np.arange(nseed, dtype=np.int32)
    return dict(nx=n, ny=n, nz=n, dx=d, dt=dt, lasers=lasers, rec=rec, ids=ids, nsteps=200)


CASCADE_BINS = np.geomspace(1.0, 3000.0, 9)   # photon gamma histogram edges of the cascade
CASCADE_PAIR_BINS = np.geomspace(1.0, 3000.0, 7)   # positron gamma histogram edges of the cascade


def summary_cascade(d, photon_g, positron_g):
    """Yields of the cascade (electron, positron and photon counts and
    energies from the last step's diagnostics d) and the photon and positron
    gamma histograms."""
    return np.r_[float(d["n_e"]), float(d["n_p"]), float(d["n_ph"]), float(d["energy_e"]), float(d["energy_p"]),
                 float(d["energy_ph"]), np.histogram(photon_g, bins=CASCADE_BINS)[0].astype(float),
                 np.histogram(positron_g, bins=CASCADE_PAIR_BINS)[0].astype(float)]


NAMES_CASCADE = (["n_e", "n_p", "n_ph", "energy_e", "energy_p", "energy_ph"] +
                 [f"photon hist[{a:.3g},{b:.3g})" for a, b in zip(CASCADE_BINS[:-1], CASCADE_BINS[1:])] +
                 [f"positron hist[{a:.3g},{b:.3g})" for a, b in zip(CASCADE_PAIR_BINS[:-1],
                                                                    CASCADE_PAIR_BINS[1:])])
YIELDS_CASCADE = 6


def summary_pairs(n_ph, n_p, energy_ph, energy_p, photon_g, positron_g):
    """Yields (photon and positron counts and energies), then the photon and
    positron gamma histograms."""
    return np.r_[float(n_ph), float(n_p), float(energy_ph), float(energy_p),
                 np.histogram(photon_g, bins=BINS)[0].astype(float),
                 np.histogram(positron_g, bins=PAIR_BINS)[0].astype(float)]


NAMES_PAIRS = (["n_ph", "n_p", "energy_ph", "energy_p"] +
               [f"photon hist[{a:.3g},{b:.3g})" for a, b in zip(BINS[:-1], BINS[1:])] +
               [f"positron hist[{a:.3g},{b:.3g})" for a, b in zip(PAIR_BINS[:-1], PAIR_BINS[1:])])
YIELDS = 4   # leading columns of summary_pairs that are yields


def spectrum_chi2_p(z):
    """p-value of sum z^2 over the spectrum bins (chi-square with one degree
    of freedom per bin): the spectra agree within 3 sigma as a whole when
    p > 0.0027."""
    from scipy.stats import chi2
    z = np.asarray(z, float)
    return float(chi2.sf((z * z).sum(), len(z)))


def summary(n_ph, n_p, energy_ph, photon_g):
    """Per-run numbers compared across implementations: final photon and
    positron counts, photon energy, photon-gamma histogram."""
    return np.r_[float(n_ph), float(n_p), float(energy_ph), np.histogram(photon_g, bins=BINS)[0].astype(float)]


NAMES = ["n_ph", "n_p", "energy_ph"] + [f"hist[{a:.3g},{b:.3g})" for a, b in zip(BINS[:-1], BINS[1:])]
YIELDS_STATS = 3   # leading columns of summary that are yields


def zscores(a, b):
    """Two-sample z of the means, per column (runs are rows)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    return (a.mean(0) - b.mean(0)) / np.where(se > 0, se, 1.0)
