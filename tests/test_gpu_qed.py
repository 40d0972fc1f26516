"""QED on the GPU: photon emission, Breit-Wheeler pairs and vacuum pairs.

Event sequences cannot match the reference (its get_rand is a sequential
lagged-Fibonacci stream, src/mc.cpp:12-23; the GPU uses counter-based Philox
streams), so parity is statistical: 16 seeded GPU runs against 16 runs of
the reference's own code (tests/golden/make_golden.py).  Three cases: the
chi ~ 1 beam of qed_stats.npz, the pair-rich chi ~ 2 case of qed_pairs.npz
(about 1400 pairs per run) and the colliding-pulse cascade of
qed_cascade.npz.  Each compares yields within 3 sigma (|z| < 3) and its
spectra each within 3 sigma as a whole (sum of z^2 over the bins,
chi-square p > 0.0027).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import golden_io as gio  # noqa: E402
import oracle  # noqa: E402
import qed_case  # noqa: E402
import quill_amd as qa  # noqa: E402


def gpu_summaries(precision):
    c = qed_case.inputs()
    h = np.frombuffer(gio.load("qed_stats")["header"].tobytes(), gio.HDR)[0]
    out = []
    for seed in range(1, qed_case.SEEDS + 1):
        with qa.Region(c["nx"], c["ny"], c["nz"], c["dx"], c["dx"], c["dx"], c["dt"], solver="fdtd", pusher="vay",
                       precision=precision, qed=True, e_s=qed_case.E_S, seed=seed, norm_n=float(h["norm_n"]),
                       norm_ef=float(h["norm_ef"]), norm_ep=float(h["norm_ep"])) as r:
            r.set_fields(c["E"], c["B"])
            r.add_records(c["rec"], c["ids"])
            r.init_boundaries()
            d = r.step(c["nsteps"], diag=True)[-1]
            ph, _ = r.particles(qa.PHOTON)
            assert d["n_ph"] == len(ph)
            out.append(qed_case.summary(d["n_ph"], d["n_p"], d["energy_ph"], ph[:, 6]))
    return np.array(out)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_qed_statistics_match_reference(precision):
    print(f"QED statistics, {precision}, {qed_case.SEEDS} seeds", flush=True)
    ref = gio.load("qed_stats")["summary"]
    mine = gpu_summaries(precision)
    z = qed_case.zscores(mine, ref)
    y = qed_case.YIELDS_STATS
    print("z: " + ", ".join(f"{n} {v:.2f}" for n, v in zip(qed_case.NAMES, z)), flush=True)
    bad = [(n, round(float(v), 2)) for n, v in zip(qed_case.NAMES[:y], z[:y]) if not abs(v) < 3]
    assert not bad, f"GPU QED yields differ from the reference: {bad}"
    p = qed_case.spectrum_chi2_p(z[y:])
    assert p > 0.0027, (p, np.round(z[y:], 2))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_qed_pairs_match_reference(precision):
    print(f"QED pair case, {precision}, {qed_case.SEEDS} seeds", flush=True)
    c = qed_case.inputs_pairs()
    fx = gio.load("qed_pairs")
    h = np.frombuffer(fx["header"].tobytes(), gio.HDR)[0]
    mine = []
    for seed in range(1, qed_case.SEEDS + 1):
        with qa.Region(c["nx"], c["ny"], c["nz"], c["dx"], c["dx"], c["dx"], c["dt"], solver="fdtd", pusher="vay",
                       precision=precision, qed=True, e_s=qed_case.E_S, seed=seed, norm_n=float(h["norm_n"]),
                       norm_ef=float(h["norm_ef"]), norm_ep=float(h["norm_ep"])) as r:
            r.set_fields(c["E"], c["B"])
            r.add_records(c["rec"], c["ids"])
            r.init_boundaries()
            d = r.step(c["nsteps"], diag=True)[-1]
            ph, _ = r.particles(qa.PHOTON)
            po, _ = r.particles(qa.POSITRON)
            mine.append(qed_case.summary_pairs(d["n_ph"], d["n_p"], d["energy_ph"], d["energy_p"], ph[:, 6],
                                               po[:, 6]))
    z = qed_case.zscores(np.array(mine), fx["summary"])
    y = qed_case.YIELDS
    print("yields z: " + ", ".join(f"{n} {v:.2f}" for n, v in zip(qed_case.NAMES_PAIRS[:y], z[:y])), flush=True)
    bad = [(n, round(float(v), 2)) for n, v in zip(qed_case.NAMES_PAIRS[:y], z[:y]) if not abs(v) < 3]
    assert not bad, f"GPU QED yields differ from the reference: {bad}"
    nb = len(qed_case.BINS) - 1
    for name, zz in (("photon", z[y:y + nb]), ("positron", z[y + nb:])):
        p = qed_case.spectrum_chi2_p(zz)
        print(f"{name} spectrum: chi-square p = {p:.3g}", flush=True)
        assert p > 0.0027, (name, p, np.round(zz, 2))


@pytest.mark.gpu
def test_vacuum_pair_rate():
    """birth_from_vacuum (src/mc.cpp:142-196): in a uniform E the number of
    pairs born in one step is binomial over the cells 2 .. n-3 with the
    per-cell probability W dt dV of the Schwinger rate (oracle rate, pinned
    to the reference by rates.npz): the mean over 8 seeds within 3 standard
    errors."""
    n = 24
    dx = 0.05 * 2 * np.pi
    dt = 0.025 * 2 * np.pi
    e_s, e0 = 1.0, 10.0
    E = np.zeros((n, n, n, 3))
    E[..., 0] = e0
    B = np.zeros_like(E)
    w = oracle.lib().oq_vacuum_rate(e_s, oracle._dp(np.array([e0, 0.0, 0.0])), oracle._dp(np.zeros(3)))
    p = w * dt * dx ** 3
    cells = (n - 4) ** 3
    counts = []
    for seed in range(1, 9):
        with qa.Region(n, n, n, dx, dx, dx, dt, solver="fdtd", pusher="boris", precision="fp32", qed=True, e_s=e_s,
                       seed=seed, q_vac=1e-3) as r:
            r.set_fields(E, B)
            r.step(1)
            counts.append((r.count(qa.ELECTRON), r.count(qa.POSITRON)))
    counts = np.array(counts, float)
    assert (counts[:, 0] == counts[:, 1]).all()   # pairs
    mean, sd = cells * p, np.sqrt(cells * p * (1 - p))
    z = (counts[:, 1].mean() - mean) / (sd / np.sqrt(len(counts)))
    print(f"vacuum pairs: mean {counts[:, 1].mean():.1f} expected {mean:.1f}, z {z:.2f}", flush=True)
    assert mean > 20 and abs(z) < 3, (counts[:, 1], mean, z)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_qed_cascade_matches_reference(precision):
    """The reduced colliding-pulse cascade (qed_case.inputs_cascade: 32^3
    cells, two circularly polarised pulses of peak |a| = 1000 from the
    device's f_init_cos, five-point + Vay, 200 steps, about 1600 pairs per
    run) on the GPU against 16 runs of the reference: electron, positron and
    photon yields and energies within 3 sigma (|z| < 3 each), the photon and
    positron gamma spectra each within 3 sigma as a whole (chi-square
    p > 0.0027)."""
    print(f"QED cascade case, {precision}, {qed_case.SEEDS} seeds", flush=True)
    c = qed_case.inputs_cascade()
    fx = gio.load("qed_cascade")
    h = np.frombuffer(fx["header"].tobytes(), gio.HDR)[0]
    n, d = c["nx"], c["dx"]
    mine = []
    for seed in range(1, qed_case.SEEDS + 1):
        with qa.Region(n, n, n, d, d, d, c["dt"], solver="five_point", pusher="vay", precision=precision, qed=True,
                       e_s=qed_case.E_S, seed=seed, norm_n=float(h["norm_n"]), norm_ef=float(h["norm_ef"]),
                       norm_ep=float(h["norm_ep"])) as r:
            for L in c["lasers"]:
                r.init_laser_cos(*L)
            r.add_records(c["rec"], c["ids"])
            r.init_boundaries()
            last = r.step(c["nsteps"], diag=True)[-1]
            ph, _ = r.particles(qa.PHOTON)
            po, _ = r.particles(qa.POSITRON)
            mine.append(qed_case.summary_cascade(last, ph[:, 6], po[:, 6]))
    z = qed_case.zscores(np.array(mine), fx["summary"])
    y = qed_case.YIELDS_CASCADE
    print("yields z: " + ", ".join(f"{nm} {v:.2f}" for nm, v in zip(qed_case.NAMES_CASCADE[:y], z[:y])), flush=True)
    bad = [(nm, round(float(v), 2)) for nm, v in zip(qed_case.NAMES_CASCADE[:y], z[:y]) if not abs(v) < 3]
    assert not bad, f"GPU cascade yields differ from the reference: {bad}"
    nb = len(qed_case.CASCADE_BINS) - 1
    for name, zz in (("photon", z[y:y + nb]), ("positron", z[y + nb:])):
        p = qed_case.spectrum_chi2_p(zz)
        print(f"{name} spectrum: chi-square p = {p:.3g}", flush=True)
        assert p > 0.0027, (name, p, np.round(zz, 2))
