"""The particle-merging statistics case shared by make_golden.py and the tests.

A mixed population (electrons, positrons, photons, one ion species) with
log-normally spread charges, merged once with the "ti" and the "nl"
algorithm of pmerging (src/mc.cpp:202-462).  The reference merges it with
its own random stream for 16 seeds (merge_stats.npz); the oracle and the GPU
merge it with theirs, and the per-species summaries are compared as
distributions.
"""
import numpy as np

SEEDS = 16
PROTON_MASS = 1836.1526721
ION_CMR = 1.0 / PROTON_MASS
PPD = {"ti": [0.3, 0.0, 0.0, 0.0], "nl": [0.45, 0.25, 0.6, 0.35]}
N = {-1.0: 6000, 1.0: 2500, 0.0: 2500, ION_CMR: 1500}


def inputs():
    nx = ny = nz = 12
    rng = np.random.default_rng(4242)
    recs = []
    for cmr, n in N.items():
        r = np.zeros((n, 8))
        r[:, 0:3] = rng.uniform(1.0, 9.0, (n, 3))
        r[:, 3:6] = rng.normal(0, 0.5, (n, 3))
        q = np.exp(rng.normal(np.log(1e-3), 0.8, n))
        r[:, 6] = -q if cmr == -1.0 else q
        r[:, 7] = cmr
        recs.append(r)
    rec = np.concatenate(recs)
    return dict(nx=nx, ny=ny, nz=nz, rec=rec, ids=np.arange(len(rec), dtype=np.int32))


SPECIES = [-1.0, 1.0, 0.0, ION_CMR]
NAMES = [f"{s}_{k}" for s in ("e", "p", "ph", "ion") for k in ("n", "q", "q2")]


def summary(rec):
    """Per species: surviving count, total |q|, sum of q^2 (rec rows: x y z
    ux uy uz g q cmr ...)."""
    out = []
    for cmr in SPECIES:
        q = np.abs(rec[rec[:, 8] == cmr][:, 7])
        out += [float(len(q)), float(q.sum()), float((q ** 2).sum())]
    return np.array(out)


def zscores(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    return (a.mean(0) - b.mean(0)) / np.where(se > 0, se, 1.0)
