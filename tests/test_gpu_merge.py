"""Macro-particle merging on the GPU (pmerging, src/mc.cpp:202-462).

The deletion is random, so parity is statistical: 16 seeded GPU merges of
the case in tests/golden/merge_case.py against the 16 merges of the
reference in merge_stats.npz (made by the reference's own pmerging with its
own random stream): surviving counts and sum q^2 per species within |z| < 4,
and each species' total charge kept (the final rescale).  The trigger of
evaluate_merging_condition (src/main.cpp:1691-1748) is checked on a run.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import golden_io as gio  # noqa: E402
import merge_case  # noqa: E402
import quill_amd as qa  # noqa: E402


def region(c, precision, seed, **kw):
    return qa.Region(c["nx"], c["ny"], c["nz"], 0.3, 0.3, 0.3, 0.1, precision=precision, seed=seed,
                     ion_cmr=[merge_case.ION_CMR], **kw)


def all_particles(r):
    recs = [r.particles(s)[0] for s in (qa.ELECTRON, qa.POSITRON, qa.PHOTON, qa.ION0)]
    return np.concatenate([x for x in recs if len(x)])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("mode", ["ti", "nl"])
def test_merging_matches_reference(mode, precision):
    print(f"merging {mode} {precision}, {merge_case.SEEDS} seeds", flush=True)
    c = merge_case.inputs()
    ref = gio.load("merge_stats")[mode]
    mine = []
    for seed in range(1, merge_case.SEEDS + 1):
        with region(c, precision, seed) as r:
            r.add_records(c["rec"])
            r.pmerging(mode, merge_case.PPD[mode])
            mine.append(merge_case.summary(all_particles(r)))
    mine = np.array(mine)
    z = merge_case.zscores(mine, ref)
    bad = [(n, round(float(v), 2)) for n, v in zip(merge_case.NAMES, z) if not n.endswith("_q") and not abs(v) < 4]
    assert not bad, f"GPU merging differs from the reference: {bad}"
    np.testing.assert_allclose(mine[:, 1::3], ref[:, 1::3], rtol=1e-5 if precision == "fp32" else 1e-12)


@pytest.mark.gpu
def test_merging_trigger_ti():
    """"ti": after a step with N macro-particles above (3 + n_ion) crnp, the
    next step starts with a merge that deletes the fraction
    (N - (3 + n_ion) crnp) / N (massive particles at rest in zero fields;
    photons fly along +x by 1/3 cell per step and stay inside)."""
    c = merge_case.inputs()
    rec = c["rec"].copy()
    rec[:, 3:6] = 0.0
    rec[rec[:, 7] == 0, 3] = 1e-3
    crnp = 2000.0
    with region(c, "fp32", 7, merging="ti", crnp=crnp) as r:
        r.add_records(rec)
        d = r.step(2, diag=True)
    n0 = d[0]["n_e"] + d[0]["n_p"] + d[0]["n_ph"] + d[0]["n_ion"]
    n1 = d[1]["n_e"] + d[1]["n_p"] + d[1]["n_ph"] + d[1]["n_ion"]
    assert n0 == len(rec)
    keep = 4 * crnp / n0
    sd = np.sqrt(n0 * keep * (1 - keep))
    assert abs(n1 - 4 * crnp) < 5 * sd, (n0, n1)
    # charge per species is kept: compute_N is the species' total charge
    for k in ("N_e", "N_p", "N_ph"):
        assert d[1][k] == pytest.approx(d[0][k], rel=1e-5)
