"""Pin the FP64 oracle against golden vectors of the compiled reference.

The fixtures in tests/golden were produced by oracle/_ref/quill_ref, which is
the reference's own src/*.cpp hot-path units linked with a single-rank driver
(tests/golden/make_golden.py).  Every oracle function on the path is checked
here: pushers, rate functions, the laser initial condition, and full steps of
all three solvers (gather, interpolate_be, jdeposition, boundaries, place,
diagnostics) on a mixed population of electrons, positrons, ions and photons.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
sys.path.insert(0, os.path.dirname(HERE))

import golden_io as gio  # noqa: E402
import oracle  # noqa: E402

RUN_CASES = ["run_fdtd_boris", "run_ndfx_vay", "run_five_vay", "run_fdtd_vay_cold", "run_edges"]


def hdr(fx, key="header"):
    return np.frombuffer(fx[key].tobytes(), gio.HDR)[0]


def make_oracle(h):
    return oracle.Oracle(int(h["nx"]), int(h["ny"]), int(h["nz"]), float(h["dx"]), float(h["dy"]),
                         float(h["dz"]), float(h["dt"]), solver=int(h["solver"]), pusher=int(h["pusher"]),
                         qed=bool(h["qed"]), e_s=float(h["e_s"]), external_bz=float(h["external_bz"]),
                         q_vac=float(h["q_vac"]), seed=int(h["seed"]), norm_n=float(h["norm_n"]),
                         norm_ef=float(h["norm_ef"]), norm_ep=float(h["norm_ep"]))


def rel_err(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", RUN_CASES)
def test_run_matches_reference(name):
    fx = gio.load(name)
    h = hdr(fx)
    o = make_oracle(h)
    o.field(oracle.E)[:] = fx["E_in"]
    o.field(oracle.B)[:] = fx["B_in"]
    o.add_particles(fx["part_in"], fx["ids_in"])
    o.init_boundaries()
    diag = o.step(int(h["nsteps"]), diag=True)
    for key, which in (("E", oracle.E), ("B", oracle.B), ("J", oracle.J), ("Be", oracle.BE)):
        err = rel_err(o.field(which), fx[key])
        assert err < 1e-11, f"{name} {key}: rel err {err:.3e}"
    rec, ids = o.particles()
    order = np.argsort(ids, kind="stable")
    rec, ids = rec[order], ids[order]
    np.testing.assert_array_equal(ids, fx["ids"])
    np.testing.assert_allclose(rec[:, :3], fx["part"][:, :3], rtol=0, atol=1e-11)
    np.testing.assert_allclose(rec[:, 3:6], fx["part"][:, 3:6], rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(rec[:, 6], fx["part"][:, 6], rtol=1e-12)
    np.testing.assert_array_equal(rec[:, 7:9], fx["part"][:, 7:9])
    ref = fx["diag"]
    for l, d in enumerate(diag):
        mine = [d.N_e, d.N_p, d.N_ph, d.n_e, d.n_p, d.n_ph, d.energy_f, d.energy_e, d.energy_p,
                d.energy_ph, d.energy_ion, d.n_ion]
        np.testing.assert_allclose(mine, ref[l], rtol=1e-11, atol=0, err_msg=f"{name} step {l}")


@pytest.mark.parametrize("mode", ["on", "auto"])
def test_moving_window_matches_reference(mode):
    """moving_window (src/advance.cpp:389-454), get_max_w
    (src/low-level.cpp:512-532) and fill_cell_by_particles (src/init.cpp:402-478)
    under main()'s window logic, against the reference harness."""
    import mw_case
    fx = gio.load(f"mw_{mode}")
    h = hdr(fx)
    c = mw_case.inputs(mode)
    o = make_oracle(h)
    o.field(oracle.E)[:] = fx["E_in"]
    o.field(oracle.B)[:] = fx["B_in"]
    o.add_particles(fx["part_in"], fx["ids_in"])
    o.init_boundaries()
    diag, trace = mw_case.drive(o, c, lambda cells, dens, npic, first: o.fill_cells(-1.0, cells, dens, npic, 0.0,
                                                                                   first))
    np.testing.assert_array_equal(trace[:, 0], fx["trace"][:, 0])
    np.testing.assert_allclose(trace[:, 1], fx["trace"][:, 1], rtol=1e-12)
    assert fx["trace"][-1, 0] >= 4   # the window moved
    for key, which in (("E", oracle.E), ("B", oracle.B), ("J", oracle.J), ("Be", oracle.BE)):
        err = rel_err(o.field(which), fx[key])
        assert err < 1e-11, f"mw_{mode} {key}: rel err {err:.3e}"
    rec, ids = o.particles()
    order = np.argsort(ids, kind="stable")
    rec, ids = rec[order], ids[order]
    np.testing.assert_array_equal(ids, fx["ids"])
    assert (ids >= mw_case.ID_BASE).sum() > 100   # column fills survive
    np.testing.assert_allclose(rec[:, :3], fx["part"][:, :3], rtol=0, atol=1e-11)
    np.testing.assert_allclose(rec[:, 3:6], fx["part"][:, 3:6], rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(rec[:, 7], fx["part"][:, 7], rtol=1e-14)
    for l, d in enumerate(diag):
        mine = [d.N_e, d.N_p, d.N_ph, d.n_e, d.n_p, d.n_ph, d.energy_f, d.energy_e, d.energy_p,
                d.energy_ph, d.energy_ion, d.n_ion]
        np.testing.assert_allclose(mine, fx["diag"][l], rtol=1e-11, atol=0, err_msg=f"mw_{mode} step {l}")


def test_laser_matches_reference():
    fx = gio.load("laser_cos")
    for hk, ek, bk in (("header", "E", "B"), ("header2", "E2", "B2")):
        h = hdr(fx, hk)
        o = make_oracle(h)
        L = h["laser"]
        o.laser_cos(L[0], L[1], L[2], L[3], L[4], L[5], int(L[6]), L[7], L[8], L[9], L[10], bool(L[11]), L[12],
                    L[13], L[14], L[15])
        assert rel_err(o.field(oracle.E), fx[ek]) < 1e-12
        assert rel_err(o.field(oracle.B), fx[bk]) < 1e-12
        assert np.abs(fx[ek]).max() > 0.1


def test_rates_match_reference():
    fx = gio.load("rates")
    h = hdr(fx)
    e_s = float(h["e_s"])
    L = oracle.lib()
    for row, out in zip(fx["table"], fx["out"]):
        g, r, chi = row[0], row[1], row[2]
        e, b, u = (np.ascontiguousarray(row[3:6]), np.ascontiguousarray(row[6:9]),
                   np.ascontiguousarray(row[9:12]))
        c = L.oq_chi(e_s, oracle._dp(e), oracle._dp(b), oracle._dp(u), g)
        assert c == pytest.approx(out[0], rel=1e-12, abs=1e-300) or (np.isnan(c) and np.isnan(out[0]))
        rr = np.array([r])
        w = L.oq_rate_w(e_s, g, oracle._dp(rr), chi)
        assert w == pytest.approx(out[1], rel=1e-11, abs=0)
        assert rr[0] == pytest.approx(out[2], rel=1e-14, abs=0)
        tw = L.oq_rate_tilde_w(e_s, g, r, chi)
        assert tw == pytest.approx(out[3], rel=1e-11, abs=0)
        vw = L.oq_vacuum_rate(e_s, oracle._dp(e), oracle._dp(b))
        assert vw == pytest.approx(out[4], rel=1e-11, abs=0)
    # the table exercises the nonzero branches too
    assert (fx["out"][:, 1] != 0).sum() > 50 and (fx["out"][:, 3] != 0).sum() > 50
    assert (fx["out"][:, 4] > 0).sum() > 5


def test_pushers_match_reference():
    fx = gio.load("push")
    for row, out in zip(fx["table"], fx["out"]):
        cmr, dt, e, b, u, g = row[0], row[1], row[2:5], row[5:8], row[8:11], row[11]
        ub, gb = oracle.push(oracle.BORIS, cmr, dt, e, b, u, g)
        uv, gv = oracle.push(oracle.VAY, cmr, dt, e, b, u, g)
        np.testing.assert_allclose(np.r_[ub, gb], out[:4], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(np.r_[uv, gv], out[4:], rtol=1e-12, atol=1e-14)


def qed_oracle_summaries():
    import qed_case
    c = qed_case.inputs()
    h = hdr(gio.load("qed_stats"))
    out = []
    for seed in range(1, qed_case.SEEDS + 1):
        o = oracle.Oracle(c["nx"], c["ny"], c["nz"], c["dx"], c["dx"], c["dx"], c["dt"], solver=oracle.FDTD,
                          pusher=oracle.VAY, qed=True, e_s=qed_case.E_S, seed=seed, norm_n=float(h["norm_n"]),
                          norm_ef=float(h["norm_ef"]), norm_ep=float(h["norm_ep"]))
        o.field(oracle.E)[:] = c["E"]
        o.field(oracle.B)[:] = c["B"]
        o.add_particles(c["rec"], c["ids"])
        o.init_boundaries()
        d = o.step(c["nsteps"], diag=True)[-1]
        rec, _ = o.particles()
        out.append(qed_case.summary(d.n_ph, d.n_p, d.energy_ph, rec[rec[:, 8] == 0][:, 6]))
    return np.array(out)


def test_oracle_qed_statistics_match_reference():
    """Photon emission and pair creation of the oracle (its own random stream)
    against 16 runs of the reference (get_rand, src/mc.cpp:12-23): final
    photon and positron counts and the photon energy within 3 sigma (|z| < 3
    each), and the photon-gamma spectrum within 3 sigma as a whole (sum of
    z^2 over the bins, chi-square p > 0.0027)."""
    import qed_case
    ref = gio.load("qed_stats")["summary"]
    mine = qed_oracle_summaries()
    z = qed_case.zscores(mine, ref)
    assert ref[:, 0].mean() > 10000 and ref[:, 1].mean() > 10   # the case exercises both processes
    y = qed_case.YIELDS_STATS
    bad = [(n, round(float(v), 2)) for n, v in zip(qed_case.NAMES[:y], z[:y]) if not abs(v) < 3]
    assert not bad, f"oracle QED yields differ from the reference: {bad}"
    p = qed_case.spectrum_chi2_p(z[y:])
    assert p > 0.0027, (p, np.round(z[y:], 2))


def test_oracle_qed_pairs_match_reference():
    """The pair-rich case (about 1400 Breit-Wheeler pairs per run) of the
    oracle against 16 runs of the reference: photon and positron yields and
    energies within 3 sigma (|z| < 3 each), and the photon and positron
    gamma spectra each within 3 sigma as a whole (sum of z^2 over the bins,
    chi-square p > 0.0027)."""
    import qed_case
    c = qed_case.inputs_pairs()
    ref = gio.load("qed_pairs")["summary"]
    h = hdr(gio.load("qed_pairs"))
    mine = []
    for seed in range(1, qed_case.SEEDS + 1):
        o = oracle.Oracle(c["nx"], c["ny"], c["nz"], c["dx"], c["dx"], c["dx"], c["dt"], solver=oracle.FDTD,
                          pusher=oracle.VAY, qed=True, e_s=qed_case.E_S, seed=seed, norm_n=float(h["norm_n"]),
                          norm_ef=float(h["norm_ef"]), norm_ep=float(h["norm_ep"]))
        o.field(oracle.E)[:] = c["E"]
        o.field(oracle.B)[:] = c["B"]
        o.add_particles(c["rec"], c["ids"])
        o.init_boundaries()
        d = o.step(c["nsteps"], diag=True)[-1]
        rec, _ = o.particles()
        mine.append(qed_case.summary_pairs(d.n_ph, d.n_p, d.energy_ph, d.energy_p, rec[rec[:, 8] == 0][:, 6],
                                           rec[rec[:, 8] == 1][:, 6]))
    z = qed_case.zscores(np.array(mine), ref)
    assert ref[:, 1].mean() > 1000
    y = qed_case.YIELDS
    bad = [(n, round(float(v), 2)) for n, v in zip(qed_case.NAMES_PAIRS[:y], z[:y]) if not abs(v) < 3]
    assert not bad, f"oracle QED yields differ from the reference: {bad}"
    nb = len(qed_case.BINS) - 1
    for name, zz in (("photon", z[y:y + nb]), ("positron", z[y + nb:])):
        p = qed_case.spectrum_chi2_p(zz)
        assert p > 0.0027, (name, p, np.round(zz, 2))


@pytest.mark.parametrize("mode", ["ti", "nl"])
def test_oracle_merging_matches_reference(mode):
    """pmerging (src/mc.cpp:202-462) of the oracle against 16 merges of the
    reference with its own random stream: surviving counts and sum q^2 per
    species agree as distributions (|z| < 4); each species' total charge is
    kept to round-off, as the final rescale makes it."""
    import merge_case
    c = merge_case.inputs()
    ref = gio.load("merge_stats")[mode]
    mine = []
    for seed in range(1, merge_case.SEEDS + 1):
        o = oracle.Oracle(c["nx"], c["ny"], c["nz"], 0.3, 0.3, 0.3, 0.1, seed=seed)
        o.add_particles(c["rec"], c["ids"])
        o.pmerging(mode, merge_case.PPD[mode], [merge_case.ION_CMR])
        rec, _ = o.particles()
        mine.append(merge_case.summary(rec))
    mine = np.array(mine)
    z = merge_case.zscores(mine, ref)
    bad = [(n, round(float(v), 2)) for n, v in zip(merge_case.NAMES, z) if not n.endswith("_q") and not abs(v) < 4]
    assert not bad, f"oracle merging differs from the reference: {bad}"
    np.testing.assert_allclose(mine[:, 1::3], ref[:, 1::3], rtol=1e-12)
    assert (ref[:, 0] < len(c["rec"])).all()


def _oracle_cascade_run(seed):
    """One seed of the reduced cascade on the oracle (a worker process)."""
    import qed_case
    c = qed_case.inputs_cascade()
    h = hdr(gio.load("qed_cascade"))
    n, d = c["nx"], c["dx"]
    o = oracle.Oracle(n, n, n, d, d, d, c["dt"], solver=oracle.FIVE_POINT, pusher=oracle.VAY, qed=True,
                      e_s=qed_case.E_S, seed=seed, norm_n=float(h["norm_n"]), norm_ef=float(h["norm_ef"]),
                      norm_ep=float(h["norm_ep"]))
    for L in c["lasers"]:
        o.laser_cos(*L)
    o.add_particles(c["rec"], c["ids"])
    o.init_boundaries()
    last = o.step(c["nsteps"], diag=True)[-1]
    rec, _ = o.particles()
    dg = {k: getattr(last, k) for k in ("n_e", "n_p", "n_ph", "energy_e", "energy_p", "energy_ph")}
    return qed_case.summary_cascade(dg, rec[rec[:, 8] == 0][:, 6], rec[rec[:, 8] == 1][:, 6])


def test_oracle_qed_cascade_matches_reference():
    """The reduced colliding-pulse cascade (32^3 cells, two circularly
    polarised pulses of peak |a| = 1000, five-point + Vay, 200 steps, about
    1600 pairs and 3e4 photons per run) of the oracle against 16 runs of the
    reference: electron, positron and photon yields and energies within 3
    sigma (|z| < 3 each), the photon and positron gamma spectra each within 3
    sigma as a whole (chi-square p > 0.0027)."""
    from concurrent.futures import ProcessPoolExecutor
    import qed_case
    ref = gio.load("qed_cascade")["summary"]
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        mine = np.array(list(ex.map(_oracle_cascade_run, range(1, qed_case.SEEDS + 1))))
    assert ref[:, 1].mean() > 1000 and ref[:, 2].mean() > 1e4
    z = qed_case.zscores(mine, ref)
    y = qed_case.YIELDS_CASCADE
    bad = [(n, round(float(v), 2)) for n, v in zip(qed_case.NAMES_CASCADE[:y], z[:y]) if not abs(v) < 3]
    assert not bad, f"oracle cascade yields differ from the reference: {bad}"
    nb = len(qed_case.CASCADE_BINS) - 1
    for name, zz in (("photon", z[y:y + nb]), ("positron", z[y + nb:])):
        p = qed_case.spectrum_chi2_p(zz)
        assert p > 0.0027, (name, p, np.round(zz, 2))
