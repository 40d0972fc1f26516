"""quillk-run on the GPU against the oracle on the same set-up.

The driver sets up a cold film and a cos laser from a .conf file, runs the
step loop on libquillk.so and writes the reference's N / energy records
(write_N / write_energy, src/main.cpp:160-187).  The same set-up, dumped by
the driver (--dump-setup), is run through the FP64 oracle; the per-step
records agree to the FP32 bar.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "quill_amd", "bin", "quillk-run")

CONF = """\
dx = 0.05
dy = 0.1
dz = 0.1
dt = 0.025
lambda = 1 um
ne = 0.01 ncr
xlength = 2.4001
ylength = 2.4001
zlength = 2.4001
t_end = 0.5
output_period = 0.25
f_envelope = cos
polarization = circular
a0 = 2
xsigma = 0.5
ysigma = 0.6
zsigma = 0.6
x0 = 0.3
y0 = 0.1
mwindow = off
qed = off
solver = {solver}
pusher = {pusher}
ions = on
film = on
x0film = 0.6
filmwidth = 0.6
gradwidth = 0.2
nfilm = 0.5 ncr
mcr = 2
xnpic = 2
ynpic = 1
znpic = 2
"""


@pytest.mark.gpu
@pytest.mark.parametrize("solver, pusher", [("fdtd", "boris"), ("ndfx", "vay")])
def test_driver_run_matches_oracle(tmp_path, solver, pusher):
    if not os.path.exists(BIN):
        pytest.fail("quillk-run not built")
    conf = tmp_path / "run.conf"
    conf.write_text(CONF.format(solver=solver, pusher=pusher))
    setup, out = tmp_path / "setup", tmp_path / "out"
    dry = subprocess.run([BIN, str(conf), "--dry-run"], capture_output=True, text=True, timeout=60)
    assert dry.returncode == 0, dry.stderr
    j = json.loads(dry.stdout)
    print(f"driver case {solver}/{pusher}: {j['nx']}x{j['ny']}x{j['nz']}, {j['steps']} steps", flush=True)
    r = subprocess.run([BIN, str(conf), "--dump-setup", str(setup), "--out", str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    steps = j["steps"]

    o = oracle.Oracle(j["nx"], j["ny"], j["nz"], j["dx"], j["dy"], j["dz"], j["dt"],
                      solver={"fdtd": oracle.FDTD, "ndfx": oracle.NDFX}[solver],
                      pusher={"boris": oracle.BORIS, "vay": oracle.VAY}[pusher],
                      norm_n=j["norm_n"], norm_ef=j["norm_ef"], norm_ep=j["norm_ep"])
    lasers = np.fromfile(setup / "lasers.f64").reshape(-1, 16)
    assert len(lasers) == 1
    for L in lasers:
        o.laser_cos(*L[:6], b_sign=L[6], x1=L[7], phase=L[8], y0=L[9], z0=L[10], append=bool(L[11]), phi=L[12],
                    xtarget=L[13], ytarget=L[14], ztarget=L[15])
    cmr = {0: -1.0, 3: j["ion_cmr"][0]}
    n_in = 0
    for slot, c in cmr.items():
        rec = np.fromfile(setup / f"species{slot}.f64").reshape(-1, 7)
        assert len(rec) > 1000
        n_in += len(rec)
        o.add_particles(np.c_[rec, np.full(len(rec), c)])
    o.init_boundaries()
    ref = o.step(steps, diag=True)

    N = np.loadtxt(out / "N").reshape(-1, 3)
    E = np.loadtxt(out / "energy").reshape(-1, 5)
    assert N.shape[0] == steps and E.shape[0] == steps
    N_ref = np.array([[d.N_e, d.N_p, d.N_ph] for d in ref])
    E_ref = np.array([[d.energy_f, d.energy_e, d.energy_p, d.energy_ph, d.energy_ion] for d in ref])
    # records are printed with 6 significant digits (ostream default)
    np.testing.assert_allclose(N, N_ref, rtol=2e-5)
    assert E_ref[:, 0].min() > 0 and E_ref[-1, 1] > 0
    np.testing.assert_allclose(E[:, 0], E_ref[:, 0], rtol=1e-4)
    np.testing.assert_allclose(E[:, 1], E_ref[:, 1], rtol=1e-3, atol=1e-4 * E_ref[:, 1].max())
    np.testing.assert_allclose(E[:, 4], E_ref[:, 4], rtol=1e-3, atol=1e-4 * E_ref[:, 4].max())

    # output iterations t = 0, 0.25, 0.5 lambda/c: plane files w, inv, rho
    # (the formats themselves are checked against the reference's writers in
    # test_output_writers.py); rho of the last one against the oracle
    nx, ny, nz = j["nx"], j["ny"], j["nz"]
    lines = nx * (ny + nz) + ny * nz
    for t in ("0", "0.25", "0.5"):
        for name in ("w", "inv", "rho"):
            path = out / (name + t)
            assert path.exists(), path
            assert len(path.read_text().splitlines()) == lines
    rho_gpu = np.loadtxt(out / "rho0.5")
    rho_ref = o.compute_rho()[..., 0]
    iyz = int((float(np.float32(2.4001)) * 2 * np.pi / 2) / j["dx"])
    expect = np.concatenate([np.concatenate([rho_ref[i, :, nz // 2], rho_ref[i, ny // 2, :]]) for i in range(nx)] +
                            [rho_ref[iyz].ravel()])
    np.testing.assert_allclose(rho_gpu, expect, rtol=2e-5, atol=1e-4 * np.abs(expect).max())


MW_CONF = """\
dx = 0.05
dy = 0.1
dz = 0.1
dt = 0.025
lambda = 1 um
ne = 0.05 ncr
xlength = 2.4001
ylength = 1.2001
zlength = 1.2001
t_end = 0.5
t_add_mw = 0.4
output_period = 0.5
f_envelope = cos
polarization = linear
a0 = 1
xsigma = 0.5
ysigma = 0.4
zsigma = 0.4
x0 = 0.4
mwindow = on
ne_profile_x_coords = [0;1]
ne_profile_x_values = [1;2]
qed = off
solver = ndfx
pusher = vay
xnpic = 1
ynpic = 2
znpic = 2
"""


@pytest.mark.gpu
def test_driver_moving_window_matches_oracle(tmp_path):
    """mwindow = on with a seeded plasma column (src/main.cpp:1750-1786) and
    no initial plasma: the driver's N, energy and mwcoordinate records
    against the oracle stepped with the same shifts (check_moving_window,
    src/main.cpp:1958-1960) and the column fills the driver reports
    (--mw-fills; their cells and densities are pinned against the reference
    in test_driver.py and test_driver_films.py)."""
    if not os.path.exists(BIN):
        pytest.fail("quillk-run not built")
    conf = tmp_path / "mw.conf"
    conf.write_text(MW_CONF)
    j = json.loads(subprocess.run([BIN, str(conf), "--dry-run"], capture_output=True, text=True, timeout=60).stdout)
    print(f"driver moving window: {j['nx']}x{j['ny']}x{j['nz']}, {j['steps']} steps", flush=True)
    out = tmp_path / "out"
    r = subprocess.run([BIN, str(conf), "--out", str(out), "--no-output"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    steps, dx, dt = j["steps"], j["dx"], j["dt"]
    o = oracle.Oracle(j["nx"], j["ny"], j["nz"], dx, j["dy"], j["dz"], dt, solver=oracle.NDFX, pusher=oracle.VAY,
                      norm_n=j["norm_n"], norm_ef=j["norm_ef"], norm_ep=j["norm_ep"])
    for L in j["lasers"]:
        o.laser_cos(*L[:6], b_sign=L[6], x1=L[7], phase=L[8], y0=L[9], z0=L[10], append=bool(L[11]), phi=L[12],
                    xtarget=L[13], ytarget=L[14], ztarget=L[15])
    o.init_boundaries()
    nmw, ref, mwc, fills = 1, [], [], 0
    for l in range(steps):
        ref.append(o.step(1, diag=True)[0])
        moves = (l + 1) * dt * j["mwspeed"] > nmw * dx
        if moves:
            o.moving_window()
            nmw += 1
        mwc.append((nmw - 1) * dx / 2 / np.pi)
        if moves and (l + 1) * dt < j["t_add_mw"]:
            col = json.loads(subprocess.run([BIN, str(conf), "--mw-fills", str(nmw)], capture_output=True, text=True,
                                            timeout=60).stdout)
            for f in col:
                assert f["species"] == 0
                o.fill_cells(-1.0, f["cells"], f["density"], f["npic"], f["ux0"])
            fills += 1
    assert nmw > 8 and fills > 5
    N = np.loadtxt(out / "N").reshape(-1, 3)
    E = np.loadtxt(out / "energy").reshape(-1, 4)
    W = np.loadtxt(out / "mwcoordinate")
    assert N.shape[0] == steps and E.shape[0] == steps and W.shape[0] == steps
    np.testing.assert_allclose(W, mwc, rtol=1e-5, atol=1e-12)
    N_ref = np.array([[d.N_e, d.N_p, d.N_ph] for d in ref])
    assert N_ref[-1, 0] > 0
    np.testing.assert_allclose(N, N_ref, rtol=2e-5)
    E_ref = np.array([[d.energy_f, d.energy_e, d.energy_p, d.energy_ph] for d in ref])
    np.testing.assert_allclose(E[:, 0], E_ref[:, 0], rtol=1e-4)
    np.testing.assert_allclose(E[:, 1], E_ref[:, 1], rtol=1e-3, atol=1e-4 * E_ref[:, 1].max())
