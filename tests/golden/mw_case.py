"""The moving-window cases shared by make_golden.py and the tests.

A mixed population (electrons, positrons, one ion species, photons) in
random fields, stepped with a moving window the way main() drives it for
n_sr = 1 (src/main.cpp:2333-2336, 2396-2398): after each step
check_moving_window (src/main.cpp:1953-1978) decides, moving_window()
(src/advance.cpp:389-454) shifts, and while (l + 1) dt < t_add_mw a column of
fill_cell_by_particles (src/init.cpp:402-478) at i = nx - 3 adds electrons
(add_moving_window_particles, src/main.cpp:1750-1786, uniform density).

"on" moves with mwspeed = 1 (a shift every 1 / 0.45 steps on average);
"auto" moves when the front slab [nx - nm - 2, nx - nm) holds more than
mwtolerance of the largest E^2 + B^2 (a field bump in the rear half spreads
forward).  drive() runs either on any object with step / moving_window /
max_w and a fill callback, so the reference harness (mode 7), the oracle and
the GPU run the same host logic.
"""
import numpy as np

NM = 4                      # nm, src/main.cpp:2232
ID_BASE = 1 << 24           # ids of particles added by the window
PROTON_MASS = 1836.1526721
ION_CMR = 1.0 / PROTON_MASS
MODES = {"on": 1, "auto": 2}


def inputs(mode):
    nx, ny, nz = 20, 10, 10
    dx = 0.05 * 2 * np.pi
    dt = 0.45 * dx
    rng = np.random.default_rng(7100 + MODES[mode])
    if mode == "on":
        E = rng.uniform(-0.4, 0.4, (nx, ny, nz, 3))
        B = rng.uniform(-0.4, 0.4, (nx, ny, nz, 3))
        tol = 0.0
    else:
        env = np.exp(-((np.arange(nx) - 6.0) / 2.5) ** 2)[:, None, None, None]
        E = rng.uniform(-0.4, 0.4, (nx, ny, nz, 3)) * env
        B = rng.uniform(-0.4, 0.4, (nx, ny, nz, 3)) * env
        tol = 0.005
    recs = []
    for cmr, n, q in ((-1.0, 900, -0.2), (1.0, 60, 0.2), (ION_CMR, 60, 0.25), (0.0, 40, 0.1)):
        r = np.zeros((n, 8))
        r[:, 0] = rng.uniform(1.0, nx - 2.0, n)
        r[:, 1] = rng.uniform(1.0, ny - 2.0, n)
        r[:, 2] = rng.uniform(1.0, nz - 2.0, n)
        r[:, 3:6] = rng.normal(0, 0.3, (n, 3))
        r[:, 6] = q
        r[:, 7] = cmr
        recs.append(r)
    rec = np.concatenate(recs)
    ids = np.arange(1, len(rec) + 1, dtype=np.int32)
    return dict(mode=mode, mw=MODES[mode], nx=nx, ny=ny, nz=nz, dx=dx, dt=dt, E=E, B=B, rec=rec, ids=ids,
                mwspeed=1.0, tol=tol, density=0.3, npic=np.array([2, 1, 2], np.int32), t_add=10 * dt,
                nsteps=14)


def column(c, fills):
    """Cells (nx - 3, j, k), j < ny, k < nz (in the reference's loop order),
    their densities and the first id of the fills-th fill."""
    ny, nz = c["ny"], c["nz"]
    j, k = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([np.full(ny * nz, c["nx"] - 3), j.ravel(), k.ravel()], 1).astype(np.int32)
    npt = int(np.prod(c["npic"]))
    return cells, np.full(ny * nz, c["density"]), ID_BASE + fills * ny * nz * npt


def drive(r, c, fill):
    """Step r c["nsteps"] times with the window; fill(cells, density, npic,
    first_id) adds electrons.  Returns the per-step diagnostics and the
    per-step (nmw, ratio) trace."""
    nx, dx, dt = c["nx"], c["dx"], c["dt"]
    nmw, fills = 1, 0
    diags, trace = [], []
    for l in range(c["nsteps"]):
        diags.append(r.step(1, diag=True)[0])
        ratio, shift = 0.0, False
        if c["mw"] == 1:
            shift = (l + 1) * dt * c["mwspeed"] > nmw * dx
        else:
            mx = r.max_w(0, nx)
            ratio = r.max_w(nx - NM - 2, nx - NM) / mx if mx > 0 else 0.0
            shift = ratio > c["tol"]
        if shift:
            r.moving_window()
            nmw += 1
        trace.append((nmw, ratio))
        if shift and (l + 1) * dt < c["t_add"] and c["density"] > 0:
            cells, dens, first = column(c, fills)
            fill(cells, dens, c["npic"], first)
            fills += 1
    return diags, np.array(trace)
