"""Generate the golden fixtures of tests/golden from the compiled reference.

Runs oracle/_ref/quill_ref (the reference's own hot-path translation units,
built by oracle/Makefile from the reference tree) on small seeded cases and
stores inputs plus the reference's outputs as .npz files next to this script.
Run in the build container only (the reference tree is not on the GPU host):

    make -C oracle && python tests/golden/make_golden.py [case ...]

Cases
  run_*      full steps (src/main.cpp:2319-2332 order) from random fields and a
             mixed particle population (electrons, positrons, ions, photons;
             QED off); outputs E, B, J, Be, particles (by id) and per-step
             diagnostics
  run_edges  the same from zero fields with hand-placed particles on the
             deposit's edge cases (tests/golden/edge_case.py)
  laser_cos  f_init_cos (src/init.cpp:6-210), sscos = 0, tilted and rotated
  rates      chi / w / tilde_w / mathcal_W (src/mc.cpp:26-116) on a table
  push       push_boris / push_vay (src/pusher.cpp) on a table
  fout       the reference's plane writers (ex .. bz, w, inv, rho, rho_p,
             rho_ph; src/fout.cpp:180-660) in text and binary output mode
  merge_stats  pmerging "ti" and "nl" (src/mc.cpp:202-462) over 16 seeds of the
             reference's random stream (statistical parity)
  mw_on, mw_auto  a moving window (src/advance.cpp:389-454) with "on" and
             "auto" triggers and column fills (tests/golden/mw_case.py)
  film       film() on the films quillk-run derives from the texts of
             tests/golden/film_case.py, initial and moving-window columns
  qed_stats  QED on (photon emission, Breit-Wheeler): per-step diagnostics and
             the photon-gamma histogram of 16 runs with distinct seeds of the
             reference's own random stream (statistical parity)
  qed_pairs  a harder QED case (about 1400 pairs per run): yields and the
             photon and positron spectra of 16 reference runs
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import golden_io as gio  # noqa: E402
import film_case  # noqa: E402
import edge_case  # noqa: E402
import merge_case  # noqa: E402
import mw_case  # noqa: E402
import qed_case  # noqa: E402
import oracle  # noqa: E402

PI = np.pi
LAMBDA_CM = 1e-4
E_S = LAMBDA_CM / 2.4263086e-10   # src/main.cpp:2276 (Sauter field for lambda = 1 um)
PROTON_MASS = 1836.1526721        # include/physical_constants.h:4


def norms(dx, dy, dz):
    """compute_N / compute_energy multipliers of src/main.cpp:2326-2327."""
    v = dx * dy * dz
    return dict(norm_n=v * 1.11485e13 * LAMBDA_CM / (8 * PI ** 3),
                norm_ef=0.5 * v * 3.691e4 * LAMBDA_CM / 1e7,
                norm_ep=8.2e-14 * v * 1.11485e13 * LAMBDA_CM / (8 * PI ** 3))


def run_ref(h, *arrays):
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        gio.write_input(fi, h, *arrays)
        subprocess.run([oracle.REF_BIN, fi, fo], check=True, stdout=subprocess.DEVNULL)
        if int(h["mode"]) == gio.MODE_RUN:
            return gio.read_output_run(fo, h)
        if int(h["mode"]) == gio.MODE_LASER:
            return gio.read_output_fields(fo, h)
        width = 5 if int(h["mode"]) == gio.MODE_RATES else 8
        return gio.read_output_table(fo, width)


def mixed_particles(nx, ny, nz, seed, sigma, ppc=2):
    """Electrons from the shared loader plus positrons, ions and photons."""
    rec, ids = oracle.synth_plasma(nx, ny, nz, ppc, 0.5, sigma, seed)
    rng = np.random.default_rng(seed + 17)
    extra = []
    icmr = 1.0 / PROTON_MASS
    for cmr, count in ((1.0, 40), (icmr, 40), (0.0, 30)):
        pick = rng.choice(len(rec), count, replace=False)
        r = rec[pick].copy()
        r[:, 3:6] = rng.normal(0, sigma, (count, 3))
        r[:, 6] = 0.25 if cmr != 0 else 0.1
        r[:, 7] = cmr
        extra.append(r)
    ex = np.concatenate(extra)
    ex_ids = np.arange(len(ex), dtype=np.int32) + 10_000_000
    return np.concatenate([rec, ex]), np.concatenate([ids, ex_ids]), icmr


def case_run(name, nx, ny, nz, dx, dy, dz, dt, solver, pusher, nsteps, seed, sigma, amp,
             external_bz=0.0, fields=None):
    rng = np.random.default_rng(seed)
    shape = (nx, ny, nz, 3)
    if fields is None:
        E = rng.uniform(-amp, amp, shape)
        B = rng.uniform(-amp, amp, shape)
    else:
        E, B = fields
    rec, ids, icmr = mixed_particles(nx, ny, nz, seed, sigma)
    h = gio.header(mode=gio.MODE_RUN, nx=nx, ny=ny, nz=nz, solver=solver, pusher=pusher, qed=0,
                   nsteps=nsteps, n_particles=len(rec), seed=seed, n_ion=1, dx=dx, dy=dy, dz=dz, dt=dt,
                   e_s=E_S, external_bz=external_bz, **norms(dx, dy, dz))
    h["ion_cmr"][0] = icmr
    (Eo, Bo, Jo, BEo), prec, pids, diag = run_ref(h, E, B, rec, ids.astype(np.int32))
    order = np.argsort(pids, kind="stable")
    np.savez_compressed(os.path.join(HERE, name + ".npz"), header=h.tobytes(), E_in=E, B_in=B, part_in=rec,
                        ids_in=ids, E=Eo, B=Bo, J=Jo, Be=BEo, part=prec[order], ids=pids[order], diag=diag)
    print(f"{name}: {len(rec)} particles in, {len(prec)} out, {nsteps} steps")


def case_edges():
    """Hand-placed particles on the deposit's edge cases (edge_case.py): ties
    between crossing fractions, starts on a face, ends within rounding of
    one, exits through every side, three crossings, one crowded cell."""
    nx, ny, nz = edge_case.SHAPE
    rec, ids, icmr = edge_case.particles()
    zero = np.zeros((nx, ny, nz, 3))
    dx = edge_case.DX
    h = gio.header(mode=gio.MODE_RUN, nx=nx, ny=ny, nz=nz, solver=oracle.FDTD, pusher=oracle.BORIS, qed=0,
                   nsteps=edge_case.NSTEPS, n_particles=len(rec), seed=1, n_ion=1, dx=dx, dy=dx, dz=dx,
                   dt=edge_case.DT, e_s=E_S, external_bz=0.0, **norms(dx, dx, dx))
    h["ion_cmr"][0] = icmr
    (Eo, Bo, Jo, BEo), prec, pids, diag = run_ref(h, zero, zero, rec, ids)
    order = np.argsort(pids, kind="stable")
    np.savez_compressed(os.path.join(HERE, "run_edges.npz"), header=h.tobytes(), E_in=zero, B_in=zero, part_in=rec,
                        ids_in=ids, E=Eo, B=Bo, J=Jo, Be=BEo, part=prec[order], ids=pids[order], diag=diag)
    print(f"run_edges: {len(rec)} particles in, {len(prec)} out, {edge_case.NSTEPS} steps")


def case_laser():
    nx, ny, nz = 24, 16, 16
    dx, dy, dz = 0.1 * PI, 0.12 * PI, 0.1 * PI
    L = np.zeros(16)
    # a0y a0z xsigma ysigma zsigma x0 b_sign x1 phase y0 z0 append phi xt yt zt
    L[:] = [1.0, 0.3, 0.8 * PI, 1.0 * PI, 1.2 * PI, 12 * dx - 1.0, 1, -1.0, 0.3, 0.2, -0.1, 1, 0.4, 0.0, 0.0, 0.0]
    h = gio.header(mode=gio.MODE_LASER, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, dt=0.5 * dx, laser=L)
    E, B = run_ref(h)
    L2 = L.copy()
    L2[6] = 0       # b_sign = -1 branch
    L2[11] = 1
    L2[12] = 0.0    # no rotation
    L2[9] = L2[10] = 0.0
    h2 = gio.header(mode=gio.MODE_LASER, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, dt=0.5 * dx, laser=L2)
    E2, B2 = run_ref(h2)
    np.savez_compressed(os.path.join(HERE, "laser_cos.npz"), header=h.tobytes(), header2=h2.tobytes(),
                        E=E, B=B, E2=E2, B2=B2)
    print("laser_cos: done")


def case_rates():
    rng = np.random.default_rng(7)
    n = 240
    g = 10 ** rng.uniform(0, 4.5, n)
    r = rng.uniform(0, 1, n)
    chi = 10 ** rng.uniform(-3, 1.3, n)
    e = rng.normal(0, 1, (n, 3)) * 10 ** rng.uniform(-1, 5.5, (n, 1))
    b = rng.normal(0, 1, (n, 3)) * 10 ** rng.uniform(-1, 5.5, (n, 1))
    u = rng.normal(0, 1, (n, 3)) * g[:, None] / np.sqrt(3)
    # special rows: r = 0, chi = 0, (1-r) g < 1, r g < 1, chi near the 0.133 switch, B = 0, E || B
    g[:8] = [10, 10, 1.5, 1.5, 100, 100, 50, 50]
    r[:8] = [0, 0.5, 0.6, 0.5, 0.3, 0.3, 0.2, 0.9]
    chi[:8] = [0.5, 0, 0.5, 0.5, 0.1329, 0.1331, 0.01, 2.0]
    b[8] = 0
    b[9] = e[9] * 0.5
    e[10] = 0
    tab = np.column_stack([g, r, chi, e, b, u])
    h = gio.header(mode=gio.MODE_RATES, nx=4, ny=4, nz=4, dx=0.1, dy=0.1, dz=0.1, dt=0.05, e_s=E_S,
                   n_particles=n)
    out = run_ref(h, tab)
    np.savez_compressed(os.path.join(HERE, "rates.npz"), header=h.tobytes(), table=tab, out=out)
    print("rates: done")


def case_push():
    rng = np.random.default_rng(11)
    n = 240
    cmr = rng.choice([-1.0, 1.0, 1 / PROTON_MASS, -0.5], n)
    dt = rng.uniform(0.01, 0.5, n)
    e = rng.normal(0, 1, (n, 3)) * 10 ** rng.uniform(-2, 3, (n, 1))
    b = rng.normal(0, 1, (n, 3)) * 10 ** rng.uniform(-2, 3, (n, 1))
    u = rng.normal(0, 1, (n, 3)) * 10 ** rng.uniform(-3, 3, (n, 1))
    g = np.sqrt(1 + (u ** 2).sum(1))
    tab = np.column_stack([cmr, dt, e, b, u, g])
    h = gio.header(mode=gio.MODE_PUSH, nx=4, ny=4, nz=4, n_particles=n)
    out = run_ref(h, tab)
    np.savez_compressed(os.path.join(HERE, "push.npz"), header=h.tobytes(), table=tab, out=out)
    print("push: done")


def case_qed_stats():
    """QED statistics of the reference over qed_case.SEEDS seeds of its own
    random stream: per-step diagnostics and the run summaries."""
    c = qed_case.inputs()
    diags, summ = [], []
    for seed in range(1, qed_case.SEEDS + 1):
        h = gio.header(mode=gio.MODE_RUN, nx=c["nx"], ny=c["ny"], nz=c["nz"], solver=oracle.FDTD,
                       pusher=oracle.VAY, qed=1, nsteps=c["nsteps"], n_particles=len(c["rec"]), seed=seed, n_ion=0,
                       dx=c["dx"], dy=c["dx"], dz=c["dx"], dt=c["dt"], e_s=qed_case.E_S,
                       **norms(c["dx"], c["dx"], c["dx"]))
        _, prec, _, diag = run_ref(h, c["E"], c["B"], c["rec"], c["ids"])
        ph = prec[prec[:, 8] == 0]
        # diag row: N_e N_p N_ph n_e n_p n_ph energy_f energy_e energy_p energy_ph ...
        diags.append(diag)
        summ.append(qed_case.summary(diag[-1][5], diag[-1][4], diag[-1][9], ph[:, 6]))
    np.savez_compressed(os.path.join(HERE, "qed_stats.npz"), header=h.tobytes(), diag=np.array(diags),
                        summary=np.array(summ))
    s = np.array(summ)
    print(f"qed_stats: {qed_case.SEEDS} seeds, final photons {s[:, 0].mean():.0f}, positrons {s[:, 1].mean():.1f}")


def case_qed_pairs():
    """The pair-rich QED case (qed_case.inputs_pairs) over qed_case.SEEDS
    seeds of the reference's own random stream: yields and the photon and
    positron spectra."""
    c = qed_case.inputs_pairs()
    summ = []
    for seed in range(1, qed_case.SEEDS + 1):
        h = gio.header(mode=gio.MODE_RUN, nx=c["nx"], ny=c["ny"], nz=c["nz"], solver=oracle.FDTD,
                       pusher=oracle.VAY, qed=1, nsteps=c["nsteps"], n_particles=len(c["rec"]), seed=seed, n_ion=0,
                       dx=c["dx"], dy=c["dx"], dz=c["dx"], dt=c["dt"], e_s=qed_case.E_S,
                       **norms(c["dx"], c["dx"], c["dx"]))
        _, prec, _, diag = run_ref(h, c["E"], c["B"], c["rec"], c["ids"])
        ph, po = prec[prec[:, 8] == 0], prec[prec[:, 8] == 1]
        summ.append(qed_case.summary_pairs(diag[-1][5], diag[-1][4], diag[-1][9], diag[-1][8], ph[:, 6], po[:, 6]))
    np.savez_compressed(os.path.join(HERE, "qed_pairs.npz"), header=h.tobytes(), summary=np.array(summ))
    s = np.array(summ)
    print(f"qed_pairs: {qed_case.SEEDS} seeds, photons {s[:, 0].mean():.0f}, positrons {s[:, 1].mean():.1f}"
          f" +- {s[:, 1].std():.1f}")


def case_qed_cascade():
    """The reduced colliding-pulse cascade (qed_case.inputs_cascade) over
    qed_case.SEEDS seeds of the reference's own random stream: the two
    pulses from the reference's f_init_cos (appended), then its step with
    photon emission and Breit-Wheeler pairs on."""
    c = qed_case.inputs_cascade()
    nx, ny, nz, d = c["nx"], c["ny"], c["nz"], c["dx"]
    E = np.zeros((nx, ny, nz, 3))
    B = np.zeros((nx, ny, nz, 3))
    for L in c["lasers"]:
        hl = gio.header(mode=gio.MODE_LASER, nx=nx, ny=ny, nz=nz, dx=d, dy=d, dz=d, dt=c["dt"], laser=L)
        El, Bl = run_ref(hl)
        E += El
        B += Bl
    summ = []
    for seed in range(1, qed_case.SEEDS + 1):
        h = gio.header(mode=gio.MODE_RUN, nx=nx, ny=ny, nz=nz, solver=oracle.FIVE_POINT, pusher=oracle.VAY, qed=1,
                       nsteps=c["nsteps"], n_particles=len(c["rec"]), seed=seed, n_ion=0, dx=d, dy=d, dz=d,
                       dt=c["dt"], e_s=qed_case.E_S, **norms(d, d, d))
        _, prec, _, diag = run_ref(h, E, B, c["rec"], c["ids"])
        last = dict(zip(gio.DIAG_NAMES, diag[-1]))
        ph, po = prec[prec[:, 8] == 0], prec[prec[:, 8] == 1]
        summ.append(qed_case.summary_cascade(last, ph[:, 6], po[:, 6]))
        print(f"qed_cascade seed {seed}: electrons {last['n_e']:.0f}, positrons {last['n_p']:.0f}, "
              f"photons {last['n_ph']:.0f}", flush=True)
    np.savez_compressed(os.path.join(HERE, "qed_cascade.npz"), header=h.tobytes(), summary=np.array(summ))


def case_merge_stats():
    """pmerging of the reference (src/mc.cpp:202-462), "ti" and "nl", over
    merge_case.SEEDS seeds of its own random stream: run summaries."""
    c = merge_case.inputs()
    out = {}
    for mode in ("ti", "nl"):
        summ = []
        for seed in range(1, merge_case.SEEDS + 1):
            h = gio.header(mode=gio.MODE_MERGE, nx=c["nx"], ny=c["ny"], nz=c["nz"], n_particles=len(c["rec"]),
                           seed=seed, n_ion=1, dx=0.3, dy=0.3, dz=0.3, dt=0.1, e_s=E_S)
            h["ion_cmr"][0] = merge_case.ION_CMR
            h["laser"][:4] = merge_case.PPD[mode]
            h["extra"][0] = 1 if mode == "ti" else 2
            with tempfile.TemporaryDirectory() as td:
                fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
                gio.write_input(fi, h, c["rec"], c["ids"])
                subprocess.run([oracle.REF_BIN, fi, fo], check=True, stdout=subprocess.DEVNULL)
                with open(fo, "rb") as f:
                    f.read(gio.HDR.itemsize)
                    n = int(np.frombuffer(f.read(8), "<i8")[0])
                    rec = np.frombuffer(f.read(n * 80), "<f8").reshape(n, 10)
            summ.append(merge_case.summary(rec))
        out[mode] = np.array(summ)
    np.savez_compressed(os.path.join(HERE, "merge_stats.npz"), **out)
    print(f"merge_stats: {merge_case.SEEDS} seeds, ti electrons kept {out['ti'][:, 0].mean():.0f}, "
          f"nl electrons kept {out['nl'][:, 0].mean():.0f}")


def case_mw(mode):
    """A run with a moving window (tests/golden/mw_case.py) by the reference
    harness (mode 7): fields, particles by id, diagnostics and the per-step
    (nmw, ratio) trace."""
    c = mw_case.inputs(mode)
    nx, ny, nz, dx, dt = c["nx"], c["ny"], c["nz"], c["dx"], c["dt"]
    solver, pusher = (oracle.FDTD, oracle.BORIS) if mode == "on" else (oracle.NDFX, oracle.VAY)
    h = gio.header(mode=gio.MODE_MW, nx=nx, ny=ny, nz=nz, solver=solver, pusher=pusher, qed=0,
                   nsteps=c["nsteps"], n_particles=len(c["rec"]), seed=1, n_ion=1, dx=dx, dy=dx, dz=dx, dt=dt,
                   e_s=E_S, **norms(dx, dx, dx))
    h["ion_cmr"][0] = mw_case.ION_CMR
    h["extra"][:] = [c["mw"], c["mwspeed"], c["tol"], c["density"], *c["npic"], c["t_add"]]
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        gio.write_input(fi, h, c["E"], c["B"], c["rec"], c["ids"])
        subprocess.run([oracle.REF_BIN, fi, fo], check=True, stdout=subprocess.DEVNULL)
        (Eo, Bo, Jo, BEo), prec, pids, diag, trace = gio.read_output_mw(fo, h)
    order = np.argsort(pids, kind="stable")
    np.savez_compressed(os.path.join(HERE, f"mw_{mode}.npz"), header=h.tobytes(), E_in=c["E"], B_in=c["B"],
                        part_in=c["rec"], ids_in=c["ids"], E=Eo, B=Bo, J=Jo, Be=BEo, part=prec[order],
                        ids=pids[order], diag=diag, trace=trace)
    print(f"mw_{mode}: {len(c['rec'])} particles in, {len(prec)} out, shifts {int(trace[-1, 0]) - 1}, "
          f"ratios {np.round(trace[:, 1], 4).tolist()}")


def case_film():
    """film() (src/init.cpp:537-646) of the reference with the arguments
    quillk-run derives from the texts of tests/golden/film_case.py: each
    film's initial particles, and for the window text the column film()
    fills at i = nx - 3 after nmw shifts (append_for_moving_window)."""
    import json
    drv = os.path.join(ROOT, "quill_amd", "bin", "quillk-run")
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, text in film_case.CONFS.items():
            conf = os.path.join(td, name + ".conf")
            with open(conf, "w") as f:
                f.write(text)
            j = json.loads(subprocess.run([drv, conf, "--dry-run"], check=True, capture_output=True,
                                          text=True).stdout)
            calls = []
            for k, a in enumerate(j["film_args"]):
                calls.append((f"{name}_init_{k}", a, 0))
                if name == "window":
                    for nmw in film_case.NMW:
                        b = list(a)
                        b[0] -= j["dx"] * nmw
                        b[1] -= j["dx"] * nmw
                        calls.append((f"{name}_col_{nmw}_{k}", b, 1))
            for key, a, append in calls:
                h = gio.header(mode=gio.MODE_FILM, nx=j["nx"], ny=j["ny"], nz=j["nz"], dx=j["dx"], dy=j["dy"],
                               dz=j["dz"], dt=j["dt"], e_s=E_S)
                h["laser"][:] = a[:16]
                h["extra"][0] = append
                h["extra"][1] = a[16]
                fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
                gio.write_input(fi, h)
                subprocess.run([oracle.REF_BIN, fi, fo], check=True, stdout=subprocess.DEVNULL)
                with open(fo, "rb") as f:
                    f.read(gio.HDR.itemsize)
                    n = int(np.frombuffer(f.read(8), "<i8")[0])
                    out[key] = np.frombuffer(f.read(n * 80), "<f8").reshape(n, 10).copy()
                out[key + "_args"] = np.array(a)
    np.savez_compressed(os.path.join(HERE, "film.npz"), **out)
    print("film: " + ", ".join(f"{k} {len(v)}" for k, v in out.items() if not k.endswith("_args")))


FOUT_FILES = ["ex", "ey", "ez", "bx", "by", "bz", "w", "inv", "rho", "rho_p", "rho_ph"]


def case_fout():
    """The reference's own plane writers (fout_ex .. fout_bz, fout_w, fout_inv,
    fout_rho with their yz-plane parts) on random fields and a mixed particle
    population, in text and in binary output mode."""
    nx, ny, nz, iyz = 14, 12, 10, 6
    dx = 0.05 * 2 * PI
    rng = np.random.default_rng(55)
    E = rng.uniform(-0.7, 0.7, (nx, ny, nz, 3))
    B = rng.uniform(-0.7, 0.7, (nx, ny, nz, 3))
    rec, ids, icmr = mixed_particles(nx, ny, nz, 55, 0.3)
    out = {"E": E, "B": B, "i_yz": np.int32(iyz)}
    for binary in (0, 1):
        h = gio.header(mode=gio.MODE_FOUT, nx=nx, ny=ny, nz=nz, n_particles=len(rec), n_ion=1, dx=dx, dy=dx, dz=dx,
                       dt=0.5 * dx, e_s=E_S)
        h["ion_cmr"][0] = icmr
        h["extra"][0] = iyz
        h["extra"][1] = binary
        with tempfile.TemporaryDirectory() as td:
            fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out")
            gio.write_input(fi, h, E, B, rec, ids.astype(np.int32))
            subprocess.run([oracle.REF_BIN, fi, fo], check=True, stdout=subprocess.DEVNULL)
            with open(fo, "rb") as f:
                f.read(gio.HDR.itemsize)
                nf = nx * ny * nz * 3
                _, _, rho, be = [np.frombuffer(f.read(nf * 8), "<f8").reshape(nx, ny, nz, 3) for _ in range(4)]
            out["rho"], out["Be"] = rho, be
            tag = "bin" if binary else "text"
            for name in FOUT_FILES:
                with open(fo + "." + name, "rb") as f:
                    out[f"{tag}_{name}"] = np.frombuffer(f.read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "fout.npz"), **out)
    print(f"fout: {len(FOUT_FILES)} files x 2 modes")


def main(names):
    """Generate the named fixtures (all when none are named)."""
    if not os.path.exists(oracle.REF_BIN):
        sys.exit("oracle/_ref/quill_ref missing: run `make -C oracle` with the reference tree present")
    dx = 0.05 * 2 * PI
    dt = 0.025 * 2 * PI
    cases = {
        "run_fdtd_boris": lambda: case_run("run_fdtd_boris", 14, 12, 10, dx, dx, dx, dt, oracle.FDTD, oracle.BORIS,
                                           10, 101, 0.8, 0.5),
        "run_ndfx_vay": lambda: case_run("run_ndfx_vay", 12, 14, 10, dx, 1.2 * dx, 1.4 * dx, 0.9 * dt, oracle.NDFX,
                                         oracle.VAY, 10, 202, 0.8, 0.5, external_bz=0.3),
        "run_five_vay": lambda: case_run("run_five_vay", 12, 10, 12, dx, dx, dx, dt, oracle.FIVE_POINT, oracle.VAY,
                                         10, 303, 0.8, 0.5),
        "run_fdtd_vay_cold": lambda: case_run("run_fdtd_vay_cold", 12, 12, 12, dx, dx, dx, dt, oracle.FDTD,
                                              oracle.VAY, 12, 404, 0.01, 0.05),
        "run_edges": case_edges,
        "laser_cos": case_laser,
        "rates": case_rates,
        "push": case_push,
        "qed_stats": case_qed_stats,
        "qed_pairs": case_qed_pairs,
        "qed_cascade": case_qed_cascade,
        "fout": case_fout,
        "merge_stats": case_merge_stats,
        "mw_on": lambda: case_mw("on"),
        "mw_auto": lambda: case_mw("auto"),
        "film": case_film,
    }
    for name in names or list(cases):
        cases[name]()


if __name__ == "__main__":
    main(sys.argv[1:])
