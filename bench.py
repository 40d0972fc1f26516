"""Benchmark of the per-step PIC loop on one MI355X.

Headline workload "wakefield-fdtd-boris-512x256x256", the largest case that
fits one GPU: 512 x 256 x 256 cells, dx = dy = dz = 0.05 lambda, dt = 0.025
lambda/c, FDTD solver, Boris pusher, QED off, uniform electron plasma at
0.001 n_cr with 8 particles per cell (is_inside cells only: 260.6 M
electrons), thermal momenta N(0, 0.01^2), linearly polarised laser a0 = 4,
waist 6 lambda, duration 8 lambda/c set as the reference's cos-envelope
initial condition at x-min, cell-tile sort at least every 20 steps.  FP32 on
the GPU.

A "step" is one pass of src/main.cpp:2319-2332 (padvance, compute_N,
compute_energy, fadvance) over the whole subdomain.  The timed region holds K
steps with every input resident in HBM; it is bracketed by stream
synchronisation of the library's stream (all work of a step is enqueued
there).  torch is not imported: the library brings its own HIP runtime.

Printed: one JSON line with value = particle-steps/s, the roofline of the
dominant kernel (the fused gather/push/deposit kernel) from HIP events on the
library's stream, the plain FP64 GPU method on the same workload, the FP32 vs
FP64 field error, and the reference CPU path timed on this host (on one
GPU's x-slab of the grid, a bounded sample).  Two more blocks: the
256 x 128 x 128 NDFX / Vay case with its 500-step FP32-vs-FP64 errors, and the
64 x 32 x 32 FDTD / Boris case with the reference's own 200 steps on the CPU.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import quill_amd as qa  # noqa: E402

TWO_PI = 2 * math.pi
HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s on 1 MI355X (push+deposit+Maxwell), % HBM roofline vs FP64"

WORKLOADS = {
    # name: grid, solver, pusher, density (n_cr), laser a0, waist (lambda), duration (lambda/c)
    "uniform-ndfx-vay-256x128x128": dict(grid=(256, 128, 128), solver="ndfx", pusher="vay", density=0.01, a0=1.0,
                                         waist=2.0, duration=4.0),
    "uniform-fdtd-boris-64x32x32": dict(grid=(64, 32, 32), solver="fdtd", pusher="boris", density=0.01, a0=1.0,
                                        waist=2.0, duration=4.0, source=True),
    "wakefield-fdtd-boris-512x256x256": dict(grid=(512, 256, 256), solver="fdtd", pusher="boris", density=0.001,
                                             a0=4.0, waist=6.0, duration=8.0),
}
PPC = 8
SIGMA_U = 0.01
SEED = 20261202
DX_LAMBDA = 0.05
DT_LAMBDA = 0.025


def laser_params(w, dx):
    """f_init_cos arguments for a pulse whose rear edge sits at x-min, moving +x.

    xsigma / ysigma are the 1/e^2 half sizes the reference's cos envelope is
    matched to (src/init.cpp:34-37): xsigma = duration / 2, ysigma = zsigma =
    waist.  x0 places the pulse centre one envelope half-length xs from the
    left wall; x1 < 0 gives propagation along +x (src/init.cpp:56-59)."""
    xsigma = w["duration"] / 2 * TWO_PI
    ysig = w["waist"] * TWO_PI
    xs = xsigma * 2 * math.sqrt(2 * math.pi) / 3
    nx = w["grid"][0]
    x0 = xs
    x1 = x0 - nx * dx / 2   # pulse centre relative to the box centre
    if x1 >= 0:
        x1 = -1e-6
    # a0y a0z xsigma ysigma zsigma x0 b_sign x1 phase y0 z0 append phi xt yt zt
    return dict(a0y=w["a0"], a0z=0.0, xsigma=xsigma, ysigma=ysig, zsigma=ysig, x0=x0, b_sign=1, x1=x1)


# cell-tile sorts of the FP32 run: whenever more than 1 % of the electrons
# were outside their tile windows, and at least every SORT_EVERY steps
# (the large wakefield case asks for a sort every 20 steps).  Measured at
# 256x128x128 over 500 steps: 1.011 ms per step, against 1.028 with a fixed
# sort every 10 steps; thresholds of 0.5 % and 2 % gave 1.020 and 1.024 ms
SORT_EVERY = {"uniform-ndfx-vay-256x128x128": 40, "uniform-fdtd-boris-64x32x32": 20,
              "wakefield-fdtd-boris-512x256x256": 20}
SORT_FALLBACK = 0.01
HEADLINE = "wakefield-fdtd-boris-512x256x256"
# the CPU sample of the headline: one GPU's x-slab of an 8-way split
CPU_SAMPLE_GRID = {"wakefield-fdtd-boris-512x256x256": (64, 256, 256)}


SOURCE_PLANE = 2


def source_params(w, dx):
    """A laser source at x plane 2 (side 0) whose pulse front stands at that
    plane when the run starts: the pulse of laser_params with its centre at
    2 dx - xs, moving +x, entering through x-min."""
    lp = laser_params(w, dx)
    lp["x0"] = SOURCE_PLANE * dx - lp["xsigma"] * 2 * math.sqrt(2 * math.pi) / 3
    lp["x1"] = -1.0
    return lp


def make_region(w, precision, deposit, sort_every=20, sort_fallback=SORT_FALLBACK):
    nx, ny, nz = w["grid"]
    dx = DX_LAMBDA * TWO_PI
    dt = DT_LAMBDA * TWO_PI
    r = qa.Region(nx, ny, nz, dx, dx, dx, dt, solver=w["solver"], pusher=w["pusher"], precision=precision,
                  deposit=deposit, sort_every=sort_every, sort_fallback=sort_fallback, device=0)
    n = r.load_uniform_plasma(qa.ELECTRON, PPC, w["density"], SIGMA_U, SEED)
    if w.get("source"):
        r.add_laser_source(0, SOURCE_PLANE, **source_params(w, dx))
    else:
        r.init_laser_cos(**laser_params(w, dx))
    r.init_boundaries()
    return r, n


def timed_run(w, precision, deposit, steps, warmup, sort_every=20, sort_fallback=SORT_FALLBACK):
    r, n0 = make_region(w, precision, deposit, sort_every, sort_fallback)
    if warmup:
        r.step(warmup)
    r.sync()
    n_start = r.count(qa.ELECTRON)
    t0 = time.perf_counter()
    diag = r.step(steps, diag=True)
    r.sync()
    t1 = time.perf_counter()
    ms_total, ms_part, ms_field = r.timing()
    fallback = r.fallback_fraction()
    # particles pushed in step l are those alive before it
    pushed = n_start + sum(d["n_e"] for d in diag[:-1])
    r.close()
    return dict(seconds=t1 - t0, pushed=pushed, n_start=n_start, n_loaded=n0, ms_events=ms_total,
                ms_particles=ms_part, ms_fields=ms_field, energy_f_last=diag[-1]["energy_f"], fallback=fallback)


def field_error(w, steps, moments=False, sort_every=20):
    """Per-component relative L2 error of FP32 vs the plain FP64 method after
    `steps`; with `moments`, also the relative errors of the field energy,
    the electrons' kinetic energy and charge (compute_energy / compute_N)."""
    out = {}
    res, dg = {}, {}
    for prec, dep in (("fp32", "tiled"), ("fp64", "global")):
        r, _ = make_region(w, prec, dep, sort_every)
        d = r.step(steps, diag=moments)
        res[prec] = r.get_fields()
        dg[prec] = d[-1] if moments else None
        r.close()
    for name, a, b in zip(("E", "B", "J"), res["fp32"][:3], res["fp64"][:3]):
        for c, axis in enumerate("xyz"):
            den = np.linalg.norm(b[..., c])
            out[f"{name}{axis}"] = float(np.linalg.norm(a[..., c] - b[..., c]) / den) if den > 0 else 0.0
    if moments:
        for k in ("energy_f", "energy_e", "N_e"):
            out[k] = abs(dg["fp32"][k] - dg["fp64"][k]) / abs(dg["fp64"][k])
    return out


def cpu_model():
    """The host CPU's model name (the reference CPU path runs there)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def ref_binary():
    """The reference harness build for this host: the x86-64-v4 (AVX-512)
    build when the CPU has that feature set, else the x86-64-v3 build."""
    import oracle

    v4 = oracle.REF_BIN + "_v4"
    need = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((set(line.split(":", 1)[1].split()) for line in f if line.startswith("flags")), set())
    except OSError:
        flags = set()
    if os.path.exists(v4) and need <= flags:
        return v4, "x86-64-v4"
    return oracle.REF_BIN, "x86-64-v3"


def cpu_baseline(w, budget_steps=3, grid=None, warm=1):
    """The reference's own step (oracle/_ref/quill_ref*: the reference's
    src/*.cpp units, -O3, single-threaded as the reference runs per rank) on
    the workload's plasma and laser over `grid` (default: the workload's
    grid): `warm` warm-up steps, then `budget_steps` timed steps.  Falls back
    to the FP64 C restatement when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import golden_io as gio
    import oracle

    nx, ny, nz = grid or w["grid"]
    dx = DX_LAMBDA * TWO_PI
    dt = DT_LAMBDA * TWO_PI
    lp = laser_params(dict(w, grid=(nx, ny, nz)), dx)
    L = np.zeros(16)
    L[:8] = [lp["a0y"], lp["a0z"], lp["xsigma"], lp["ysigma"], lp["zsigma"], lp["x0"], lp["b_sign"], lp["x1"]]
    L[11] = 1
    solver = {"fdtd": 0, "ndfx": 1, "five_point": 2}[w["solver"]]
    pusher = {"vay": 0, "boris": 1}[w["pusher"]]
    cores = 1
    what = "full grid" if (nx, ny, nz) == tuple(w["grid"]) else \
        f"x-slab {nx}x{ny}x{nz} of the {'x'.join(map(str, w['grid']))} grid (one GPU's share of an 8-way split)"
    binary, march = ref_binary()
    if os.path.exists(binary):
        h = gio.header(mode=gio.MODE_TIME, nx=nx, ny=ny, nz=nz, solver=solver, pusher=pusher, nsteps=budget_steps,
                       seed=SEED & 0x7FFFFFFF, dx=dx, dy=dx, dz=dx, dt=dt, laser=L)
        h["extra"][:6] = [PPC, w["density"], SIGMA_U, warm, 1, nx]
        with tempfile.TemporaryDirectory() as td:
            fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            gio.write_input(fi, h)
            res = subprocess.run([binary, fi, fo], capture_output=True, text=True, check=True)
        rec = json.loads(res.stdout.strip().splitlines()[-1])
        rate = rec["pushed"] / rec["seconds"]
        return dict(value=rate, unit="particle-steps/s", cores=cores, kind="reference",
                    sample=f"reference src/*.cpp step (oracle/_ref, -O3 -march={march}), {what}, "
                           f"{rec['particles']} particles loaded, {warm} warm-up + {rec['steps']} timed steps "
                           f"pushing {rec['pushed']} particle-steps, "
                           f"{rec['seconds']:.2f} s, single thread of {cpu_model()} ({os.cpu_count()} threads)")
    o = oracle.Oracle(nx, ny, nz, dx, dx, dx, dt, solver=solver, pusher=pusher)
    n = o.load_uniform(PPC, w["density"], SIGMA_U, SEED)
    o.laser_cos(*L)
    o.init_boundaries()
    o.step(warm)
    t0 = time.perf_counter()
    o.step(budget_steps)
    sec = time.perf_counter() - t0
    return dict(value=n * budget_steps / sec, unit="particle-steps/s", cores=cores, kind="port",
                sample=f"FP64 C restatement (oracle/), {what}, {n} particles, {budget_steps} steps, {sec:.2f} s, "
                       f"single thread of {cpu_model()}")


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime the library
    uses.  Profiled with FETCH_SIZE / WRITE_SIZE, it gives the factor between
    those counters and bytes actually moved for a kernel of known traffic."""
    import ctypes as C
    qa.load_library()   # the same libamdhip64 as the library's
    hip = C.CDLL("libamdhip64.so.7")   # already loaded by libquillk.so
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "rc": rc}), flush=True)


def measured_traffic(workload):
    """HBM bytes per launch of k_push_tiled from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE in KiB per dispatch,
    scaled by the factors a plain copy of known size reads on the same
    counters, both stated in the summary), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"k_push_tiled_{workload}.md")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            for key in ("FETCH_SIZE_KiB_per_dispatch", "WRITE_SIZE_KiB_per_dispatch", "FETCH_SIZE_factor",
                        "WRITE_SIZE_factor"):
                if line.strip().startswith(key + ":"):
                    vals[key] = float(line.split(":", 1)[1].split()[0])
    if len(vals) < 4:
        return None
    return 1024.0 * (vals["FETCH_SIZE_KiB_per_dispatch"] * vals["FETCH_SIZE_factor"] +
                     vals["WRITE_SIZE_KiB_per_dispatch"] * vals["WRITE_SIZE_factor"])


def adjoint_check(steps, every):
    """The differentiable laser-plasma case (quill_amd.adjoint.LaserPlasma).

    Full size: 128 x 64 x 64 cells, 8 per cell (3.72M electrons), density
    ramp of 32 nodes about 0.005 n_cr, a0 = 2 pulse with waist, chirp and 16
    transverse-phase coefficients, NDFX + Vay, `steps` steps with a snapshot
    every `every`; the objective is the electrons' kinetic energy above
    5 mc^2 (softplus-smoothed).  The FP32 adjoint gradient (all 51
    parameters) against the FP64 adjoint gradient, and the timings.

    Central differences: the discrete step jumps when a particle crosses a
    cell face (each field component is interpolated across its own axis
    only) or is deleted, and at 3.7M particles x 300 steps every practical
    interval holds such jumps, so the full-size quotients are reported for
    information only.  The adjoint is checked against FP64 central
    differences (FP64 forward runs of this library) on the reduced case of
    the tests (32 x 16 x 16 cells, 2 per cell, 30 steps), where quotients at
    h and h / 10 agree to 1e-7, which shows that the interval is free of
    jumps."""
    import torch   # the adjoint's host side is torch; the plain bench needs none
    from quill_amd.adjoint import LaserPlasma

    def log(msg):
        print(f"adjoint check: {msg}", file=sys.stderr, flush=True)

    def dot(g, v):
        return sum(float((g[k] * v[k]).sum()) for k in v)

    def directions(lp):
        out = []
        for name, keep in (("all", None), ("laser", ("a0", "ysigma", "chirp", "c")), ("density", ("nodes",))):
            v = lp.direction(5)
            out.append((name, {k: x for k, x in v.items() if keep is None or k in keep}))
        return out

    lp = LaserPlasma(steps=steps, every=every)
    log(f"{len(lp.pos)} electrons, {steps} steps")
    t = time.perf_counter()
    L32, g32 = lp.gradient("fp32")
    t32 = time.perf_counter() - t
    log(f"FP32 forward + adjoint {t32:.1f} s")
    t = time.perf_counter()
    L64, g64 = lp.gradient("fp64")
    t64 = time.perf_counter() - t
    log(f"FP64 forward + adjoint {t64:.1f} s")
    a = torch.cat([g64[k].reshape(-1) for k in g64])
    b = torch.cat([g32[k].reshape(-1) for k in g32])
    full_fd = {}
    name, v = directions(lp)[1]
    for h in (1e-4, 1e-5, 1e-6):
        full_fd[h] = lp.central_difference(v, h)
        log(f"full size, {name} direction, h = {h}: {full_fd[h]:.9e}")

    small = LaserPlasma(shape=(32, 16, 16), ppc=2, steps=30, every=10, waist=0.2, duration=0.3, x0=0.6,
                        threshold=0.5, width=0.5, margin=2)
    _, s64 = small.gradient("fp64")
    _, s32 = small.gradient("fp32")
    checks = []
    for name, v in directions(small):
        fd = small.central_difference(v, 1e-6)
        fd10 = small.central_difference(v, 1e-7)
        checks.append({"direction": name, "fp64_central_difference": fd, "h": 1e-6,
                       "jump_free": abs(fd10 - fd) <= 1e-7 * abs(fd),
                       "fp32_adjoint": dot(s32, v), "fp64_adjoint": dot(s64, v),
                       "fp32_rel_err": abs(dot(s32, v) - fd) / abs(fd),
                       "fp64_rel_err": abs(dot(s64, v) - fd) / abs(fd)})
        log(f"reduced case, {name}: {checks[-1]}")
    nx, ny, nz = lp.shape
    name, v = directions(lp)[1]
    return {"metric": "adjoint gradient error vs FP64 central differences (max over directions, reduced case)",
            "value": max(c["fp32_rel_err"] for c in checks), "unit": "relative", "higher_is_better": False,
            "n_gpus": 1, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"laser-plasma-adjoint-{nx}x{ny}x{nz}", "particles": len(lp.pos), "ppc": lp.ppc,
                       "steps": steps, "checkpoint_every": every, "solver": lp.solver, "pusher": lp.pusher,
                       "density_nodes": len(lp.theta0["nodes"]), "phase_coefficients": len(lp.theta0["c"]),
                       "objective": "sum |q| softplus(gamma - 6) over electrons (kinetic energy above 5 mc^2)"},
            "full_size": {"objective_fp32": L32, "objective_fp64": L64,
                          "seconds_fp32_forward_adjoint": t32, "seconds_fp64_forward_adjoint": t64,
                          "fp32_vs_fp64_gradient_rel_l2": float((a - b).norm() / a.norm()),
                          "laser_direction": {"fp32_adjoint": dot(g32, v), "fp64_adjoint": dot(g64, v),
                                              "fp64_central_differences_with_jumps": full_fd}},
            "reduced_case": {"grid": list(small.shape), "particles": len(small.pos), "steps": small.steps,
                             "checkpoint_every": small.every, "checks": checks}}


def qed_cascade_run(steps, seed, precision="fp32", plane=4):
    """One run of the QED cascade case on one GPU: 128^3 cells spanning a
    4 lambda cube, five-point solver, Vay pusher, photon emission,
    Breit-Wheeler pairs and vacuum pairs on; lambda = 1 um
    (e_s = lambda / lambda_C, src/main.cpp:2276).  dt = dx / 4: the
    five-point solver's interior update is unstable above about dt = 0.41 dx
    in 3-D (its grid-scale mode, k = pi on every axis, has the difference
    factor 2 (b - 2a) with a < 0), as the reference's is.

    Two counter-propagating circularly polarised pulses (f_init_cos, peak
    |a| = 1000, waist 1.5 lambda, duration 5 lambda/c, 8.4 lambda long) are
    injected from -x and +x through laser sources at the planes `plane` and
    nx - `plane`; their peaks meet at the centre at step 0: the part of each
    pulse that has already entered (its total-field side of its source
    plane) is set as the initial field, the rest enters through the planes
    during the run, so the centre holds the colliding peaks for the whole
    run.  1e4 electrons at rest, uniform in the central lambda^3; the
    reference's "ti" merging holds each species at crnp = 1e7
    macro-particles.  Throughput counts every live particle (electrons,
    positrons, photons) pushed per step."""
    nx = 128
    lam = TWO_PI
    d = 4 * lam / nx
    dt = 0.25 * d
    e_s = 1e-4 / 2.4263086e-10
    a = 1000 / math.sqrt(2)
    xsigma, waist = 2.5 * lam, 1.5 * lam
    rng = np.random.default_rng(seed)
    n0 = 10000
    pos = nx / 2 + (rng.random((n0, 3)) - 0.5) * (lam / d)
    crnp = 1e7
    r = qa.Region(nx, nx, nx, d, d, d, dt, solver="five_point", pusher="vay", precision=precision, qed=True,
                  e_s=e_s, seed=seed, sort_every=0, merging="ti", crnp=crnp)
    try:
        # pulse moving +x (x1 = -1) enters through side 0, the other through
        # side 1; each one's initial field only on its total-field side
        E, B = np.zeros((nx, nx, nx, 3)), np.zeros((nx, nx, nx, 3))
        for side, x1 in ((0, -1.0), (1, 1.0)):
            L = dict(a0y=a, a0z=a, xsigma=xsigma, ysigma=waist, zsigma=waist, x0=nx * d / 2, b_sign=1, x1=x1)
            r.init_laser_cos(append=False, **L)
            El, Bl, _, _ = r.get_fields()
            sl = slice(plane, None) if side == 0 else slice(None, nx - plane)
            E[sl] += El[sl]
            B[sl] += Bl[sl]
            r.add_laser_source(side, plane if side == 0 else nx - plane, **L)
        r.set_fields(E, B)
        del E, B, El, Bl
        z = np.zeros(n0)
        r.add_particles(qa.ELECTRON, pos[:, 0], pos[:, 1], pos[:, 2], z, z, z, np.full(n0, -1e-3))
        r.init_boundaries()
        r.sync()
        t0 = time.perf_counter()
        diag = r.step(steps, diag=True)
        r.sync()
        sec = time.perf_counter() - t0
    finally:
        r.close()
    pushed = sum(x["n_e"] + x["n_p"] + x["n_ph"] for x in diag[:-1]) + n0
    return sec, pushed, diag


def qed_cascade(steps, seeds, precision="fp32"):
    """The QED cascade case over `seeds` seeds (1..seeds): throughput of each
    run, and the growth of the electron, positron and photon counts as mean
    and standard deviation over the seeds."""
    runs = []
    for seed in range(1, seeds + 1):
        sec, pushed, diag = qed_cascade_run(steps, seed, precision)
        runs.append((sec, pushed, diag))
        last = diag[-1]
        log(f"cascade seed {seed}: {sec:.2f} s, electrons {last['n_e']}, positrons {last['n_p']}, "
            f"photons {last['n_ph']}")
    sec = sum(x[0] for x in runs)
    pushed = sum(x[1] for x in runs)
    marks = sorted(set(list(range(0, steps, max(1, steps // 8))) + [steps - 1]))
    growth = []
    for k in marks:
        v = np.array([[x[2][k]["n_e"], x[2][k]["n_p"], x[2][k]["n_ph"]] for x in runs], float)
        growth.append({"step": k + 1, "mean": v.mean(0).tolist(), "std": v.std(0, ddof=1).tolist() if seeds > 1 else None})
    fin = np.array([[x[2][-1]["n_e"], x[2][-1]["n_p"], x[2][-1]["n_ph"], x[2][-1]["energy_ph"]] for x in runs], float)
    return {"metric": "QED cascade particle-steps/s on 1 MI355X", "value": pushed / sec,
            "unit": "particle-steps/s", "higher_is_better": True, "n_gpus": 1, "steps": steps,
            "ms_per_step": sec * 1e3 / (steps * seeds), "dtype": "f32" if precision == "fp32" else "f64",
            "data": "synthetic",
            "config": {"workload": "qed-cascade-five-vay-128", "grid": [128, 128, 128], "dx_lambda": 4 / 128,
                       "dt_lambda": 1 / 128, "a0": 1000, "polarisation": "circular", "waist_lambda": 1.5,
                       "duration_lambda": 5.0, "injection": "laser sources at x planes 4 and 124, peaks meeting "
                       "at the centre at step 0", "seed_electrons": 10000, "rng_seeds": list(range(1, seeds + 1)),
                       "merging": "ti", "crnp": 1e7},
            "final_mean": {"electrons": fin[:, 0].mean(), "positrons": fin[:, 1].mean(), "photons": fin[:, 2].mean(),
                           "photon_energy": fin[:, 3].mean()},
            "final_std": {"electrons": fin[:, 0].std(ddof=1), "positrons": fin[:, 1].std(ddof=1),
                          "photons": fin[:, 2].std(ddof=1)} if seeds > 1 else None,
            "growth": growth}


def push_roofline(run, steps, ncell, workload):
    """Roofline of the fused gather/push/deposit kernel from the HIP events
    of the particle phase: algorithmic bytes 64 per particle-step (32 B
    record read and written) + 36 per cell (E and Be gathered, J flushed)."""
    np_avg = run["pushed"] / steps
    push_ms = run["ms_particles"] / steps
    push_bytes = 64.0 * np_avg + 36.0 * ncell
    achieved = push_bytes / (push_ms * 1e-3) / 1e9
    traffic = measured_traffic(workload)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": (f"profiles/k_push_tiled_{workload}.md: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                               "of an earlier run of this workload, copy-calibrated; not measured in this run")
            if traffic is not None else None,
            "kernel": "k_push_tiled (gather+push+deposit)", "avg_launch_ms": push_ms,
            "algorithmic_bytes_per_launch": push_bytes}


def workload_block(name, steps, warmup, fp64_steps, err_steps, moments_steps=0):
    """Throughput, push roofline, plain FP64 method and FP32-vs-FP64 errors
    of one workload."""
    w = WORKLOADS[name]
    nx, ny, nz = w["grid"]
    ncell = nx * ny * nz
    se = SORT_EVERY[name]
    run = timed_run(w, "fp32", "tiled", steps, warmup, se)
    ms_step = run["seconds"] * 1e3 / steps
    value = run["pushed"] / run["seconds"]
    step_bytes = 64.0 * run["pushed"] / steps + 96.0 * ncell
    out = {
        "value": value, "unit": "particle-steps/s", "steps": steps, "warmup": warmup, "ms_per_step": ms_step,
        "config": {"workload": name, "grid": [nx, ny, nz], "cells": ncell, "ppc": PPC, "particles": run["n_start"],
                   "solver": w["solver"], "pusher": w["pusher"], "qed": False, "density_ncr": w["density"],
                   "laser_a0": w["a0"], "laser": "source at x plane 2 (enters through x-min)" if w.get("source")
                   else "f_init_cos initial condition, rear edge at x-min", "deposit": "tiled", "sort_every": se,
                   "sort_fallback": SORT_FALLBACK, "dx_lambda": DX_LAMBDA, "dt_lambda": DT_LAMBDA},
        "roofline": push_roofline(run, steps, ncell, name),
        "step_hbm": {"algorithmic_bytes_per_step": step_bytes,
                     "achieved_GBs": step_bytes / (ms_step * 1e-3) / 1e9,
                     "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        # field-only rate: the four field kernels of a step (B half step,
        # E step, B half step, interpolate_be) over their own time; at FP32
        # they move 36 + 60 + 36 + 24 = 156 B per cell
        "field_cell_updates_per_s": ncell * steps / (run["ms_fields"] * 1e-3),
        "field_phase_GBs": 156.0 * ncell * steps / (run["ms_fields"] * 1e-3) / 1e9,
        "ms_particles_per_step": run["ms_particles"] / steps,
        "ms_fields_per_step": run["ms_fields"] / steps,
        # between the first step's particle phase and the last step's field
        # phase, what is neither: the cell-tile sorts of steps 2 .. K (and
        # the gaps between launches); outside that window: the first step's
        # sort and the host's own time
        "ms_sort_in_steps_per_step": (run["ms_events"] - run["ms_particles"] - run["ms_fields"]) / steps,
        "ms_outside_events_per_step": ms_step - run["ms_events"] / steps,
        "deposit_fallback_fraction": run["fallback"],
    }
    if fp64_steps:
        plain = timed_run(w, "fp64", "global", fp64_steps, 2, se)
        v64 = plain["pushed"] / plain["seconds"]
        out["fp64_plain"] = {"value": v64, "unit": "particle-steps/s", "steps": fp64_steps,
                             "ms_per_step": plain["seconds"] * 1e3 / fp64_steps, "speedup_fp32": value / v64,
                             "method": "FP64 SoA, one thread per particle, global FP64 atomics"}
    if err_steps:
        out["fp32_vs_fp64_rel_l2"] = field_error(w, err_steps, moments=True, sort_every=se)
        out["fp32_vs_fp64_steps"] = err_steps
    if moments_steps:
        # the whole run length: the components carried by plasma noise
        # decorrelate (face crossings one step apart change the current),
        # the laser's components and the global moments do not
        out[f"fp32_vs_fp64_rel_l2_{moments_steps}_steps"] = field_error(w, moments_steps, moments=True,
                                                                          sort_every=se)
    return out


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=HEADLINE, choices=sorted(WORKLOADS))
    ap.add_argument("--fp64-steps", type=int, default=20, help="steps of the plain FP64 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 comparison runs")
    ap.add_argument("--no-extra", action="store_true", help="skip the 256x128x128 and 64x32x32 blocks")
    ap.add_argument("--adjoint", action="store_true",
                    help="only run the adjoint check of the laser-plasma case (128x64x64, 300 steps)")
    ap.add_argument("--adjoint-steps", type=int, default=300)
    ap.add_argument("--qed-cascade", action="store_true",
                    help="only run the QED cascade case (128^3, two colliding CP pulses, a0 = 1000)")
    ap.add_argument("--seeds", type=int, default=16, help="seeds of the QED cascade case")
    args = ap.parse_args()
    if args.qed_cascade:
        print(json.dumps(qed_cascade(args.steps if args.steps != 20 else 400, args.seeds)), flush=True)
        return
    if args.adjoint:
        print(json.dumps(adjoint_check(args.adjoint_steps, 25)), flush=True)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the multi-GPU slab decomposition is a design, see the sharding notes)", file=sys.stderr)
        sys.exit(2)
    name = args.workload
    log(f"{name}: {args.steps} steps after {args.warmup} warm-up")
    head = workload_block(name, args.steps, args.warmup, 0 if args.no_compare else min(args.fp64_steps, args.steps),
                          0 if args.no_compare else min(50, args.steps))
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
    }
    line.update({k: v for k, v in head.items() if k not in ("value", "unit", "steps", "warmup", "ms_per_step")})
    if not args.no_cpu:
        log("CPU baseline")
        w = WORKLOADS[name]
        line["cpu_baseline"] = cpu_baseline(w, budget_steps=2, grid=CPU_SAMPLE_GRID.get(name))
    if not args.no_extra and name == HEADLINE:
        # the 256x128x128 NDFX / Vay case: 500 steps, FP32 against FP64
        # field and moment errors over the whole run
        c2 = "uniform-ndfx-vay-256x128x128"
        log(c2)
        line["uniform_ndfx_vay_256x128x128"] = workload_block(
            c2, args.steps, args.warmup, 0 if args.no_compare else min(args.fp64_steps, args.steps),
            0 if args.no_compare else min(50, args.steps), 0 if args.no_compare else 500)
        # the 64x32x32 FDTD / Boris case: the GPU (laser entering through a
        # source at x-min), and the reference's own 200 steps on one host
        # core (the reference has no source: its laser is f_init_cos with
        # the pulse's rear edge at x-min; the per-step work is the same)
        c1 = "uniform-fdtd-boris-64x32x32"
        log(c1)
        blk = workload_block(c1, 200, 5, 0, 0)
        if not args.no_cpu:
            blk["cpu_reference_200_steps"] = cpu_baseline(WORKLOADS[c1], budget_steps=200, warm=0)
        line["uniform_fdtd_boris_64x32x32"] = blk
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
