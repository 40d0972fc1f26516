"""quill_amd -- MI355X-native per-step PIC loop of QUILL, Python host binding.

The compute path is libquillk.so (HIP kernels for gfx950 behind the C ABI of
include/quillk.h).  This module binds that ABI with ctypes and mirrors the
reference's spatial_region interface (include/main.h:17-170): create the
subdomain, set fields, place particles, f_init_boundaries / interpolate_be,
and step in the order of src/main.cpp:2319-2332.

There is no fallback: if the library is missing or no GPU is present, the
calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libquillk.so")

QK_OK = 0
ERRORS = {1: "invalid argument", 2: "capacity exceeded", 3: "HIP error", 4: "bad state", 5: "not built"}

SOLVERS = {"fdtd": 0, "ndfx": 1, "five_point": 2}
PUSHERS = {"vay": 0, "boris": 1}
PRECISIONS = {"fp32": 0, "fp64": 1}
DEPOSITS = {"tiled": 0, "global": 1}
ELECTRON, POSITRON, PHOTON, ION0 = 0, 1, 2, 3
MAX_ION = 4
MAX_SPECIES = ION0 + MAX_ION
SPECIES_NAMES = {"electron": ELECTRON, "positron": POSITRON, "photon": PHOTON}

EXPORTS = ["qk_create", "qk_destroy", "qk_last_error", "qk_set_fields", "qk_get_fields", "qk_add_particles",
           "qk_load_uniform_plasma", "qk_init_laser_cos", "qk_init_boundaries", "qk_step", "qk_sync",
           "qk_particle_count", "qk_get_particles", "qk_compute_rho", "qk_compute_rho_noalloc", "qk_last_timing",
           "qk_last_fallback", "qk_get_stream", "qk_pmerging", "qk_moving_window", "qk_max_w", "qk_fill_cells",
           "qk_tape_begin", "qk_adjoint", "qk_add_laser_source", "qk_tangent_begin", "qk_tangent_get"]


class QuillError(RuntimeError):
    pass


class Params(C.Structure):
    """qk_params of include/quillk.h."""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("solver", C.c_int32), ("pusher", C.c_int32), ("qed", C.c_int32), ("precision", C.c_int32),
                ("deposit", C.c_int32), ("sort_every", C.c_int32), ("track_ids", C.c_int32),
                ("device", C.c_int32),
                ("e_s", C.c_double), ("external_bz", C.c_double), ("q_vac", C.c_double),
                ("seed", C.c_uint64), ("capacity", C.c_int64),
                ("n_ion", C.c_int32), ("ion_cmr", C.c_double * MAX_ION),
                ("norm_n", C.c_double), ("norm_ef", C.c_double), ("norm_ep", C.c_double),
                ("merging", C.c_int32), ("crnp", C.c_double), ("sort_fallback", C.c_double)]


class Diag(C.Structure):
    """qk_diag of include/quillk.h (compute_N / compute_energy, src/fout.cpp:41-138)."""
    _fields_ = [("N_e", C.c_double), ("N_p", C.c_double), ("N_ph", C.c_double),
                ("n_e", C.c_int64), ("n_p", C.c_int64), ("n_ph", C.c_int64), ("n_ion", C.c_int64),
                ("energy_f", C.c_double), ("energy_e", C.c_double), ("energy_p", C.c_double),
                ("energy_ph", C.c_double), ("energy_ion", C.c_double), ("energy_ions", C.c_double * MAX_ION),
                ("n_nonfinite", C.c_int64)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["energy_ions"] = list(self.energy_ions)
        return d


_lib = None


def load_library(path=LIB_PATH):
    """Load libquillk.so and declare the ABI; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise QuillError(f"{path} not found: build it with `make -C quill_amd/csrc` or __graft_entry__.build()")
    L = C.CDLL(path)
    vp, dp, ip, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    L.qk_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.qk_destroy.argtypes = [vp]
    L.qk_last_error.restype = C.c_char_p
    L.qk_last_error.argtypes = []
    L.qk_set_fields.argtypes = [vp, dp, dp]
    L.qk_get_fields.argtypes = [vp, dp, dp, dp, dp]
    L.qk_add_particles.argtypes = [vp, C.c_int, C.c_int64] + [dp] * 7 + [ip, i64p]
    L.qk_load_uniform_plasma.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                         C.c_int, i64p]
    L.qk_init_laser_cos.argtypes = [vp, dp]
    L.qk_init_boundaries.argtypes = [vp]
    L.qk_step.argtypes = [vp, C.c_int, C.POINTER(Diag)]
    L.qk_sync.argtypes = [vp]
    L.qk_particle_count.argtypes = [vp, C.c_int, i64p]
    L.qk_get_particles.argtypes = [vp, C.c_int, C.c_int64, dp, ip, i64p]
    L.qk_compute_rho.argtypes = [vp, dp]
    L.qk_compute_rho_noalloc.argtypes = [vp, dp]
    L.qk_add_laser_source.argtypes = [vp, dp, C.c_int, C.c_int]
    L.qk_last_timing.argtypes = [vp, dp, dp, dp]
    L.qk_last_fallback.argtypes = [vp, dp]
    L.qk_get_stream.argtypes = [vp, C.POINTER(vp)]
    L.qk_pmerging.argtypes = [vp, C.c_int, dp]
    L.qk_moving_window.argtypes = [vp]
    L.qk_max_w.argtypes = [vp, C.c_int, C.c_int, dp]
    L.qk_fill_cells.argtypes = [vp, C.c_int, C.c_int64, ip, dp, ip, C.c_double, C.c_int32, i64p]
    L.qk_tape_begin.argtypes = [vp, C.c_int]
    L.qk_adjoint.argtypes = [vp, dp, dp, C.POINTER(dp), dp, dp, C.POINTER(dp)]
    L.qk_tangent_begin.argtypes = [vp, dp, dp, C.POINTER(dp)]
    L.qk_tangent_get.argtypes = [vp, dp, dp, C.POINTER(dp)]
    for name in EXPORTS:
        if name not in ("qk_last_error",):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(rc, what):
    if rc != QK_OK:
        msg = load_library().qk_last_error().decode(errors="replace")
        raise QuillError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Region:
    """One subdomain on one GPU: the reference's spatial_region (include/main.h:17-170).

    Parameters follow qk_params; solver / pusher / precision / deposit are
    given by name.  Fields are numpy arrays (nx, ny, nz, 3) in FP64 on the
    host side of the boundary.
    """

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, solver="fdtd", pusher="boris", precision="fp32",
                 deposit="tiled", qed=False, sort_every=0, track_ids=False, device=0, e_s=1e10,
                 external_bz=0.0, q_vac=0.0, seed=1, capacity=0, ion_cmr=(), norm_n=1.0, norm_ef=1.0,
                 norm_ep=1.0, merging=None, crnp=0.0, sort_fallback=0.0):
        L = load_library()
        p = Params()
        p.nx, p.ny, p.nz = nx, ny, nz
        p.dx, p.dy, p.dz, p.dt = dx, dy, dz, dt
        p.solver = SOLVERS[solver] if isinstance(solver, str) else int(solver)
        p.pusher = PUSHERS[pusher] if isinstance(pusher, str) else int(pusher)
        p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        p.deposit = DEPOSITS[deposit] if isinstance(deposit, str) else int(deposit)
        p.qed = int(bool(qed))
        p.sort_every = sort_every
        p.track_ids = int(bool(track_ids))
        p.device = device
        p.e_s, p.external_bz, p.q_vac = e_s, external_bz, q_vac
        p.seed = seed
        p.capacity = capacity
        p.n_ion = len(ion_cmr)
        for n, v in enumerate(ion_cmr):
            p.ion_cmr[n] = v
        p.norm_n, p.norm_ef, p.norm_ep = norm_n, norm_ef, norm_ep
        p.merging = {None: 0, "off": 0, "ti": 1, "nl": 2}[merging]
        p.crnp = crnp
        p.sort_fallback = sort_fallback
        self.params = p
        self.shape = (nx, ny, nz)
        self.precision = precision
        h = C.c_void_p()
        _check(L.qk_create(C.byref(p), C.byref(h)), "qk_create")
        self._h = h

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _check(load_library().qk_destroy(self._h), "qk_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- fields -------------------------------------------------------------
    def set_fields(self, E=None, B=None):
        e = np.ascontiguousarray(E, np.float64) if E is not None else None
        b = np.ascontiguousarray(B, np.float64) if B is not None else None
        for a in (e, b):
            if a is not None and a.shape != self.shape + (3,):
                raise ValueError(f"field shape {a.shape} != {self.shape + (3,)}")
        _check(load_library().qk_set_fields(self._h, _dp(e) if e is not None else None,
                                            _dp(b) if b is not None else None), "qk_set_fields")

    def get_fields(self):
        """Return (E, B, J, Be) as (nx, ny, nz, 3) FP64 arrays."""
        out = [np.zeros(self.shape + (3,)) for _ in range(4)]
        _check(load_library().qk_get_fields(self._h, *[_dp(a) for a in out]), "qk_get_fields")
        return tuple(out)

    def init_laser_cos(self, a0y, a0z, xsigma, ysigma, zsigma, x0, b_sign=1, x1=0.0, phase=0.0, y0=0.0, z0=0.0,
                       append=True, phi=0.0, xtarget=0.0, ytarget=0.0, ztarget=0.0):
        L = np.array([a0y, a0z, xsigma, ysigma, zsigma, x0, float(b_sign), x1, phase, y0, z0, float(append), phi,
                      xtarget, ytarget, ztarget], np.float64)
        _check(load_library().qk_init_laser_cos(self._h, _dp(L)), "qk_init_laser_cos")

    def add_laser_source(self, side, plane, a0y, a0z, xsigma, ysigma, zsigma, x0, b_sign=1, x1=0.0, phase=0.0,
                         y0=0.0, z0=0.0, phi=0.0, xtarget=0.0, ytarget=0.0, ztarget=0.0):
        """A laser entering through plane `plane` (side 0: x-min, total field at
        i >= plane; side 1: x-max, total field at i < plane); the pulse is
        that of init_laser_cos with the same arguments, at the position x0
        (etc.) when the next step starts."""
        L = np.array([a0y, a0z, xsigma, ysigma, zsigma, x0, float(b_sign), x1, phase, y0, z0, 1.0, phi,
                      xtarget, ytarget, ztarget], np.float64)
        _check(load_library().qk_add_laser_source(self._h, _dp(L), int(side), int(plane)), "qk_add_laser_source")

    def init_boundaries(self):
        _check(load_library().qk_init_boundaries(self._h), "qk_init_boundaries")

    # -- particles ----------------------------------------------------------
    def add_particles(self, species, x, y, z, ux, uy, uz, q, ids=None):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, ux, uy, uz, q)]
        n = len(arrs[0])
        idv = np.ascontiguousarray(ids, np.int32) if ids is not None else None
        kept = C.c_int64()
        _check(load_library().qk_add_particles(self._h, species, n, *[_dp(a) for a in arrs],
                                               _ip(idv) if idv is not None else None, C.byref(kept)),
               "qk_add_particles")
        return kept.value

    def add_records(self, rec, ids=None, species_of_cmr=None):
        """Add rows x y z ux uy uz q cmr, choosing the species from cmr."""
        rec = np.asarray(rec, np.float64)
        kept = 0
        for cmr in np.unique(rec[:, 7]):
            if species_of_cmr is not None:
                s = species_of_cmr(cmr)
            elif cmr == -1:
                s = ELECTRON
            elif cmr == 1:
                s = POSITRON
            elif cmr == 0:
                s = PHOTON
            else:
                cm = list(self.params.ion_cmr[:self.params.n_ion])
                s = ION0 + cm.index(cmr)
            sel = rec[:, 7] == cmr
            kept += self.add_particles(s, *[rec[sel, c] for c in range(7)],
                                       ids=ids[sel] if ids is not None else None)
        return kept

    def load_uniform_plasma(self, species, ppc, density, sigma_u, seed, ilo=1, ihi=1 << 30):
        added = C.c_int64()
        _check(load_library().qk_load_uniform_plasma(self._h, species, ppc, density, sigma_u, seed, ilo, ihi,
                                                     C.byref(added)), "qk_load_uniform_plasma")
        return added.value

    def count(self, species):
        n = C.c_int64()
        _check(load_library().qk_particle_count(self._h, species, C.byref(n)), "qk_particle_count")
        return n.value

    def particles(self, species):
        """Return (rec[n,10] = x y z ux uy uz g q cmr chi, ids[n])."""
        cap = max(self.count(species), 1)
        rec = np.zeros((cap, 10))
        ids = np.zeros(cap, np.int32)
        n = C.c_int64()
        _check(load_library().qk_get_particles(self._h, species, cap, _dp(rec), _ip(ids), C.byref(n)),
               "qk_get_particles")
        return rec[:n.value], ids[:n.value]

    def compute_rho(self):
        rho = np.zeros(self.shape + (3,))
        _check(load_library().qk_compute_rho(self._h, _dp(rho)), "qk_compute_rho")
        return rho

    def compute_rho_noalloc(self):
        """compute_rho through qk_compute_rho_noalloc: no device allocation
        (the library's fixed scratch; grids of at most 2**21 cells)."""
        rho = np.zeros(self.shape + (3,))
        _check(load_library().qk_compute_rho_noalloc(self._h, _dp(rho)), "qk_compute_rho_noalloc")
        return rho

    def pmerging(self, mode, ppd):
        """spatial_region::pmerging: mode "ti" or "nl", deletion fractions
        ppd for electrons, positrons, photons and each ion population."""
        a = np.zeros(MAX_SPECIES)
        a[:len(ppd)] = ppd
        _check(load_library().qk_pmerging(self._h, 1 if mode == "ti" else 2, _dp(a)), "qk_pmerging")

    # -- moving window ------------------------------------------------------
    def moving_window(self):
        """spatial_region::moving_window: shift fields and particles one cell toward -x."""
        _check(load_library().qk_moving_window(self._h), "qk_moving_window")

    def max_w(self, ilo, ihi):
        """spatial_region::get_max_w over the planes [ilo, ihi)."""
        v = C.c_double()
        _check(load_library().qk_max_w(self._h, ilo, ihi, C.byref(v)), "qk_max_w")
        return v.value

    def fill_cells(self, species, cells, density, npic, ux0=0.0, first_id=0):
        """fill_cell_by_particles (T = 0) on listed cells (int (m, 3)) with densities (m,)."""
        cells = np.ascontiguousarray(cells, np.int32)
        density = np.ascontiguousarray(density, np.float64)
        npic = np.ascontiguousarray(npic, np.int32)
        added = C.c_int64()
        _check(load_library().qk_fill_cells(self._h, species, len(cells), _ip(cells), _dp(density), _ip(npic),
                                            ux0, first_id, C.byref(added)), "qk_fill_cells")
        return added.value

    # -- adjoint ------------------------------------------------------------
    def tape_begin(self, every):
        """Record the following steps for adjoint(): a snapshot now and every
        `every` steps.  With the tiled deposit the particles are first sorted
        into tiles, which permutes their slots; tape_ids() gives the order."""
        _check(load_library().qk_tape_begin(self._h, every), "qk_tape_begin")
        self._tape_n = {s: self.count(s) for s in range(ION0 + self.params.n_ion)}
        self._tape_ids = {s: self.particles(s)[1] for s, n in self._tape_n.items() if n > 0} \
            if self.params.track_ids else {}

    def tape_ids(self, species):
        """Ids of the particles in slot order after tape_begin (needs track_ids)."""
        return self._tape_ids[species]

    def adjoint(self, lam_E=None, lam_B=None, lam_parts=None):
        """Adjoint of the recorded steps.  lam_E, lam_B: (nx, ny, nz, 3) adjoints
        of the final fields; lam_parts: {species: (n, 7)} adjoints of the final
        x y z ux uy uz q of the n slots the species had after tape_begin, in
        slot order (tape_ids; the order added with the global deposit; zero
        for deleted particles).  Returns (grad_E0, grad_B0, {species: (n, 7)})
        at tape_begin, in the same order."""
        L = load_library()
        shape = self.shape + (3,)
        lE = np.ascontiguousarray(lam_E, np.float64) if lam_E is not None else None
        lB = np.ascontiguousarray(lam_B, np.float64) if lam_B is not None else None
        gE, gB = np.zeros(shape), np.zeros(shape)
        lp = (C.POINTER(C.c_double) * MAX_SPECIES)()
        gp = (C.POINTER(C.c_double) * MAX_SPECIES)()
        keep, grads = [], {}
        for s, n in self._tape_n.items():
            if n == 0:
                continue
            if lam_parts is not None and s in lam_parts:
                a = np.ascontiguousarray(lam_parts[s], np.float64)
                if a.shape != (n, 7):
                    raise ValueError(f"species {s}: adjoint shape {a.shape} != {(n, 7)}")
                keep.append(a)
                lp[s] = _dp(a)
            grads[s] = np.zeros((n, 7))
            gp[s] = _dp(grads[s])
        _check(L.qk_adjoint(self._h, _dp(lE) if lE is not None else None, _dp(lB) if lB is not None else None, lp,
                            _dp(gE), _dp(gB), gp), "qk_adjoint")
        return gE, gB, grads

    def tangent_begin(self, dE=None, dB=None, dparts=None):
        """Seed the tangent-linear run right after tape_begin: dE, dB
        (nx, ny, nz, 3) perturb the fields as set before init_boundaries;
        dparts {species: (n, 7)} the x y z ux uy uz q of the tape's slots."""
        keep = []

        def arr(a, shape):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64)
            if a.shape != shape:
                raise ValueError(f"tangent seed shape {a.shape} != {shape}")
            keep.append(a)
            return _dp(a)

        shape = self.shape + (3,)
        dp_ = (C.POINTER(C.c_double) * MAX_SPECIES)()
        for s, n in self._tape_n.items():
            if dparts is not None and s in dparts and n > 0:
                dp_[s] = arr(dparts[s], (n, 7))
        _check(load_library().qk_tangent_begin(self._h, arr(dE, shape), arr(dB, shape), dp_), "qk_tangent_begin")

    def tangent_get(self):
        """(dE, dB, {species: (n, 7)}): the tangent of the current state."""
        shape = self.shape + (3,)
        dE, dB = np.zeros(shape), np.zeros(shape)
        gp = (C.POINTER(C.c_double) * MAX_SPECIES)()
        out = {}
        for s, n in self._tape_n.items():
            if n > 0:
                out[s] = np.zeros((n, 7))
                gp[s] = _dp(out[s])
        _check(load_library().qk_tangent_get(self._h, _dp(dE), _dp(dB), gp), "qk_tangent_get")
        return dE, dB, out

    # -- stepping -----------------------------------------------------------
    def step(self, nsteps=1, diag=False):
        d = (Diag * max(nsteps, 1))() if diag else None
        _check(load_library().qk_step(self._h, nsteps, d), "qk_step")
        return [d[i].as_dict() for i in range(nsteps)] if diag else None

    def sync(self):
        _check(load_library().qk_sync(self._h), "qk_sync")

    def timing(self):
        """(ms_total, ms_particles, ms_fields) of the last step() call, from HIP events."""
        t = [C.c_double(), C.c_double(), C.c_double()]
        _check(load_library().qk_last_timing(self._h, *[C.byref(v) for v in t]), "qk_last_timing")
        return tuple(v.value for v in t)

    def fallback_fraction(self):
        """Share of the last step() call's pushes that deposited through global atomics."""
        v = C.c_double()
        _check(load_library().qk_last_fallback(self._h, C.byref(v)), "qk_last_fallback")
        return v.value

    def stream(self):
        s = C.c_void_p()
        _check(load_library().qk_get_stream(self._h, C.byref(s)), "qk_get_stream")
        return s.value
