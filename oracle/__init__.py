"""ctypes binding of the FP64 oracle (oracle/quill_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of bench.py, as the checker.  The product never imports
this package.

Field arrays are numpy views of the oracle's AoS storage, shape
(nx, ny, nz, 3), k fastest, as the reference's ``field3d`` (containers.h:34-52).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "quill_ref")

FDTD, NDFX, FIVE_POINT = 0, 1, 2
VAY, BORIS = 0, 1
E, B, J, BE = 0, 1, 2, 3


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
                ("solver", C.c_int), ("pusher", C.c_int), ("qed", C.c_int),
                ("e_s", C.c_double), ("external_bz", C.c_double), ("q_vac", C.c_double),
                ("seed", C.c_uint64),
                ("norm_n", C.c_double), ("norm_ef", C.c_double), ("norm_ep", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("N_e", C.c_double), ("N_p", C.c_double), ("N_ph", C.c_double),
                ("n_e", C.c_int64), ("n_p", C.c_int64), ("n_ph", C.c_int64), ("n_ion", C.c_int64),
                ("energy_f", C.c_double), ("energy_e", C.c_double), ("energy_p", C.c_double),
                ("energy_ph", C.c_double), ("energy_ion", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.oq_create.restype = vp
        L.oq_create.argtypes = [C.POINTER(Params)]
        L.oq_destroy.argtypes = [vp]
        L.oq_field.restype = dp
        L.oq_field.argtypes = [vp, C.c_int]
        L.oq_add_particles.restype = C.c_int64
        L.oq_add_particles.argtypes = [vp, C.c_int64] + [dp] * 7 + [C.c_double, ip]
        L.oq_count.restype = C.c_int64
        L.oq_count.argtypes = [vp]
        L.oq_get_particles.argtypes = [vp, dp, ip]
        for name in ("oq_init_boundaries", "oq_interpolate_be", "oq_maxwell", "oq_padvance",
                     "oq_birth_from_vacuum", "oq_compute_rho"):
            getattr(L, name).argtypes = [vp]
        L.oq_diagnostics.argtypes = [vp, C.POINTER(Diag)]
        L.oq_step.argtypes = [vp, C.c_int, C.POINTER(Diag)]
        L.oq_init_laser_cos.argtypes = [vp] + [C.c_double] * 6 + [C.c_int] + [C.c_double] * 4 + \
            [C.c_int] + [C.c_double] * 4
        L.oq_load_uniform.restype = C.c_int64
        L.oq_load_uniform.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_int, C.c_int]
        L.oq_synth_plasma.restype = C.c_int64
        L.oq_synth_plasma.argtypes = [C.c_int] * 4 + [C.c_double] * 2 + [C.c_uint64, C.c_int, C.c_int, dp, ip]
        L.oq_push.argtypes = [C.c_int, C.c_double, C.c_double, dp, dp, dp, dp]
        L.oq_chi.restype = C.c_double
        L.oq_chi.argtypes = [C.c_double, dp, dp, dp, C.c_double]
        L.oq_rate_w.restype = C.c_double
        L.oq_rate_w.argtypes = [C.c_double, C.c_double, dp, C.c_double]
        L.oq_rate_tilde_w.restype = C.c_double
        L.oq_rate_tilde_w.argtypes = [C.c_double] * 4
        L.oq_pmerging.argtypes = [vp, C.c_int, dp, C.c_int, dp]
        L.oq_moving_window.argtypes = [vp]
        L.oq_max_w.restype = C.c_double
        L.oq_max_w.argtypes = [vp, C.c_int, C.c_int]
        L.oq_fill_cells.restype = C.c_int64
        L.oq_fill_cells.argtypes = [vp, C.c_double, C.c_int64, ip, dp, ip, C.c_double, C.c_int32]
        L.oq_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.oq_add_laser_source.restype = C.c_int
        L.oq_add_laser_source.argtypes = [vp, dp, C.c_int, C.c_int]
        L.oq_laser_field.argtypes = [vp, dp, C.c_double, C.c_int, dp, dp]
        L.oq_vacuum_rate.restype = C.c_double
        L.oq_vacuum_rate.argtypes = [C.c_double, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def philox(ctr, key):
    """Philox4x32-10 of the loader on one block: 4 counter words, 2 key words."""
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    lib().oq_philox(c, k, out)
    return list(out)


def synth_plasma(nx, ny, nz, ppc, density, sigma_u, seed, ilo=1, ihi=1 << 30):
    """Shared Philox loader: returns (rec[n,8] = x,y,z,ux,uy,uz,q,cmr ; ids[n])."""
    L = lib()
    n = L.oq_synth_plasma(nx, ny, nz, ppc, density, sigma_u, seed, ilo, ihi, None, None)
    rec = np.zeros((max(n, 1), 8))
    ids = np.zeros(max(n, 1), np.int32)
    L.oq_synth_plasma(nx, ny, nz, ppc, density, sigma_u, seed, ilo, ihi, _dp(rec), _ip(ids))
    return rec[:n], ids[:n]


class Oracle:
    """One subdomain of the FP64 restatement."""

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, solver=FDTD, pusher=BORIS, qed=False,
                 e_s=1e10, external_bz=0.0, q_vac=0.0, seed=1, norm_n=1.0, norm_ef=1.0, norm_ep=1.0):
        self.shape = (nx, ny, nz)
        self.p = Params(nx, ny, nz, dx, dy, dz, dt, solver, pusher, int(qed), e_s, external_bz,
                        q_vac, seed, norm_n, norm_ef, norm_ep)
        self.h = lib().oq_create(C.byref(self.p))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oq_destroy(self.h)
            self.h = None

    def field(self, which):
        ptr = lib().oq_field(self.h, which)
        n = self.shape[0] * self.shape[1] * self.shape[2] * 3
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.shape + (3,))

    def add_particles(self, rec, ids=None):
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        kept = 0
        for cmr in np.unique(rec[:, 7]):
            sel = rec[:, 7] == cmr
            cols = [np.ascontiguousarray(rec[sel, c]) for c in range(7)]
            idv = np.ascontiguousarray(ids[sel], dtype=np.int32) if ids is not None else None
            kept += lib().oq_add_particles(self.h, int(sel.sum()), *[_dp(c) for c in cols], float(cmr),
                                           _ip(idv) if idv is not None else None)
        return kept

    def count(self):
        return lib().oq_count(self.h)

    def particles(self):
        n = self.count()
        rec = np.zeros((max(n, 1), 10))
        ids = np.zeros(max(n, 1), np.int32)
        lib().oq_get_particles(self.h, _dp(rec), _ip(ids))
        return rec[:n], ids[:n]

    def init_boundaries(self):
        lib().oq_init_boundaries(self.h)

    def step(self, nsteps=1, diag=False):
        d = (Diag * max(nsteps, 1))()
        lib().oq_step(self.h, nsteps, d if diag else None)
        return [d[i] for i in range(nsteps)] if diag else None

    def diagnostics(self):
        d = Diag()
        lib().oq_diagnostics(self.h, C.byref(d))
        return d

    def compute_rho(self):
        lib().oq_compute_rho(self.h)
        return self.field(J).copy()

    def laser_cos(self, a0y, a0z, xsigma, ysigma, zsigma, x0, b_sign=1, x1=0.0, phase=0.0, y0=0.0, z0=0.0,
                  append=True, phi=0.0, xtarget=0.0, ytarget=0.0, ztarget=0.0):
        lib().oq_init_laser_cos(self.h, a0y, a0z, xsigma, ysigma, zsigma, x0, int(b_sign), x1, phase, y0, z0,
                                int(append), phi, xtarget, ytarget, ztarget)

    def laser_source(self, L, side, plane):
        """A laser entering through plane `plane` (side 0: total field at
        i >= plane, side 1: at i < plane); L as laser_cos's 16 arguments."""
        L = np.ascontiguousarray(L, np.float64)
        if lib().oq_add_laser_source(self.h, _dp(L), int(side), int(plane)) != 0:
            raise ValueError("oq_add_laser_source: bad arguments")

    def laser_field(self, L, shift, staggered=True):
        """(E, B) of the pulse L after travelling `shift` along its direction."""
        L = np.ascontiguousarray(L, np.float64)
        E = np.zeros(self.shape + (3,))
        B = np.zeros(self.shape + (3,))
        lib().oq_laser_field(self.h, _dp(L), float(shift), int(staggered), _dp(E), _dp(B))
        return E, B

    def pmerging(self, mode, ppd, ion_cmr=()):
        """One merge (mode "ti" or "nl") with deletion fractions ppd[3 + n_ion]."""
        ppd = np.ascontiguousarray(ppd, np.float64)
        icmr = np.ascontiguousarray(list(ion_cmr) + [0.0], np.float64)
        lib().oq_pmerging(self.h, 1 if mode == "ti" else 2, _dp(ppd), len(ion_cmr), _dp(icmr))

    def moving_window(self):
        lib().oq_moving_window(self.h)

    def max_w(self, ilo, ihi):
        return lib().oq_max_w(self.h, ilo, ihi)

    def fill_cells(self, cmr, cells, density, npic, ux0=0.0, first_id=0):
        """fill_cell_by_particles on listed cells (int (m, 3)), density (m,)."""
        cells = np.ascontiguousarray(cells, np.int32)
        density = np.ascontiguousarray(density, np.float64)
        npic = np.ascontiguousarray(npic, np.int32)
        return lib().oq_fill_cells(self.h, cmr, len(cells), _ip(cells), _dp(density), _ip(npic), ux0, first_id)

    def load_uniform(self, ppc, density, sigma_u, seed, ilo=1, ihi=1 << 30):
        return lib().oq_load_uniform(self.h, ppc, density, sigma_u, seed, ilo, ihi)


def push(pusher, cmr, dt, e, b, u, g):
    e = np.ascontiguousarray(e, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    u = np.array(u, np.float64)
    gg = np.array([g], np.float64)
    lib().oq_push(pusher, cmr, dt, _dp(e), _dp(b), _dp(u), _dp(gg))
    return u, gg[0]
