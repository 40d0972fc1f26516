"""Case-file I/O for the compiled reference harness and the committed fixtures.

Mirrors oracle/golden_format.h (keep the two in step).  Fixtures are .npz
files holding the inputs and the reference's outputs of one case.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

MAGIC = 0x4C4C5551
MODE_RUN, MODE_LASER, MODE_RATES, MODE_PUSH, MODE_TIME, MODE_FOUT, MODE_MERGE, MODE_MW, MODE_FILM = range(9)
DIAG_WIDTH = 12
DIAG_NAMES = ["N_e", "N_p", "N_ph", "n_e", "n_p", "n_ph", "energy_f", "energy_e", "energy_p",
              "energy_ph", "energy_ion", "n_ion"]

HDR = np.dtype([("magic", "<i4"), ("mode", "<i4"), ("nx", "<i4"), ("ny", "<i4"), ("nz", "<i4"),
                ("solver", "<i4"), ("pusher", "<i4"), ("qed", "<i4"), ("nsteps", "<i4"),
                ("n_particles", "<i4"), ("seed", "<i4"), ("n_ion", "<i4"),
                ("dx", "<f8"), ("dy", "<f8"), ("dz", "<f8"), ("dt", "<f8"),
                ("e_s", "<f8"), ("external_bz", "<f8"), ("q_vac", "<f8"),
                ("norm_n", "<f8"), ("norm_ef", "<f8"), ("norm_ep", "<f8"),
                ("laser", "<f8", (16,)), ("ion_cmr", "<f8", (4,)), ("extra", "<f8", (8,))])
assert HDR.itemsize == 352


def header(**kw):
    h = np.zeros((), HDR)
    h["magic"] = MAGIC
    h["norm_n"] = h["norm_ef"] = h["norm_ep"] = 1.0
    h["e_s"] = 1e10
    for k, v in kw.items():
        h[k] = v
    return h


def header_dict(h):
    return {name: (h[name].tolist()) for name in HDR.names}


def write_input(path, h, *arrays):
    with open(path, "wb") as f:
        f.write(h.tobytes())
        for a in arrays:
            f.write(np.ascontiguousarray(a).tobytes())


def read_output_run(path, h):
    nx, ny, nz = int(h["nx"]), int(h["ny"]), int(h["nz"])
    nf = nx * ny * nz * 3
    with open(path, "rb") as f:
        f.read(HDR.itemsize)
        fields = [np.frombuffer(f.read(nf * 8), "<f8").reshape(nx, ny, nz, 3) for _ in range(4)]
        n = int(np.frombuffer(f.read(8), "<i8")[0])
        rec = np.frombuffer(f.read(n * 80), "<f8").reshape(n, 10)
        ids = np.frombuffer(f.read(n * 4), "<i4")
        diag = np.frombuffer(f.read(int(h["nsteps"]) * DIAG_WIDTH * 8), "<f8").reshape(-1, DIAG_WIDTH)
    return fields, rec, ids, diag


def read_output_mw(path, h):
    """RUN output followed by the per-step (nmw, ratio) trace."""
    nx, ny, nz = int(h["nx"]), int(h["ny"]), int(h["nz"])
    nf = nx * ny * nz * 3
    nsteps = int(h["nsteps"])
    with open(path, "rb") as f:
        f.read(HDR.itemsize)
        fields = [np.frombuffer(f.read(nf * 8), "<f8").reshape(nx, ny, nz, 3) for _ in range(4)]
        n = int(np.frombuffer(f.read(8), "<i8")[0])
        rec = np.frombuffer(f.read(n * 80), "<f8").reshape(n, 10)
        ids = np.frombuffer(f.read(n * 4), "<i4")
        diag = np.frombuffer(f.read(nsteps * DIAG_WIDTH * 8), "<f8").reshape(-1, DIAG_WIDTH)
        trace = np.frombuffer(f.read(nsteps * 16), "<f8").reshape(-1, 2)
    return fields, rec, ids, diag, trace


def read_output_fields(path, h):
    nx, ny, nz = int(h["nx"]), int(h["ny"]), int(h["nz"])
    nf = nx * ny * nz * 3
    with open(path, "rb") as f:
        f.read(HDR.itemsize)
        return [np.frombuffer(f.read(nf * 8), "<f8").reshape(nx, ny, nz, 3) for _ in range(2)]


def read_output_table(path, width):
    with open(path, "rb") as f:
        f.read(HDR.itemsize)
        return np.frombuffer(f.read(), "<f8").reshape(-1, width)


def load(name):
    """Load a committed fixture as a dict of arrays (header fields unpacked)."""
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}
