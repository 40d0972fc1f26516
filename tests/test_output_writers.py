"""quillk-run's output files against the reference's own writers.

tests/golden/fout.npz holds the plane files that the reference's
spatial_region writers (fout_ex .. fout_bz, fout_w, fout_inv, fout_rho and
their yz-plane parts, src/fout.cpp:180-660, driven as write_fields /
write_density do for one subdomain, src/main.cpp:369-476, 745-995) produced
from given fields and particles, in text and in binary output mode.  The
driver's writers, given the same E, B, Be and rho arrays, must produce the
same bytes.  No GPU is involved.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(ROOT, "quill_amd", "bin", "quillk-run")

import golden_io as gio  # noqa: E402

FILES = ["ex", "ey", "ez", "bx", "by", "bz", "w", "inv", "rho", "rho_p", "rho_ph"]


@pytest.mark.parametrize("mode", ["text", "bin"])
def test_plane_files_match_reference_writers(tmp_path, mode):
    if not os.path.exists(BIN):
        pytest.skip("quillk-run not built (make -C quill_amd/driver)")
    fx = gio.load("fout")
    nx, ny, nz, _ = fx["E"].shape
    inp = tmp_path / "planes.in"
    with open(inp, "wb") as f:
        f.write(np.array([nx, ny, nz, int(fx["i_yz"]), mode == "bin"], np.int32).tobytes())
        for key in ("E", "B", "Be", "rho"):
            f.write(np.ascontiguousarray(fx[key], np.float64).tobytes())
    out = tmp_path / "out"
    r = subprocess.run([BIN, "--write-planes", str(inp), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    for name in FILES:
        mine = (out / name).read_bytes()
        ref = bytes(fx[f"{mode}_{name}"])
        assert mine == ref, f"{mode} {name}: {len(mine)} bytes vs reference {len(ref)}"
    # the rho files are not empty of particles
    assert np.abs(fx["rho"][..., 0]).max() > 0


# --- spectra and phase space (write_spectrum_phasespace, src/main.cpp:478-743)

PI = np.pi
ION_CMR = 1.0 / 1836.1526721


def write_particle_files(tmp_path, species, conf, dx=0.3, lam=1e-4, icmr=(ION_CMR,)):
    """Run the driver's writer on species = [electrons, positrons, photons,
    ions...] (each (n, 10): x y z ux uy uz g q cmr chi)."""
    inp = tmp_path / "particles.bin"
    with open(inp, "wb") as f:
        f.write(np.int32(len(icmr)).tobytes())
        f.write(np.array([conf["deps"], conf["deps_p"], conf["deps_ph"], conf["deps_i"]], "<f8").tobytes())
        f.write(np.array([conf["neps"]] * 4 + [conf["enthp"]] * 4 + [conf["write_p"], conf["write_ph"]],
                         "<i4").tobytes())
        f.write(np.array([dx, dx, dx, lam], "<f8").tobytes())
        f.write(np.array(icmr, "<f8").tobytes())
        for rec in species:
            f.write(np.int64(len(rec)).tobytes())
            f.write(np.ascontiguousarray(rec, "<f8").tobytes())
    out = tmp_path / "out"
    r = subprocess.run([BIN, "--write-particles", str(inp), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return out


def particles(rng, n, cmr, gmax, q):
    rec = np.zeros((n, 10))
    rec[:, :3] = rng.uniform(1, 9, (n, 3))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    g = rng.uniform(1.0 if cmr != 0 else 0.0, gmax, n)
    umag = np.sqrt(np.maximum(g * g - (1 if cmr != 0 else 0), 0))
    rec[:, 3:6] = u * umag[:, None]
    rec[:, 6] = g
    rec[:, 7] = q
    rec[:, 8] = cmr
    rec[:, 9] = rng.uniform(0, 1, n)
    return rec


@pytest.mark.skipif(not os.path.exists(BIN), reason="quillk-run not built")
def test_spectrum_bins_and_normalisation(tmp_path):
    """One particle per species: its charge lands in bin
    int((g - 1) 0.511 / deps) (g 0.511 / deps for photons, times the proton
    mass for ions, src/main.cpp:538, 563, 584, 611), scaled to dN/deps by
    1.11485e13 lambda dx dy dz / (8 pi^3) / deps (:690-720)."""
    conf = dict(deps=0.5, deps_p=0.5, deps_ph=0.5, deps_i=0.5, neps=20, enthp=0, write_p=1, write_ph=1)
    dx, lam = 0.3, 1e-4
    g = {0: 3.7, 1: 8.2, 2: 5.3, 3: 1.0011}
    species = []
    for s, cmr, q in ((0, -1.0, -0.02), (1, 1.0, 0.03), (2, 0.0, 0.05), (3, ION_CMR, 0.07)):
        rec = np.zeros((1, 10))
        rec[0, :3] = [2.5, 3.5, 4.5]
        rec[0, 6], rec[0, 7], rec[0, 8] = g[s], q, cmr
        species.append(rec)
    out = write_particle_files(tmp_path, species, conf, dx, lam)
    base = 1.11485e13 * lam * dx ** 3 / (8 * PI ** 3) / 0.5
    expect = {"spectrum": (int((g[0] - 1) * 0.511 / 0.5), 0.02 * base),
              "spectrum_p": (int((g[1] - 1) * 0.511 / 0.5), 0.03 * base),
              "spectrum_ph": (int(g[2] * 0.511 / 0.5), 0.05 * base),
              f"spectrum_{ION_CMR:g}_": (int((g[3] - 1) * 0.511 * 1836.1526721 / 0.5), 0.07 * base / ION_CMR)}
    for name, (b, v) in expect.items():
        vals = np.loadtxt(out / name)
        assert len(vals) == 20
        assert np.count_nonzero(vals) == 1 and np.argmax(np.abs(vals)) == b, (name, vals)
        assert abs(vals[b] - v) <= 1e-5 * v, (name, vals[b], v)


@pytest.mark.skipif(not os.path.exists(BIN), reason="quillk-run not built")
def test_spectrum_sum_rule_and_phasespace_records(tmp_path):
    rng = np.random.default_rng(8)
    conf = dict(deps=0.3, deps_p=0.3, deps_ph=0.3, deps_i=0.3, neps=40, enthp=2, write_p=1, write_ph=0)
    dx, lam = 0.3, 1e-4
    el = particles(rng, 600, -1.0, 30.0, -0.01)
    po = particles(rng, 200, 1.0, 30.0, 0.02)
    ph = particles(rng, 150, 0.0, 30.0, 0.04)
    io = particles(rng, 90, ION_CMR, 1.01, 0.05)
    out = write_particle_files(tmp_path, [el, po, ph, io], conf, dx, lam)
    base = 1.11485e13 * lam * dx ** 3 / (8 * PI ** 3)
    # sum rule: sum_i spectrum_i deps / base = sum |q| of the particles in range
    for name, rec, e_of, norm in (("spectrum", el, lambda g: (g - 1) * 0.511, 1.0),
                                  ("spectrum_p", po, lambda g: (g - 1) * 0.511, 1.0),
                                  (f"spectrum_{ION_CMR:g}_", io, lambda g: (g - 1) * 0.511 * 1836.1526721, ION_CMR)):
        vals = np.loadtxt(out / name)
        inr = e_of(rec[:, 6]) / 0.3 < 40
        assert abs(vals.sum() * 0.3 / base * norm - np.abs(rec[inr, 7]).sum()) <= 1e-5 * np.abs(rec[:, 7]).sum()
    # photons not asked for: their files exist and are empty
    assert (out / "spectrum_ph").read_text() == "" and (out / "phasespace_ph").read_text() == ""
    # phase space: every third particle in cell order, 9 lines each
    for name, rec in (("phasespace", el), ("phasespace_p", po)):
        lines = np.loadtxt(out / name)
        assert len(lines) % 9 == 0
        got = lines.reshape(-1, 9)
        cell = np.floor(rec[:, :3]).astype(np.int64)
        order = np.lexsort((cell[:, 2], cell[:, 1], cell[:, 0]))   # stable: records of a cell keep their order
        want = rec[order][2::3]
        assert len(got) == len(want) == len(rec) // 3
        np.testing.assert_allclose(got[:, 0], want[:, 7], rtol=1e-5)
        np.testing.assert_allclose(got[:, 1:4], want[:, :3] * dx / (2 * PI), rtol=1e-5)
        np.testing.assert_allclose(got[:, 4:7], want[:, 3:6], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(got[:, 7], want[:, 6], rtol=1e-5)
        np.testing.assert_allclose(got[:, 8], want[:, 9], rtol=1e-5, atol=1e-6)
