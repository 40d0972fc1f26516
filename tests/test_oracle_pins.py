"""Pins of the oracle functions the reference's own outputs do not cover.

* Philox4x32-10 (oracle/philox.h, the loader's and the QED streams'
  generator) against the published known-answer vectors of Salmon, Moraes,
  Dror and Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11),
  distributed with their Random123 library (kat_vectors: zero, all-ones and
  pi-digit counter/key).
* The synthetic plasma loader (oq_synth_plasma / oq_load_uniform) against its
  definition: per-cell counts and the moments of offsets and momenta
  (tests/plasma_checks.py), at a million particles.
* oq_compute_rho against the reference's own compute_rho output
  (rhodeposition, src/deposition.cpp:358-429, via src/fout.cpp:7-39) held in
  tests/golden/fout.npz.
"""
import numpy as np
import pytest

import oracle
from plasma_checks import check_uniform_plasma

# Random123 kat_vectors, philox4x32 with 10 rounds: counter, key -> output
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    assert tuple(oracle.philox(ctr, key)) == expect


GRID = (53, 53, 53)   # 50^3 is_inside cells
PPC, DENSITY, SIGMA, SEED = 8, 0.01, 0.01, 20261202


@pytest.fixture(scope="module")
def plasma():
    rec, ids = oracle.synth_plasma(*GRID, PPC, DENSITY, SIGMA, SEED)
    return rec, ids


def test_loader_statistics(plasma):
    rec, ids = plasma
    assert len(rec) == 1_000_000
    check_uniform_plasma(rec[:, :3], rec[:, 3:6], rec[:, 6], GRID, PPC, DENSITY, SIGMA)
    assert (rec[:, 7] == -1).all()
    # ids are linear_cell * ppc + slot, each once
    cell = np.floor(rec[:, :3]).astype(np.int64)
    lin = (cell[:, 0] * GRID[1] + cell[:, 1]) * GRID[2] + cell[:, 2]
    assert np.array_equal(ids // PPC, lin)
    assert len(np.unique(ids)) == len(ids)


def test_loader_checks_catch_mistakes(plasma):
    """The statistical bars are tight enough to see a 1 % error in sigma, a
    1 % squeeze of the placement, or a missing cell."""
    rec, _ = plasma
    pos, u, q = rec[:, :3], rec[:, 3:6], rec[:, 6]
    with pytest.raises(AssertionError):
        check_uniform_plasma(pos, u * 1.01, q, GRID, PPC, DENSITY, SIGMA)
    cell = np.floor(pos)
    with pytest.raises(AssertionError):
        check_uniform_plasma(cell + (pos - cell) * 0.99, u, q, GRID, PPC, DENSITY, SIGMA)
    with pytest.raises(AssertionError):
        check_uniform_plasma(pos[PPC:], u[PPC:], q[PPC:], GRID, PPC, DENSITY, SIGMA)


def test_loader_ranges_and_seed():
    """ilo/ihi select x-slabs of the same stream; another seed changes it."""
    nx, ny, nz = 20, 9, 8
    full, fid = oracle.synth_plasma(nx, ny, nz, 3, 0.5, 0.2, 7)
    part, pid = oracle.synth_plasma(nx, ny, nz, 3, 0.5, 0.2, 7, ilo=5, ihi=9)
    sel = (np.floor(full[:, 0]) >= 5) & (np.floor(full[:, 0]) < 9)
    assert np.array_equal(full[sel], part) and np.array_equal(fid[sel], pid)
    other, _ = oracle.synth_plasma(nx, ny, nz, 3, 0.5, 0.2, 8)
    assert not np.allclose(other[:, 3:6], full[:, 3:6])
    o = oracle.Oracle(nx, ny, nz, 0.3, 0.3, 0.3, 0.1)
    assert o.load_uniform(3, 0.5, 0.2, 7) == len(full)
    got, gid = o.particles()
    order = np.argsort(gid)
    assert np.array_equal(gid[order], np.sort(fid))
    ref = full[np.argsort(fid)]
    got = got[order]
    assert np.array_equal(got[:, :6], ref[:, :6]) and np.array_equal(got[:, 7], ref[:, 6])
    assert np.allclose(got[:, 6], np.sqrt(1 + (ref[:, 3:6] ** 2).sum(1)), rtol=1e-15)


def test_compute_rho_matches_reference():
    """oq_compute_rho on the fout case's particles against the reference's
    compute_rho (electrons, positrons and photons in cj.x / y / z)."""
    import make_golden as mg

    g = np.load(oracle_golden("fout.npz"))
    nx, ny, nz = g["E"].shape[:3]
    rec, ids, _ = mg.mixed_particles(nx, ny, nz, 55, 0.3)
    dx = 0.05 * 2 * np.pi
    o = oracle.Oracle(nx, ny, nz, dx, dx, dx, 0.5 * dx)
    o.add_particles(rec, ids)
    rho = o.compute_rho()
    ref = g["rho"]
    for c in range(3):
        assert np.abs(ref[..., c]).max() > 0
        err = np.abs(rho[..., c] - ref[..., c]).max() / np.abs(ref[..., c]).max()
        assert err < 1e-12, (c, err)


def oracle_golden(name):
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)
