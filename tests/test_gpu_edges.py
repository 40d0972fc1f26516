"""Empty and vacuum inputs on the GPU, through the C ABI.

* vacuum: a tilted, rotated f_init_cos pulse (src/init.cpp:6-210) and no
  particles, 40 steps of each solver against the oracle (FP64 1e-10, FP32
  1e-4 relative L2 per component);
* empty inputs: zero particles added, every species empty while the step,
  the sort, compute_rho and the particle readback run;
* a species emptied by the boundaries (every particle leaves is_inside and
  is deleted, src/plist.cpp:7-40), stepped on while empty, then refilled and
  compared with the oracle on the same history.
"""
import numpy as np
import pytest

import oracle
import quill_amd as qa

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi
DX = 0.05 * TWO_PI
SOLVERS = {"fdtd": (oracle.FDTD, 0.5), "ndfx": (oracle.NDFX, 0.5), "five_point": (oracle.FIVE_POINT, 0.35)}
PULSE = dict(a0y=1.0, a0z=0.5, xsigma=0.4 * TWO_PI, ysigma=0.3 * TWO_PI, zsigma=0.3 * TWO_PI, x0=0.9 * TWO_PI,
             b_sign=1, x1=-1.0, phase=0.3, y0=0.2, z0=-0.1, phi=0.4)


def oracle_region(shape, solver, dtf):
    code = SOLVERS[solver][0]
    return oracle.Oracle(*shape, DX, DX, DX, dtf * DX, solver=code, pusher=oracle.BORIS)


@pytest.mark.parametrize("solver", sorted(SOLVERS))
def test_vacuum_fields_match_oracle(solver):
    shape = (36, 28, 24)
    dtf = SOLVERS[solver][1]
    o = oracle_region(shape, solver, dtf)
    o.laser_cos(**PULSE)
    o.init_boundaries()
    o.step(40)
    ref = (o.field(oracle.E), o.field(oracle.B))
    assert np.abs(ref[0]).max() > 0.1
    for precision in ("fp64", "fp32"):
        print(f"vacuum {solver} {precision}", flush=True)
        with qa.Region(*shape, DX, DX, DX, dtf * DX, solver=solver, pusher="boris", precision=precision) as r:
            r.init_laser_cos(**PULSE)
            r.init_boundaries()
            r.step(40)
            E, B, J, _ = r.get_fields()
        assert not J.any()
        for name, a, b in zip("EB", (E, B), ref):
            if precision == "fp64":
                err = np.abs(a - b).max() / np.abs(b).max()
                assert err < 1e-10, (name, err)
            else:
                for c in range(3):
                    err = np.linalg.norm(a[..., c] - b[..., c]) / np.linalg.norm(b[..., c])
                    assert err < 1e-4, (name, c, err)


@pytest.mark.parametrize("deposit", ["tiled", "global"])
def test_empty_inputs(deposit):
    shape = (20, 18, 16)
    print(f"empty inputs, {deposit}", flush=True)
    with qa.Region(*shape, DX, DX, DX, 0.5 * DX, precision="fp32", deposit=deposit, sort_every=1,
                   track_ids=True) as r:
        empty = np.zeros(0)
        assert r.add_particles(qa.ELECTRON, empty, empty, empty, empty, empty, empty, empty) == 0
        # placed outside is_inside only: nothing is kept
        x = np.array([0.5, 18.5, 5.5])
        y = np.array([5.5, 5.5, 17.2])
        assert r.add_particles(qa.POSITRON, x, y, y, 0 * x, 0 * x, 0 * x, 1e-3 + 0 * x) == 0
        r.init_boundaries()
        d = r.step(5, diag=True)
        assert all(s["n_e"] == 0 and s["n_p"] == 0 and s["energy_e"] == 0 for s in d)
        for s in (qa.ELECTRON, qa.POSITRON):
            assert r.count(s) == 0
            rec, ids = r.particles(s)
            assert rec.shape == (0, 10) and ids.shape == (0,)
        assert not r.compute_rho().any()
        E, B, J, Be = r.get_fields()
        assert not (E.any() or B.any() or J.any() or Be.any())


@pytest.mark.parametrize("deposit", ["tiled", "global"])
def test_species_emptied_then_refilled(deposit):
    shape = (20, 16, 16)
    dtf = 0.5
    rng = np.random.default_rng(11)
    # fast electrons in the last is_inside cell along x (17 of 20), 0.5 cells
    # per step: all leave within 3 steps
    n = 200
    first = np.zeros((n, 8))
    first[:, 0] = rng.uniform(17.0, 17.99, n)
    first[:, 1:3] = rng.uniform(2.0, 13.0, (n, 2))
    first[:, 3] = 20.0
    first[:, 4:6] = rng.normal(0, 0.5, (n, 2))
    first[:, 6] = -1e-3
    first[:, 7] = -1.0
    ids1 = np.arange(n, dtype=np.int32)
    second, ids2 = oracle.synth_plasma(*shape, 2, 0.01, 0.3, 5)
    ids2 = ids2 + 1000
    o = oracle_region(shape, "fdtd", dtf)
    o.add_particles(first, ids1)
    o.init_boundaries()
    o.step(4)
    assert o.count() == 0
    o.step(2)
    o.add_particles(second, ids2)
    o.step(6)
    ref_E, ref_J = o.field(oracle.E).copy(), o.field(oracle.J).copy()
    ref_rec, ref_ids = o.particles()
    print(f"emptied species, {deposit}", flush=True)
    with qa.Region(*shape, DX, DX, DX, dtf * DX, solver="fdtd", pusher="boris", precision="fp64", deposit=deposit,
                   sort_every=2, track_ids=True) as r:
        r.add_records(first, ids1)
        r.init_boundaries()
        r.step(4)
        assert r.count(qa.ELECTRON) == 0
        r.step(2)
        r.add_records(second, ids2)
        r.step(6)
        E, _, J, _ = r.get_fields()
        rec, ids = r.particles(qa.ELECTRON)
    assert np.abs(E - ref_E).max() < 1e-10 * np.abs(ref_E).max()
    assert np.abs(J - ref_J).max() < 1e-10 * np.abs(ref_J).max()
    a, b = np.argsort(ids), np.argsort(ref_ids)
    np.testing.assert_array_equal(ids[a], ref_ids[b])
    np.testing.assert_allclose(rec[a, :3], ref_rec[b, :3], rtol=0, atol=1e-10)
