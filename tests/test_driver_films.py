"""quillk-run's films against the reference's own film() (src/init.cpp:537-646).

tests/golden/film.npz holds what the compiled reference's film() made from
the film() arguments quillk-run derives from the texts in
tests/golden/film_case.py: the initial particles of each film and, for the
moving-window text, the column film() fills at i = nx - 3 after nmw shifts
(append_for_moving_window).  quillk-run's initial particles (--dump-setup)
must equal them (as sets: the reference keeps per-cell lists), positions
and momenta bit for bit and charges to a few ulp, and its column fills
(--mw-fills) must give the same cells and densities.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import film_case  # noqa: E402
import golden_io as gio  # noqa: E402

ROOT = os.path.dirname(HERE)
BIN = os.path.join(ROOT, "quill_amd", "bin", "quillk-run")
SLOT = {-1.0: 0, 1.0: 1}


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(BIN):
        pytest.skip("quillk-run not built (make -C quill_amd/driver)")


def driver(tmp_path, name, *args):
    conf = tmp_path / f"{name}.conf"
    conf.write_text(film_case.CONFS[name])
    r = subprocess.run([BIN, str(conf), *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def rows(a):
    return a[np.lexsort(a.T[::-1])]


@pytest.mark.parametrize("name", sorted(film_case.CONFS))
def test_initial_films_match_reference(tmp_path, name):
    fx = gio.load("film")
    j = json.loads(driver(tmp_path, name, "--dry-run"))
    setup = tmp_path / "setup"
    driver(tmp_path, name, "--dump-setup", str(setup), "--dry-run")
    ref = np.concatenate([fx[f"{name}_init_{k}"] for k in range(len(j["film_args"]))])
    np.testing.assert_array_equal(np.array(j["film_args"]), np.array([fx[f"{name}_init_{k}_args"]
                                                                        for k in range(len(j["film_args"]))]))
    slots = {**SLOT, **{c: 3 + n for n, c in enumerate(j["ion_cmr"])}}
    seen = 0
    for cmr in np.unique(ref[:, 8]):
        want = ref[ref[:, 8] == cmr][:, [0, 1, 2, 3, 4, 5, 7]]   # x y z ux uy uz q
        got = np.fromfile(setup / f"species{slots[cmr]}.f64").reshape(-1, 7)
        # place() keeps is_inside cells only (src/plist.cpp:22-48); the set-up
        # dump holds all, qk_add_particles drops the others
        c = np.floor(got[:, :3]).astype(int)
        inside = np.all((c > 0) & (c < np.array([j["nx"], j["ny"], j["nz"]]) - 2), axis=1)
        got = got[inside]
        assert len(got) == len(want), (cmr, len(got), len(want))
        got, want = rows(got), rows(want)
        np.testing.assert_array_equal(got[:, :6], want[:, :6])
        # charges to a few ulp: the reference build (g++ -O3 on an FMA target,
        # contraction on by default) fuses i * dx - x0 of the x ramp
        np.testing.assert_allclose(got[:, 6], want[:, 6], rtol=2e-15, atol=0)
        seen += len(got)
    assert seen == len(ref) > 1000


@pytest.mark.parametrize("nmw", film_case.NMW)
def test_window_columns_match_reference(tmp_path, nmw):
    fx = gio.load("film")
    ref = fx[f"window_col_{nmw}_0"]
    j = json.loads(driver(tmp_path, "window", "--dry-run"))
    fills = json.loads(driver(tmp_path, "window", "--mw-fills", str(nmw)))
    assert {f["species"] for f in fills} == {0, 1}   # electrons and (ions = positrons) positrons
    for f in fills:
        cmr = -1.0 if f["species"] == 0 else 1.0
        r = ref[ref[:, 8] == cmr]
        npt = int(np.prod(f["npic"]))
        # qk_fill_cells skips the cells outside is_inside, as place() does
        c = np.array(f["cells"])
        inside = np.all((c > 0) & (c < np.array([j["nx"], j["ny"], j["nz"]]) - 2), axis=1)
        f["cells"] = c[inside].tolist()
        f["density"] = np.array(f["density"])[inside].tolist()
        assert len(r) == npt * len(f["cells"])
        # per cell: npt particles of charge +-density / npt
        cells = np.floor(r[:, :3]).astype(int)
        key = (cells[:, 0] * 1000 + cells[:, 1]) * 1000 + cells[:, 2]
        uk, first = np.unique(key, return_index=True)
        c = np.array(f["cells"])
        np.testing.assert_array_equal(np.sort((c[:, 0] * 1000 + c[:, 1]) * 1000 + c[:, 2]), uk)
        dens = dict(zip(key[first], np.abs(r[first, 7]) * npt))
        got = dict(zip((c[:, 0] * 1000 + c[:, 1]) * 1000 + c[:, 2], np.array(f["density"])))
        for k in uk:
            assert got[k] == pytest.approx(dens[k], rel=1e-14)
        assert np.all(c[:, 0] == 37)   # nx - 3
        assert np.all(r[:, 3] == f["ux0"])
    # the negative-gradwidth ramp is what this column samples
    assert len({round(d, 12) for d in fills[0]["density"]}) == 1 and fills[0]["density"][0] < 0.3
