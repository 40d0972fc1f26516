"""quillk-run, the .conf-driven driver: config grammar and the derived run.

The grammar and storage follow parse.py:8-66 and var::read
(src/low-level.cpp:376-444); the derived quantities follow init()
(src/main.cpp:2558-3600).  Known answers are worked out by hand from those
rules (values pass through float as stof does).  No GPU is touched
(--print-conf / --dry-run).
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "quill_amd", "bin", "quillk-run")
TWO_PI = 2 * math.pi


def f32(x):
    return float(np.float32(x))


def run(tmp_path, text, *args, ok=True):
    conf = tmp_path / "run.conf"
    conf.write_text(text)
    r = subprocess.run([BIN, str(conf), *args], capture_output=True, text=True, timeout=60)
    if ok:
        assert r.returncode == 0, r.stderr
    return r


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(BIN):
        pytest.skip("quillk-run not built (make -C quill_amd/driver)")


def test_grammar(tmp_path):
    text = ("# comment line\n"
            "a = 1.5   # trailing comment\n"
            "b = [1;2.5;-3e2]\n"
            "c = on\n"
            "d = 0.91 um\n"
            "e = 4e20 W/cm^2\n"
            "f = 1e\n"
            "g = 5 cm^{-3}\n"
            "h = 12 %\n"
            "i = .5\n"
            "j = xy&xz\n")
    v = json.loads(run(tmp_path, text, "--print-conf").stdout)
    got = {x["name"]: (x["value"], x["units"], x["array"]) for x in v}
    assert got["a"] == (1.5, "#", [])                     # plain number: units "#"
    assert got["b"] == (0.0, "#", [1.0, 2.5, -300.0])     # arrays in double
    assert got["c"] == (0.0, "on", [])                    # text: value 0, text in units
    assert got["d"] == (f32(0.91), "um", [])              # number through float (stof)
    assert got["e"] == (f32(4e20), "W/cm^2", [])
    assert got["f"] == (0.0, "1e", [])                    # not a number: text
    assert got["g"] == (5.0, "cm^{-3}", [])
    assert got["h"] == (12.0, "%", [])
    assert got["i"] == (0.5, "#", [])
    assert got["j"] == (0.0, "xy&xz", [])
    assert [x["name"] for x in v] == list("abcdefghij")   # order kept (find takes the first)


@pytest.mark.parametrize("line, msg", [("a = eval{1+1}", "eval"), ("a = 1 = 2", "key = value"),
                                       ("a = 1 um m", "incorrect"), ("a b = 1", "bad key")])
def test_grammar_errors(tmp_path, line, msg):
    r = run(tmp_path, line + "\n", "--print-conf", ok=False)
    assert r.returncode == 1 and msg in r.stderr


BASE = """\
dx = 0.05
dy = 0.1
dz = 0.1
dt = 0.025
lambda = 1 um
ne = 0.01 ncr
xlength = 3.2001
ylength = 2.4001
zlength = 2.4001
t_end = 0.5
output_period = 0.25
f_envelope = cos
polarization = linear
a0 = 1
xsigma = 1
ysigma = 1
zsigma = 1
x0 = -0.5
mwindow = off
qed = off
solver = fdtd
pusher = boris
film = on
x0film = 1
filmwidth = 1
gradwidth = 0
nfilm = 0.01 ncr
xnpic = 2
ynpic = 1
znpic = 1
"""


def test_derived_run(tmp_path):
    j = json.loads(run(tmp_path, BASE, "--dry-run").stdout)
    dx, dy, dt = f32(0.05) * TWO_PI, f32(0.1) * TWO_PI, f32(0.025) * TWO_PI
    assert j["dx"] == pytest.approx(dx, rel=1e-15) and j["dy"] == pytest.approx(dy, rel=1e-15)
    assert j["dt"] == pytest.approx(dt, rel=1e-15)
    assert (j["nx"], j["ny"], j["nz"]) == (64, 24, 24)
    # OutputIterations: last iteration round(t_end / dt) = round(0.5 / 0.025) = 20 -> 21 steps
    assert j["steps"] == 21
    assert (j["solver"], j["pusher"], j["qed"]) == ("fdtd", "boris", 0)
    lam = f32(1) * 1e-4
    assert j["lambda_cm"] == pytest.approx(lam, rel=1e-15)
    assert j["e_s"] == pytest.approx(lam / 2.4263086e-10, rel=1e-15)
    vol = dx * dy * dy
    assert j["norm_n"] == pytest.approx(vol * 1.11485e13 * lam / (8 * math.pi ** 3), rel=1e-14)
    assert j["norm_ef"] == pytest.approx(0.5 * vol * 3.691e4 * lam / 1e7, rel=1e-14)
    # one laser call: f_init_cos(a0, 0, xs, ys, zs, xlength/2 + x0, b_sign=1, x0, phase 0, 0, 0, append, phi 0)
    assert len(j["lasers"]) == 1
    L = j["lasers"][0]
    xl = f32(3.2001) * TWO_PI
    x0 = f32(-0.5) * TWO_PI
    np.testing.assert_allclose(L[:8], [1.0, 0.0, TWO_PI, TWO_PI, TWO_PI, xl / 2 + x0, 1.0, x0], rtol=1e-6)
    # film (src/init.cpp:537-646): cells i in [int(x0/dx), int(x1/dx)), j in
    # [int(y0/dy) + 1, int(y1/dy) - 1) with y0 = 0, y1 = ylength, same for k;
    # 2 particles each, q = -n / 2 per electron (place() later drops the
    # cells outside is_inside; the set-up counts all)
    i0 = int(f32(1) * TWO_PI / dx)
    i1 = int((f32(1) * TWO_PI + f32(1) * TWO_PI) / dx)
    nj = int(f32(2.4001) * TWO_PI / dy) - 2
    cells = (i1 - i0) * nj * nj
    assert j["particles"]["0"]["n"] == 2 * cells
    assert j["particles"]["0"]["q_sum"] == pytest.approx(-f32(0.01) * cells, rel=1e-12)


def test_reflections_and_circular(tmp_path):
    text = BASE.replace("polarization = linear", "polarization = circular") + "lp_reflection = xy&yz\nf_reflection = y\n"
    j = json.loads(run(tmp_path, text, "--dry-run").stdout)
    # base pulse + xy mirror + two yz mirrors (src/main.cpp:1183-1195)
    assert len(j["lasers"]) == 4
    a = np.array(j["lasers"])
    assert a[0, 0] == 1.0 and a[0, 1] == 1.0          # circular: a0z = a0y
    assert a[1, 0] == -1.0                            # f_reflection = y flips a0y of the first mirror
    assert a[2, 7] == pytest.approx(-a[0, 7]) and a[3, 7] == pytest.approx(-a[0, 7])   # yz mirrors at -x0
    assert a[1, 10] == pytest.approx(-a[0, 10])       # xy mirror flips z0


def test_merging_options(tmp_path):
    j = json.loads(run(tmp_path, "pmerging = nl\ncrpc = 2.5\n" + BASE, "--dry-run").stdout)
    # crnp = crpc * xlength ylength zlength / (dx dy dz), src/main.cpp:1701
    dx, dy = f32(0.05) * TWO_PI, f32(0.1) * TWO_PI
    cells = (f32(3.2001) * TWO_PI) * (f32(2.4001) * TWO_PI) ** 2 / (dx * dy * dy)
    assert j["merging"] == 2 and j["crnp"] == pytest.approx(f32(2.5) * cells, rel=1e-12)
    # output iterations t = 0, 0.25, 0.5
    assert j["outputs"] == 3


def test_moving_window_options(tmp_path):
    # mwindow is on unless "off" (src/main.cpp:2932-2947), mwspeed 1, t_add_mw = t_end
    j = json.loads(run(tmp_path, BASE.replace("mwindow = off\n", ""), "--dry-run").stdout)
    assert (j["mwindow"], j["mwspeed"], j["mwtolerance"]) == (1, 1.0, 1e-6)
    assert j["t_add_mw"] == pytest.approx(0.5 * TWO_PI, rel=1e-15)
    # seed column (src/main.cpp:1756-1786): every (j, k) below int(length / d),
    # density 1 / k0^2 = ne / n_cr; the film (x0film = 1, width 1) does not
    # reach the column nx - 3 = 61 after one shift
    assert j["mw_fills"] == [[0, 24 * 24, pytest.approx(24 * 24 * f32(0.01), rel=1e-12)]]
    # "auto", mwseed off, a film reaching the front: its column only,
    # cells j, k in [1, 23) (film(), src/init.cpp:548-560, 588-592)
    text = ("mwindow = auto\nmwtolerance = 0.01\nmwseed = off\nt_add_mw = 3 um\n" +
            BASE.replace("mwindow = off\n", "").replace("filmwidth = 1", "filmwidth = 5"))
    j = json.loads(run(tmp_path, text, "--dry-run").stdout)
    assert (j["mwindow"], j["mwtolerance"]) == (2, f32(0.01))
    assert j["t_add_mw"] == pytest.approx(f32(3) * 1e-4 / 1e-4 * TWO_PI, rel=1e-12)
    assert j["mw_fills"] == [[0, 22 * 22, pytest.approx(22 * 22 * f32(0.01), rel=1e-6)]]
    # density profiles (lin_interpolation, src/low-level.cpp:492-510): x at
    # (nmw - 1) dx = dx, r from the box centre
    text = ("ne_profile_x_coords = [0;10]\nne_profile_x_values = [1;3]\n"
            "ne_profile_r_coords = [0;0.5]\nne_profile_r_values = [1;0]\n" + BASE.replace("mwindow = off\n", ""))
    j = json.loads(run(tmp_path, text, "--dry-run").stdout)
    dx, dy = f32(0.05) * TWO_PI, f32(0.1) * TWO_PI
    nx_ = 1 + 2 * dx / (10 * TWO_PI)
    L = f32(2.4001) * TWO_PI
    jj, kk = np.meshgrid(np.arange(24), np.arange(24), indexing="ij")
    r = np.hypot(jj * dy - L / 2, kk * dy - L / 2)
    mod = np.where(r > 0.5 * TWO_PI, 0.0, 1 - r / (0.5 * TWO_PI))
    assert j["mw_fills"][0][:2] == [0, int((mod != 0).sum())]
    assert j["mw_fills"][0][2] == pytest.approx(f32(0.01) * nx_ * mod.sum(), rel=1e-9)


@pytest.mark.parametrize("extra, msg", [("beam = on\n", "beam"),
                                        ("pmerging = xx\n", "merging"), ("f_envelope = focused\n", "focused"),
                                        ("solver = spectral\n", "solver unknown")])
def test_unsupported_options_are_refused(tmp_path, extra, msg):
    # find() takes the first occurrence, so the override goes first
    r = run(tmp_path, extra + BASE, "--dry-run", ok=False)
    assert r.returncode == 1 and msg in r.stderr


def test_qed_defaults_on(tmp_path):
    j = json.loads(run(tmp_path, BASE.replace("qed = off\n", ""), "--dry-run").stdout)
    assert j["qed"] == 1   # find_boolean(first, "qed", true), src/main.cpp:3542


LOG_KEYS = ["dx", "dy", "dz", "dt", "nx", "ny", "nz", "lambda", "ne", "t_end", "output_period", "t_add_mw",
            "xsigma", "ysigma", "zsigma", "x0fout", "a0y", "a0z", "external_bz", "mwindow",
            "e_components_for_output", "b_components_for_output", "f_envelope", "b_sign", "x0", "y0", "z0",
            "xtarget", "ytarget", "ztarget", "phase", "lp_reflection", "f_reflection", "phi", "phi_rotate",
            "shenergy", "shphase", "beam", "beam_particles", "Nb", "epsb", "xb", "rb", "x0b", "y0b", "phib",
            "mwspeed", "nelflow", "vlflow", "mcrlflow", "Tlflow", "nerflow", "vrflow", "mcrrflow", "Trflow", "ions",
            "film", "x0film", "filmwidth", "gradwidth", "gradwidth_y", "y0film", "y1film", "z0film", "z1film",
            "nfilm", "mcr", "Tfilm", "vxfilm", "xnpic_film", "ynpic_film", "znpic_film", "n_ion_populations",
            "icmr", "xnpic", "ynpic", "znpic", "particles_for_output", "deps", "neps", "enthp", "deps_p", "neps_p",
            "enthp_p", "deps_ph", "neps_ph", "enthp_ph", "deps_i", "neps_i", "enthp_i", "n_sr", "n_tracks",
            "particles_to_track", "tr_start", "tr_init", "xtr1", "ytr1", "ztr1", "xtr2", "ytr2", "ztr2", "pmerging",
            "crpc"]


def test_log_file(tmp_path):
    """The log main() writes after init() (src/main.cpp:3611-3787): key and
    value lines in the reference's order up to "$" (the order of its
    writer: one ion population's icmr, one film), then the run switches;
    values in lambda units and cm^-3, as python/resread.py:52-133 reads
    them back."""
    text = ("dx = 0.05\ndy = 0.04\ndz = 0.05\ndt = 0.025\nlambda = 0.8 um\nne = 0.02 ncr\n"
            "xlength = 3.2\nylength = 1.6\nzlength = 1.2\nt_end = 5\noutput_period = 2.5\n"
            "a0 = 1.5\nxsigma = 2\nysigma = 3 um\nzsigma = 2\nx0 = -1\nmwindow = off\n"
            "film = on\nx0film = 0.2\nfilmwidth = 2.5\nnfilm = 0.01 ncr\nions = on\nmcr = 2\n"
            "xnpic = 2\nynpic = 1\nznpic = 1\ndeps = 1 mc^2\nneps = 100\nenthp = 10\n"
            "particles_for_output = epph\noutput_mode = binary\nsolver = fdtd\npusher = boris\n")
    run(tmp_path, text, "--write-log", str(tmp_path / "out"))
    lines = (tmp_path / "out" / "log").read_text().split("\n")
    end = lines.index("$")
    pairs = list(zip(lines[:end:2], lines[1:end:2]))
    assert [k for k, _ in pairs] == LOG_KEYS
    v = dict(pairs)
    lam_cm = f32(0.8) * 1e-4
    assert float(v["dx"]) == pytest.approx(f32(0.05), rel=1e-5)
    assert float(v["dy"]) == pytest.approx(f32(0.04), rel=1e-5)
    assert float(v["dt"]) == pytest.approx(f32(0.025), rel=1e-5)
    assert (int(v["nx"]), int(v["ny"]), int(v["nz"])) == (64, 40, 24)
    assert float(v["lambda"]) == pytest.approx(lam_cm, rel=1e-5)
    assert float(v["ne"]) == pytest.approx(f32(0.02) * 1.11485e13 / lam_cm ** 2, rel=1e-5)
    assert float(v["t_end"]) == pytest.approx(5, rel=1e-5) and float(v["output_period"]) == pytest.approx(2.5)
    assert float(v["ysigma"]) == pytest.approx(f32(3) * 1e-4 / lam_cm, rel=1e-5)   # um -> lambda
    assert float(v["a0y"]) == pytest.approx(1.5) and float(v["a0z"]) == 0
    assert float(v["deps"]) == pytest.approx(0.511, rel=1e-5)                       # mc^2 -> MeV
    assert v["particles_for_output"] == "epph" and v["n_ion_populations"] == "1"
    assert float(v["icmr"]) == pytest.approx(1 / (1836.1526721 * 2), rel=1e-5)
    assert float(v["nfilm"]) == pytest.approx(f32(0.01) * 1.11485e13 / lam_cm ** 2, rel=1e-5)
    assert v["mwindow"] == "off" and v["n_sr"] == "1"
    tail = lines[end + 1:]
    for k, want in (("catching", "off"), ("dump_photons", "off"), ("output_mode", "1")):
        assert tail[tail.index(k) + 1] == want
    assert tail[tail.index("qed") + 1] in ("on", "off")
