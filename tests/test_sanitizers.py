"""Address- and undefined-behaviour-sanitizer builds of the host code.

* the FP64 oracle (oracle/quill_oracle.c) under oracle/sanitize_check.c,
  which drives every entry point on small cases: all three solvers, QED,
  both merging modes, the moving window, column fills, laser sources on
  both walls, rho, max_w;
* the driver quillk-run (quill_amd/driver), on a config with a film and a
  moving window: --print-conf, --dry-run, --write-log, --mw-fills, and the
  plane and particle writers on given data.
Host code only (g++ / gcc); nothing here touches a GPU.  Any sanitizer
report fails the test.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
REPORTS = ("runtime error", "AddressSanitizer", "LeakSanitizer")

CONF = """dx = 0.05
dy = 0.05
dz = 0.05
dt = 0.025
lambda = 0.8 um
ne = 0.01 ncr
xlength = 2
ylength = 1
zlength = 1
t_end = 2
output_period = 1
a0 = 1
xsigma = 1
ysigma = 0.5
zsigma = 0.5
x0 = -0.5
mwindow = on
mwspeed = 1
film = on
x0film = 0.2
filmwidth = 1.5
nfilm = 0.01 ncr
ions = on
xnpic = 1
ynpic = 1
znpic = 1
deps = 1
neps = 50
enthp = 3
particles_for_output = epph
solver = ndfx
pusher = vay
"""


def clean(r):
    return r.returncode == 0 and not any(k in r.stderr for k in REPORTS)


def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_san"
    subprocess.run(["gcc", *SAN, "-std=c11", "-I", os.path.join(ROOT, "oracle"), "-o", str(exe),
                    os.path.join(ROOT, "oracle", "sanitize_check.c"), os.path.join(ROOT, "oracle", "quill_oracle.c"),
                    "-lm"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert clean(r), (r.returncode, r.stderr[-3000:])
    assert r.stdout.count("particles") == 4


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "quill_amd", "libquillk.so")), reason="library not built")
def test_driver_under_sanitizers(tmp_path):
    d = os.path.join(ROOT, "quill_amd", "driver")
    exe = tmp_path / "quillk_run_san"
    subprocess.run(["g++", *SAN, "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    *[os.path.join(d, f) for f in ("qk_conf.cpp", "qk_setup.cpp", "qk_output.cpp", "quillk_run.cpp")],
                    "-L", os.path.join(ROOT, "quill_amd"), "-lquillk", "-Wl,-rpath," + os.path.join(ROOT, "quill_amd"),
                    "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    conf = tmp_path / "run.conf"
    conf.write_text(CONF)
    for args in (["--print-conf"], ["--dry-run"], ["--write-log", str(tmp_path / "log")], ["--mw-fills", "3"]):
        r = subprocess.run([str(exe), str(conf), *args], capture_output=True, text=True, timeout=120)
        assert clean(r), (args, r.returncode, r.stderr[-3000:])
    # the plane writers on random arrays (text mode)
    nx, ny, nz = 10, 8, 6
    rng = np.random.default_rng(1)
    planes = tmp_path / "planes.bin"
    with open(planes, "wb") as f:
        f.write(np.array([nx, ny, nz, 4, 0], "<i4").tobytes())
        for _ in range(4):
            f.write(rng.normal(size=(nx, ny, nz, 3)).astype("<f8").tobytes())
    r = subprocess.run([str(exe), "--write-planes", str(planes), str(tmp_path / "p")], capture_output=True,
                       text=True, timeout=120)
    assert clean(r), r.stderr[-3000:]
    # the particle writers on a mixed population
    parts = tmp_path / "parts.bin"
    with open(parts, "wb") as f:
        f.write(np.int32(1).tobytes())
        f.write(np.array([0.5, 0.5, 0.5, 0.5], "<f8").tobytes())
        f.write(np.array([30, 30, 30, 30, 2, 2, 2, 2, 1, 1], "<i4").tobytes())
        f.write(np.array([0.3, 0.3, 0.3, 1e-4, 1.0 / 1836.0], "<f8").tobytes())
        for n, g in ((200, 5.0), (50, 3.0), (40, 8.0), (30, 1.001)):
            rec = np.zeros((n, 10))
            rec[:, :3] = rng.uniform(1, 7, (n, 3))
            rec[:, 6] = rng.uniform(1, g, n)
            rec[:, 7] = 0.01
            f.write(np.int64(n).tobytes())
            f.write(rec.astype("<f8").tobytes())
    r = subprocess.run([str(exe), "--write-particles", str(parts), str(tmp_path / "s")], capture_output=True,
                       text=True, timeout=120)
    assert clean(r), r.stderr[-3000:]
    assert (tmp_path / "s" / "spectrum").exists()
