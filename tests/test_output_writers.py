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
