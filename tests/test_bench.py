"""bench.py pieces that run without a GPU: the roofline's measured traffic
comes from the committed rocprofv3 summaries, which must parse, and the
refusal of more than one GPU."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_measured_traffic_reads_the_committed_profiles():
    for wl in ("uniform-ndfx-vay-256x128x128", "wakefield-fdtd-boris-512x256x256"):
        t = bench.measured_traffic(wl)
        assert t is not None, wl
        # at least the algorithmic bytes of one launch, at most a few times them
        nx, ny, nz = bench.WORKLOADS[wl]["grid"]
        n_p = (nx - 3) * (ny - 3) * (nz - 3) * bench.PPC
        algo = 64.0 * n_p + 36.0 * nx * ny * nz
        assert algo < t < 3 * algo, (wl, t, algo)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
