"""The reference-side binding on the GPU: oracle/_ref/bridge_check sets a
region up with the reference's own calls, hands it to libquillk.so through
quill_gpu_create (main()'s globals, a walk of the cp[i][j][k] lists), and
runs the loop body of src/main.cpp:2319-2332 on both the reference (CPU) and
the GPU (FP64, through quill_gpu_step); fields, current, Be and the per-step
diagnostics must agree to 1e-10."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "bridge_check")


@pytest.mark.skipif(not os.path.exists(BIN), reason="bridge_check not built (needs the reference tree)")
@pytest.mark.parametrize("solver,pusher", [(0, 1), (1, 0), (2, 0)])
def test_binding_runs_the_reference_region_on_the_gpu(solver, pusher):
    print(f"bridge_check solver {solver} pusher {pusher}", flush=True)
    r = subprocess.run([BIN, str(solver), str(pusher), "12"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["electrons"] > 1000
    for k in ("E", "B", "J", "Be", "diagnostics"):
        assert res[k] < 1e-10, (k, res[k])
