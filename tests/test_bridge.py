"""The reference-side binding (quill_amd/bridge/quill_gpu.cpp) builds against
the reference's own headers and objects, the way a QUILL build would link
it: oracle/Makefile compiles it with oracle/bridge_check.cpp (main()'s
globals, a region set up by the reference's own calls) and the reference's
hot-path units into oracle/_ref/bridge_check.  Without a GPU the binding's
qk_create fails and the driver reports it (exit 3); tests/test_gpu_bridge.py
runs it on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
BIN = os.path.join(ROOT, "oracle", "_ref", "bridge_check")


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "include", "main.h")), reason="reference tree absent")
def test_binding_compiles_links_and_fails_cleanly_without_gpu():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", os.path.abspath(BIN)], check=True)
    assert os.path.exists(BIN)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")   # no device even where one exists
    r = subprocess.run([BIN, "0", "1", "2"], capture_output=True, text=True, env=env)
    assert r.returncode == 3, (r.returncode, r.stderr)
    assert "no GPU context" in r.stderr
