"""CPU checks of the C ABI: the library loads, exports every declared entry
point, its structs have the C layout, and it fails cleanly without a GPU."""
import os
import re
import subprocess
import tempfile

import ctypes as C
import pytest

import quill_amd as qa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "quillk.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(qk_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = qa.load_library()
    names = declared()
    assert len(names) >= 16
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(qa.EXPORTS)


def test_struct_layout_matches_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "quillk.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(qk_params), offsetof(qk_params, dx),
         offsetof(qk_params, e_s), offsetof(qk_params, seed), offsetof(qk_params, ion_cmr),
         offsetof(qk_params, norm_ep), offsetof(qk_params, sort_fallback), sizeof(qk_diag));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "l.c")
        exe = os.path.join(td, "l")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    P = qa.Params
    want = [C.sizeof(P), P.dx.offset, P.e_s.offset, P.seed.offset, P.ion_cmr.offset, P.norm_ep.offset,
            P.sort_fallback.offset, C.sizeof(qa.Diag)]
    assert got == want


def test_create_rejects_bad_config_without_crashing():
    # invalid grids are rejected before any device call
    with pytest.raises(qa.QuillError, match="invalid"):
        qa.Region(3, 8, 8, 0.3, 0.3, 0.3, 0.1)
    with pytest.raises(qa.QuillError, match="dt must be below"):
        qa.Region(8, 8, 8, 0.3, 0.3, 0.3, 0.5)
    with pytest.raises(qa.QuillError):
        qa.Region(8, 8, 8, 0.3, 0.3, 0.3, 0.1, solver=7)
    with pytest.raises(qa.QuillError, match="sort_fallback"):
        qa.Region(8, 8, 8, 0.3, 0.3, 0.3, 0.1, sort_fallback=1.5)
    # cell and tile-key indices are 32-bit: 2048 x 1024 x 1024 cells is 2^31
    with pytest.raises(qa.QuillError, match="2\\^31"):
        qa.Region(2048, 1024, 1024, 0.3, 0.3, 0.3, 0.1)
    with pytest.raises(qa.QuillError, match="2\\^31"):
        qa.Region(2041, 1017, 1017, 0.3, 0.3, 0.3, 0.1)


def test_null_arguments_return_errors():
    L = qa.load_library()
    assert L.qk_step(None, 1, None) == 1
    assert L.qk_sync(None) == 1
    assert L.qk_destroy(None) == 0
    assert b"null" in L.qk_last_error()
