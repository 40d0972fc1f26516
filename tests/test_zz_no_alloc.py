"""qk_compute_rho_noalloc against qk_compute_rho through the C ABI.

The no-allocation variant accumulates the node charges of compute_rho
(src/fout.cpp:7-39) in the library's fixed device scratch instead of a
buffer allocated per call.  Two calls in a row, the second on a smaller grid
with fewer particles, must each give what qk_compute_rho gives for the same
context: the same FP64 deposit, so the only difference allowed is the order
of the FP64 atomic sums (relative 1e-12 of max |rho|).  The scratch keeps the
previous call's contents, so the second call also shows that it is cleared."""
import numpy as np
import pytest

import oracle
import quill_amd as qa

pytestmark = pytest.mark.gpu

DX = 0.05 * 2 * np.pi
DT = 0.025 * 2 * np.pi


def region(shape, ppc, seed, precision):
    nx, ny, nz = shape
    r = qa.Region(nx, ny, nz, DX, DX, DX, DT, solver="fdtd", pusher="boris", precision=precision,
                  deposit="tiled", device=0)
    rec, _ = oracle.synth_plasma(nx, ny, nz, ppc, 0.01, 0.05, seed)
    r.add_records(rec)
    # positrons too (charge and cmr of the opposite sign), so two of the
    # three slots are filled
    rec_p = rec[::3].copy()
    rec_p[:, 6] = -rec_p[:, 6]
    rec_p[:, 7] = 1.0
    r.add_records(rec_p)
    r.init_boundaries()
    r.step(3)
    return r


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_rho_matches_compute_rho(precision):
    for shape, ppc, seed in (((40, 28, 24), 4, 11), ((16, 12, 12), 2, 12)):
        print(f"noalloc rho: {precision} {shape} ppc {ppc}", flush=True)
        with region(shape, ppc, seed, precision) as r:
            want = r.compute_rho()
            got = r.compute_rho_noalloc()
        assert got.shape == want.shape
        scale = np.abs(want).max()
        assert scale > 0
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale)
        # the positron slot holds charge, the photon slot none
        assert np.abs(got[..., 1]).max() > 0 and not np.any(got[..., 2])


def test_noalloc_rho_refuses_large_grids():
    # 2**21 + 1 cells and more are refused before any kernel runs
    with qa.Region(130, 128, 128, DX, DX, DX, DT, precision="fp32", device=0) as r:
        with pytest.raises(qa.QuillError, match="above its limit"):
            r.compute_rho_noalloc()
