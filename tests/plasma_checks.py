"""Statistical checks of the synthetic uniform thermal plasma.

The loader (oracle/philox.h on the CPU, k_load_uniform on the GPU) has no
counterpart in the reference: it is the builder-defined input of every
benchmark workload.  What it must produce is fixed by its definition, not by
its code: exactly ppc particles in every is_inside cell
(src/plist.cpp:42-48) and none elsewhere, in-cell offsets uniform on [0, 1)
on a 2^-24 lattice, momenta N(0, sigma^2) per component and independent,
charge q = -n / ppc (src/init.cpp:413-414).  Each moment is checked to four
standard errors, so a wrong sigma, a biased placement or a missing cell fails.
"""
import numpy as np


def _within(value, expect, se, what):
    assert abs(value - expect) <= 4 * se, f"{what}: {value:.6g} vs {expect:.6g} (4 SE = {4 * se:.3g})"


def check_uniform_plasma(pos, u, q, grid, ppc, density, sigma, q_rtol=0.0):
    """pos (n, 3) absolute positions in cells, u (n, 3), q (n,)."""
    nx, ny, nz = grid
    n = len(pos)
    cell = np.floor(pos).astype(np.int64)
    off = pos - cell
    inside = ((cell[:, 0] > 0) & (cell[:, 0] < nx - 2) & (cell[:, 1] > 0) & (cell[:, 1] < ny - 2) &
              (cell[:, 2] > 0) & (cell[:, 2] < nz - 2))
    assert inside.all(), f"{(~inside).sum()} particles outside the is_inside cells"
    lin = (cell[:, 0] * ny + cell[:, 1]) * nz + cell[:, 2]
    counts = np.bincount(lin, minlength=nx * ny * nz).reshape(nx, ny, nz)
    assert (counts[1:nx - 2, 1:ny - 2, 1:nz - 2] == ppc).all(), "an is_inside cell without exactly ppc particles"
    assert counts.sum() == n == (nx - 3) * (ny - 3) * (nz - 3) * ppc
    # offsets: the 24-bit lattice of [0, 1 - 2^-24], mean 1/2, variance 1/12
    assert (off >= 0).all() and (off < 1).all()
    scaled = off * 16777216.0
    assert np.array_equal(scaled, np.round(scaled)), "offsets off the 2^-24 lattice"
    for a in range(3):
        _within(off[:, a].mean(), 0.5, np.sqrt(1 / 12 / n), f"offset mean, axis {a}")
        _within(off[:, a].var(), 1 / 12, np.sqrt((1 / 80 - 1 / 144) / n), f"offset variance, axis {a}")
    # momenta: mean 0, variance sigma^2, fourth moment 3 sigma^4 (normal), uncorrelated
    for a in range(3):
        x = u[:, a] / sigma
        _within(x.mean(), 0.0, np.sqrt(1 / n), f"u mean, axis {a}")
        _within((x * x).mean(), 1.0, np.sqrt(2 / n), f"u variance / sigma^2, axis {a}")
        _within((x ** 4).mean(), 3.0, np.sqrt(96 / n), f"u fourth moment / sigma^4, axis {a}")
    for a, b in ((0, 1), (0, 2), (1, 2)):
        _within(np.mean(u[:, a] * u[:, b]) / sigma ** 2, 0.0, np.sqrt(1 / n), f"u correlation {a}{b}")
    # offsets independent of the momenta drawn from the same counter block
    _within(np.mean((off[:, 0] - 0.5) * u[:, 0] / sigma), 0.0, np.sqrt(1 / 12 / n), "offset-momentum correlation")
    qe = -density / ppc
    if q_rtol == 0:
        assert (q == qe).all()
    else:
        assert np.allclose(q, qe, rtol=q_rtol, atol=0)
