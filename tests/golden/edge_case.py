"""Hand-placed particles for the deposit's edge cases (run_edges.npz).

A 16 x 12 x 10 grid (tiles of 8 cells end part-way along y and z), zero
initial fields, FDTD + Boris, dt = dx / 2, three steps.  Positions are in
cell units (cell index + offset), as the reference stores them.  Each
particle exercises one branch of the charge-conserving split of
src/deposition.cpp:41-356 (jdeposition) or of the boundary handling:

  ties        two or three axes crossing at the same path fraction
  faces       starting exactly on a cell face, moving away from it or across it
  near        within 1e-12 of a face, crossing it or stopping short
  exits       leaving is_inside through a low or a high face (deleted after
              depositing, src/plist.cpp:7-40), and one placed outside it
              (never added, place(), src/plist.cpp:22-31)
  fast        |u| ~ 50 along a diagonal: three crossings at distinct fractions
  pile        eight particles in one cell (colliding LDS atomics)
  species     positrons, an ion and a photon among them

plus a background of 300 electrons at random positions with |u| ~ 1, so
that one- and two-face crossings occur in every direction.
"""
import numpy as np

SHAPE = (16, 12, 10)
PI = np.pi
DX = 0.05 * 2 * PI
DT = 0.5 * DX
NSTEPS = 3
PROTON_MASS = 1836.1526721   # include/physical_constants.h:4


def _rec(pos, u, q=-1e-3, cmr=-1.0):
    return [pos[0], pos[1], pos[2], u[0], u[1], u[2], q, cmr]


def particles():
    """(rec[n, 8] = x y z ux uy uz q cmr, ids[n], ion cmr)."""
    d = lambda u: 0.5 * np.asarray(u, float) / np.sqrt(1 + np.dot(u, u))   # displacement in cells
    r = []
    # ties
    r.append(_rec((5.8, 5.8, 4.5), (1.0, 1.0, 0.0)))           # x and y at one fraction
    r.append(_rec((6.85, 6.85, 5.85), (1.0, 1.0, 1.0)))        # x, y and z at one fraction
    r.append(_rec((7.15, 5.15, 5.15), (-1.0, -1.0, -1.0)))     # the same, moving down
    r.append(_rec((4.9, 3.1, 3.9), (1.0, -1.0, 1.0)))          # ... with mixed signs
    # three crossings at distinct fractions
    r.append(_rec((8.95, 4.2, 6.75), (2.0, -2.0, 2.0)))
    r.append(_rec((3.05, 8.9, 2.2), (-3.0, 3.0, -3.0)))
    # on a face
    r.append(_rec((9.0, 6.5, 5.5), (-0.5, 0.0, 0.0)))          # offset 0, moving down: crosses at t = 0
    r.append(_rec((9.0, 7.0, 3.0), (0.3, 0.3, 0.3)))           # offset 0 on three axes, moving up
    r.append(_rec((6.0, 3.0, 7.0), (-0.4, -0.4, -0.4)))        # offset 0 on three axes, moving down
    # within 1e-12 of a face
    r.append(_rec((10.0 - 1e-12, 6.5, 5.5), (0.2, 0.0, 0.0)))  # crossing
    r.append(_rec((10.0 - 1e-12, 6.5, 4.5), (0.0, 0.1, 0.0)))  # stopping short along x
    r.append(_rec((5.5, 4.0 + 1e-12, 4.5), (0.0, -0.2, 0.0)))  # crossing down
    u = np.array([0.7, 0.0, 0.0])
    r.append(_rec((7.0 - d(u)[0], 2.5, 2.5), tuple(u)))        # ending on a face (to rounding)
    # leaving is_inside (1 <= i < nx - 2 on each axis) and starting outside it
    r.append(_rec((1.2, 5.5, 5.5), (-2.0, 0.0, 0.0)))          # to x cell 0
    r.append(_rec((5.5, 9.8, 5.5), (0.0, 2.0, 0.0)))           # to y cell ny - 2
    r.append(_rec((5.5, 5.5, 7.9), (0.0, 0.0, 3.0)))           # to z cell nz - 2
    r.append(_rec((13.9, 1.1, 1.1), (2.0, -2.0, -2.0)))        # out through a corner
    r.append(_rec((0.5, 5.5, 5.5), (1.0, 0.0, 0.0)))           # starting in cell 0
    # fast, three crossings
    r.append(_rec((9.8, 4.3, 4.9), (30.0, -40.0, 20.0)))
    r.append(_rec((3.3, 7.85, 6.8), (-45.0, 25.0, 30.0)))
    # eight in one cell
    rng = np.random.default_rng(7)
    for m in range(8):
        off = rng.uniform(0.05, 0.95, 3)
        r.append(_rec(4.0 + off, tuple(rng.normal(0, 0.8, 3))))
    # species
    icmr = 1.0 / PROTON_MASS
    r.append(_rec((8.3, 8.6, 2.9), (0.5, -0.5, 0.5), q=1e-3, cmr=1.0))       # positron, one crossing
    r.append(_rec((2.95, 2.95, 2.95), (1.0, 1.0, 1.0), q=1e-3, cmr=1.0))     # positron, triple tie
    r.append(_rec((11.6, 3.3, 7.5), (40.0, 10.0, 0.0), q=2e-3, cmr=icmr))    # ion
    r.append(_rec((6.5, 6.5, 6.99), (0.0, 0.0, 1.0), q=1e-3, cmr=0.0))        # photon across z
    rec = np.array(r, float)
    # background: random positions in is_inside, |u| ~ 1
    nb = 300
    bg = np.zeros((nb, 8))
    for a, n in enumerate(SHAPE):
        bg[:, a] = rng.uniform(1.0, n - 2.0, nb)
    bg[:, 3:6] = rng.normal(0, 1.0, (nb, 3))
    bg[:, 6] = -1e-3
    bg[:, 7] = -1.0
    rec = np.concatenate([rec, bg])
    ids = np.arange(len(rec), dtype=np.int32)
    return rec, ids, icmr
