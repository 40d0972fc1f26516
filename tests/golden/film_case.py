"""The film cases shared by make_golden.py and the tests.

Two .conf texts whose films exercise film() (src/init.cpp:537-646): density
ramps along y (nfilm_y0 / nfilm_y1), the cumulative gradwidth_y scaling,
positive and negative gradwidth, ions and positrons, drift vxfilm, and, with
the moving window on, the column film() fills at i = nx - 3 for a shifted
film (append_for_moving_window).  make_golden.py feeds the film() arguments
quillk-run derives from these texts (its --dry-run "film_args") to the
reference's own film(); the tests compare quillk-run's particles (initial
set-up) and column fills (--mw-fills) with the result.
"""

BASE = """\
dx = 0.05
dy = 0.1
dz = 0.1
dt = 0.025
lambda = 1 um
ne = 0.02 ncr
xlength = 2.0001
ylength = 1.6001
zlength = 1.4001
t_end = 0.25
output_period = 0.25
a0 = 0
xsigma = 1
ysigma = 1
zsigma = 1
qed = off
solver = fdtd
pusher = boris
"""

CONFS = {
    "ramps": BASE + """\
mwindow = off
ions = on
film = on
x0film = 0.4
filmwidth = 0.9
gradwidth = 0.3
nfilm_y0 = 0.2 ncr
nfilm_y1 = 0.5 ncr
gradwidth_y = 0.25
y0film = 0.2
y1film = 1.4
z0film = 0.1
z1film = 1.2
mcr = 3
vxfilm = 0.2
xnpic_film = 2
ynpic_film = 1
znpic_film = 2
""",
    "window": BASE + """\
mwindow = on
mwseed = off
ions = positrons
film = on
x0film = 0.7
filmwidth = 3
gradwidth = -2
nfilm = 0.3 ncr
xnpic = 1
ynpic = 2
znpic = 1
""",
}

# shifts after which the "window" film's column is compared
NMW = [2, 5, 9]
