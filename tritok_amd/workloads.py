"""Cube workloads in the reference's own input formats.

The texts produced here are .bic files (lib_bic.f:50-502 keyword format) and
tri.inp files (lib_IO.f:2-359) for the unit cube with the GiD wall numbering of
the reference meshes (1:z=0, 2:y=0, 3:x=1, 4:y=1, 5:x=0, 6:z=1):

  * lid-driven cavity: u_x = lid on wall 6, no-slip elsewhere (the reference's
    mesh/3Dldc-X.bic);
  * differentially heated cavity: no-slip walls, T = t_x0 on wall 5 (x=0) and
    t_x1 on wall 3 (x=1), adiabatic elsewhere (mesh/NaturalConvection-X.bic has
    +0.5 at x=1; the benchmark configuration uses +0.5 at x=0).
"""
from .host import Host

# edge / corner ownership (wall1 wall2 -> wall), as in the reference's cube .bic files
_EDGES_CAVITY = [(1, 2, 1), (1, 3, 1), (1, 4, 1), (1, 5, 1), (2, 3, 3), (3, 4, 3), (4, 5, 5), (5, 2, 5),
                 (6, 2, 6), (6, 3, 6), (6, 4, 6), (6, 5, 6)]
_CORNERS_CAVITY = [(1, 2, 3, 1), (1, 3, 4, 1), (1, 4, 5, 1), (1, 5, 2, 1), (6, 2, 3, 6), (6, 3, 4, 6), (6, 4, 5, 6),
                   (6, 5, 2, 6)]
_EDGES_HEATED = [(1, 2, 1), (1, 3, 3), (1, 4, 1), (1, 5, 5), (2, 3, 3), (3, 4, 3), (4, 5, 5), (5, 2, 5),
                 (6, 2, 6), (6, 3, 3), (6, 4, 6), (6, 5, 5)]
_CORNERS_HEATED = [(1, 2, 3, 3), (1, 3, 4, 3), (1, 4, 5, 5), (1, 5, 2, 5), (6, 2, 3, 3), (6, 3, 4, 3), (6, 4, 5, 5),
                   (6, 5, 2, 5)]
# wall-normal component of velocity is the kode-1 (flux) unknown on each wall
_NORMAL_WALLS = {"X": (3, 5), "Y": (2, 4), "Z": (1, 6)}


def _kin_block(comp, values):
    lines = ["%s-KINEMATICS" % comp]
    for w in range(1, 7):
        kode = 1 if w in _NORMAL_WALLS[comp] else 0
        lines += ["%d %d 0" % (w, kode), repr(float(values.get(w, 0.0)))]
    return lines + ["END", ""]


def _vort_blocks():
    out = []
    for comp in "XYZ":
        out += ["%s-VORTICITY" % comp] + ["1 %d 0.0D0" % w for w in range(1, 7)] + ["END", ""]
    return out


def bic_cavity(lid=1.0):
    """3D lid-driven cavity, u_x = lid on z = 1."""
    lines = ["# lid-driven cavity", "EDGE"] + ["%d %d %d" % e for e in _EDGES_CAVITY] + ["END", "CORNER"]
    lines += ["%d %d %d %d" % c for c in _CORNERS_CAVITY] + ["END", "INITIAL", "4 0", "END"]
    lines += _kin_block("X", {6: lid}) + _kin_block("Y", {}) + _kin_block("Z", {}) + _vort_blocks()
    return "\n".join(lines) + "\n"


def bic_heated(t_x0=0.5, t_x1=-0.5):
    """Differentially heated cube: T fixed on x = 0 and x = 1, adiabatic elsewhere."""
    lines = ["# differentially heated cube", "EDGE"] + ["%d %d %d" % e for e in _EDGES_HEATED] + ["END", "CORNER"]
    lines += ["%d %d %d %d" % c for c in _CORNERS_HEATED] + ["END", "INITIAL", "10 0", "END"]
    lines += _kin_block("X", {}) + _kin_block("Y", {}) + _kin_block("Z", {}) + _vort_blocks()
    temp = {1: "2 1 0.0D0", 2: "2 2 0.0D0", 3: "1 3 %r" % float(t_x1), 4: "2 4 0.0D0", 5: "1 5 %r" % float(t_x0),
            6: "2 6 0.0D0"}
    lines += ["TEMP"] + [temp[w] for w in range(1, 7)] + ["END", ""]
    return "\n".join(lines) + "\n"


def inp_text(re=1.0, rapr=None, urdw=0.3, urdt=0.1, eps=1e-6, nnlit=10000, lsqs_eps=1e-7, lsqs_maxit=5000,
             dlse=(1, 10.0, 0.001), gauss=(4, 4, 4, 4), iidiv=2, ibw=1):
    """tri.inp with the keywords the hot path reads (ibw = EQNS iBw: 1 solves
    the single-domain kinematics for the boundary vorticity, 0 keeps it)."""
    eq_t = 1 if rapr is not None else 0
    lines = ["MDIR ./", "URBW 1.0 1.0 1.0", "URDV 1.0 1.0 1.0", "URDW %r %r %r" % (urdw, urdw, urdw),
             "URDT %r" % urdt, "URDC 1.0"]
    if rapr is not None:
        lines.append("RAPR %r %r" % (float(rapr[0]), float(rapr[1])))
    lines += ["REYN %r" % float(re), "PARA 0 0", "GRAV 0 0 -1", "EQNS %d 1 1 %d 0 0" % (ibw, eq_t),
              "RELA %d 1.0 %r" % (nnlit, eps), "TIME 1.0D20 1 1", "LSQS %d %r" % (lsqs_maxit, lsqs_eps),
              "DLSE %d %r %r" % dlse, "GAUS %d %d %d %d" % tuple(gauss), "IIDI %d" % iidiv, "REST 0"]
    return "\n".join(lines) + "\n"


def cavity(n, re=100.0, urdw=0.1, **kw):
    """Lid-driven cavity on an n^3 cube of 27-node cells."""
    return Host.cube(n).read_bic(bic_cavity()).read_inp(inp_text(re=re, urdw=urdw, **kw))


def natural_convection(n, ra=1e5, pr=0.71, urdw=0.1, urdt=0.05, **kw):
    """Differentially heated cavity (T = +0.5 at x = 0, -0.5 at x = 1) on an n^3 cube.

    The reference's run/nc relaxation (URDW 0.3, URDT 0.1) is set for Ra = 1e3;
    from rest at Ra = 1e5 that steady Picard iteration diverges, while
    URDW 0.1, URDT 0.05 converges (205 iterations to 1e-6 at 4^3)."""
    return Host.cube(n).read_bic(bic_heated()).read_inp(inp_text(rapr=(ra, pr), urdw=urdw, urdt=urdt, **kw))
