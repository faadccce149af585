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


def bic_mh(dirichlet=(1, 3, 5), neumann=(2, 4, 6)):
    """Modified-Helmholtz boundary conditions (MHEL block, lib_bic.f:366-395):
    u prescribed on the `dirichlet` walls, the normal flux on the `neumann`
    walls; edges and corners belong to walls as in the cavity files.  The
    values written are placeholders: the manufactured problems set them."""
    lines = ["# modified Helmholtz, single domain", "EDGE"] + ["%d %d %d" % e for e in _EDGES_CAVITY]
    lines += ["END", "CORNER"] + ["%d %d %d %d" % c for c in _CORNERS_CAVITY] + ["END", "MHEL"]
    lines += ["1 %d 0.0" % w for w in dirichlet] + ["2 %d 0.0" % w for w in neumann] + ["END"]
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


def cell_batch(n_side=16, jitter=0.1, seed=20251203, mu_range=(0.1, 100.0), beta=1.0):
    """Batch of n_side^3 independent 27-node cells for the batched
    modified-Helmholtz integrals: cell (i, j, k) of an n_side^3 arrangement with
    edge h = 1/n_side, its eight corners moved by U(-jitter*h, jitter*h) per
    coordinate, its other nodes and flux points placed trilinearly from the
    corners (straight-edged cells, as the reference's CheckLinearGeometry,
    lib_mesh.f:135-196, leaves them), and mu log-uniform in mu_range per cell,
    all from one seeded stream.  Cells share no nodes, so each cell's mu is its
    nodes' and flux points' mu: diff = beta/mu^2 at all of them.

    Numbering is the reference's for one cell (the 1-cell cube), repeated with
    offsets.  Returns dict(arrays=mesh arrays as Host.mesh_arrays(), diff,
    qdiff, beta, mu (per cell))."""
    import numpy as np
    one = Host.cube(1).mesh_arrays()
    nc = n_side ** 3
    h = 1.0 / n_side
    rng = np.random.default_rng(seed)
    ijk = np.stack(np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side), indexing="ij"),
                   -1).reshape(-1, 3)[:, ::-1].astype(float)  # i fastest
    # corners of the unit cell in trilinear order (s0 fastest)
    unit = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], float)
    corners = (ijk[:, None, :] + unit[None]) * h + rng.uniform(-jitter * h, jitter * h, size=(nc, 8, 3))
    mu = np.exp(rng.uniform(np.log(mu_range[0]), np.log(mu_range[1]), size=nc))

    def place(local):  # (p, 3) local coordinates in [0,1]^3 -> (nc, p, 3)
        u, v, w = local[:, 0], local[:, 1], local[:, 2]
        N = np.stack([(u if a else 1 - u) * (v if b else 1 - v) * (w if c else 1 - w)
                      for c in (0, 1) for b in (0, 1) for a in (0, 1)], 1)  # (p, 8)
        return np.einsum("pk,ckd->cpd", N, corners)

    nn1, nq1, nb1 = len(one["x"]), len(one["xq"]), len(one["gbn"])
    x = place(one["x"]).reshape(-1, 3)
    xq = place(one["xq"]).reshape(-1, 3)
    off = np.arange(nc, dtype=np.int32)
    arrays = dict(
        x=x, xq=xq,
        idc=(one["idc"][0][None, :] + (off * nn1)[:, None]).astype(np.int32),
        ibf=(one["ibf"][0][None] + (off * nq1)[:, None, None]).astype(np.int32),
        flpm=np.repeat(one["flpm"], nc, axis=0).astype(np.int32),
        ibc=(one["ibc"][None] + (off * nn1)[:, None, None]).reshape(-1, 9).astype(np.int32),
        gbn=(one["gbn"][None, :] + (off * nn1)[:, None]).reshape(-1).astype(np.int32))
    lbn = np.zeros(nn1 * nc, np.int32)
    lbn[arrays["gbn"] - 1] = np.arange(1, nb1 * nc + 1, dtype=np.int32)
    arrays["lbn"] = lbn
    d = beta / mu ** 2
    return dict(arrays=arrays, diff=np.repeat(d, nn1), qdiff=np.repeat(d, nq1), beta=float(beta), mu=mu)


def serving_stream(count=2000, rate=20.0, seed=20251203, sizes=(6, 8, 12, 16)):
    """Seeded request stream of the serving workload: Poisson arrivals at
    `rate` per second, cube size uniform in `sizes`, half natural convection
    (Ra log-uniform in [1e3, 1e6], Pr = 0.71) and half lid-driven cavity (Re
    uniform in [10, 400]), each to a Picard change of 1e-6."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    out = []
    for k in range(count):
        n = int(rng.choice(sizes))
        if rng.random() < 0.5:
            out.append(dict(t=float(t[k]), kind="convection", n=n, ra=float(10 ** rng.uniform(3.0, 6.0))))
        else:
            out.append(dict(t=float(t[k]), kind="cavity", n=n, re=float(rng.uniform(10.0, 400.0))))
    return out


def serving_case(spec, **kw):
    """Host inputs of one request of serving_stream().  Convection runs
    with vorticity / temperature relaxation 0.1 / 0.05 up to Ra = 1e5, half
    that up to Ra = 4e5 and a quarter above (the Picard iteration of the
    coarse meshes diverges near Ra = 1e6 at the Ra = 1e5 setting);
    cavities use the relaxation 0.1."""
    if spec["kind"] == "convection":
        ra = spec["ra"]
        f = 1.0 if ra <= 1e5 else (0.5 if ra <= 4e5 else 0.25)
        return natural_convection(spec["n"], ra=ra, pr=0.71, urdw=0.1 * f, urdt=0.05 * f, **kw)
    return cavity(spec["n"], re=spec["re"], urdw=0.1, **kw)


def vte_csr(host, macro, re, eq=0):
    """See vte_csr_parts: the values for Reynolds number re."""
    ia, ja, ah, ab, n, beta = vte_csr_parts(host, macro, eq)
    return ia, ja, ah + ab * beta * re, n


def vte_csr_parts(host, macro, eq=0):
    """The vorticity-transport system of component eq (0, 1, 2) as the
    reference assembles it (sMat2crsSysRhsB_w_fill, lib_solvFDwTE.f:593-752):
    per cell row, H + B*beta*Re on the unknown node columns and -G on the
    unknown flux columns, columns in increasing unknown order, 1-based CSR.
    macro: the dict of integrals.macro_integrals for the host's mesh.
    Returns (ia, ja, a_H, a_B, n, beta): the values are a_H + a_B * beta * Re
    (a_B zero on flux columns), in the reference's order of operations."""
    import numpy as np
    v = host.view()
    m = v.mesh
    nc, nn, nq = m.nicell, m.nnodes, m.nq
    arr = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy()
    idc = arr(m.idc, nc * 27).reshape(27, nc).T - 1
    ibf = arr(m.ibf, nc * 24).reshape(4, 6, nc).transpose(2, 1, 0).reshape(nc, 24) - 1
    wk = arr(v.wkode, nn * 3).reshape(3, nn)[eq]
    wq = arr(v.wqkode, nq * 3).reshape(3, nq)[eq]
    beta = 1.5 / v.tstep
    codes = np.concatenate([wk[idc], wq[ibf]], axis=1)  # (nc, 51)
    vh = np.concatenate([macro["H"], -macro["G"]], axis=1).reshape(nc, 51, 51)
    vb = np.concatenate([macro["B"], np.zeros_like(macro["G"])], axis=1).reshape(nc, 51, 51)
    ja, ah, ab, counts = [], [], [], []
    for c in range(nc):
        unk = np.nonzero(codes[c] < 0)[0]
        order = unk[np.argsort(-codes[c][unk], kind="stable")]
        ja.append(np.tile(-codes[c][order], 51))
        ah.append(vh[c][:, order].ravel())
        ab.append(vb[c][:, order].ravel())
        counts.append(np.full(51, len(order)))
    ia = np.concatenate([[1], 1 + np.cumsum(np.concatenate(counts))]).astype(np.int32)
    return (ia, np.concatenate(ja).astype(np.int32), np.concatenate(ah), np.concatenate(ab), int(v.nunk[eq]), beta)
