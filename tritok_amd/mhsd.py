"""Single-domain modified-Helmholtz path (EQNS imH 3-7 of the reference).

The operator (tt_mhsd_* of the C-ABI): integrals, system and its
factorisation, compression of the right-hand-side matrices and the solves.
Around it, the host side that the reference's Fortran keeps, restated here in
numpy: the manufactured problems of modHelm_SetProblem, setGRAD,
Int2MacroNodes, the time loop with its fixed-point iterations
(main.f90:1124-1395) and the RMS error of modHelm_calRMS2.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import MeshView, Quad, P, PD, PI, check, i32, dbl

_codes = _lib._sig("tt_host_mh_codes", i32, P, C.POINTER(PI), C.POINTER(PI), C.POINTER(PD), C.POINTER(PD),
                   C.POINTER(PI), C.POINTER(PI), C.POINTER(PI), C.POINTER(i32), C.POINTER(i32))
_create = _lib._sig("tt_mhsd_create", i32, C.POINTER(MeshView), C.POINTER(Quad), PI, PD, PD, dbl, C.POINTER(P), PD,
                    PD)
_create_from = _lib._sig("tt_mhsd_create_from", i32, C.POINTER(MeshView), PI, PD, PD, dbl, C.POINTER(PD),
                         C.POINTER(PD), C.POINTER(P))
_matrix = _lib._sig("tt_mhsd_matrix", i32, P, i32, i32, PD)
_system = _lib._sig("tt_mhsd_system", i32, P, PI, PI, i32, i32)
_compress = _lib._sig("tt_mhsd_compress", i32, P, i32, i32, dbl, dbl, i32, dbl, i32, PD)
_solve = _lib._sig("tt_mhsd_solve", i32, P, PD, PD, PD, PD, PD, dbl)
_stats = _lib._sig("tt_mhsd_stats", i32, P, PD, PD, C.POINTER(i32))
_free = _lib._sig("tt_mhsd_free", None, P)
_write_int = _lib._sig("tt_write_mhsd_int", i32, C.c_char_p, i32, i32, i32, i32, dbl, dbl, dbl, C.POINTER(PD),
                       C.POINTER(PD), PD)

NAMES = ["H", "G", "B", "Bp", "Abdx", "Abdy", "Abdz", "Dx", "Dy", "Dz"]
MODES = {"dense": 0, "wavelet": 1, "aca": 2, "hmat": 3}
# imH of the reference -> compression mode (imH 3 and 4 both dense)
IMH_MODE = {3: 0, 4: 0, 5: 1, 6: 2, 7: 3}


def _arr(ptr, n, dtype):
    ctype = C.c_double if dtype == np.float64 else C.c_int32
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy()


def codes(host):
    """tt_host_mh_codes of a Host after read_bic: dict(kode, qkode, u, q, ibcf,
    iwn, bewn, nunk, nknown); the flux-point arrays element-major."""
    ps = [PI(), PI(), PD(), PD(), PI(), PI(), PI()]
    nunk, nknown = i32(), i32()
    check(_codes(host._h, *[C.byref(p) for p in ps], C.byref(nunk), C.byref(nknown)), "tt_host_mh_codes")
    m = host.view().mesh
    nn, nb, ne = m.nnodes, m.nbnodes, m.nbelem
    return dict(kode=_arr(ps[0], nn, np.int32), qkode=_arr(ps[1], 4 * ne, np.int32), u=_arr(ps[2], nn, np.float64),
                q=_arr(ps[3], 4 * ne, np.float64), ibcf=_arr(ps[4], 4 * ne, np.int32),
                iwn=_arr(ps[5], nb, np.int32), bewn=_arr(ps[6], ne, np.int32), nunk=nunk.value,
                nknown=nknown.value)


# ---------------------------------------------------------------------------
# host side (numpy restatements of the reference's Fortran host routines)
def time_scheme(tstep):
    """SetTimeScheme, first-order scheme (lib_util.f:2-26): beta, beta2, beta3."""
    return 1.5 / tstep, 2.0 / tstep, 0.5 / tstep


def set_problem(m, cd, f, t, mhtime):
    """modHelm_SetProblem (lib_mHelm.f:655-843) for the manufactured problems
    f = 6..12: diffusivity, velocity, right-hand side mHtime - f, the analytic
    solution, and the boundary values it prescribes -- u = anal on the known
    nodes, the outward normal derivative on the known flux points by the wall
    of their element (walls 1/6: -/+ d/dz, 5/3: -/+ d/dx, 2/4: -/+ d/dy)."""
    x, y, z = m["x"][:, 0], m["x"][:, 1], m["x"][:, 2]
    xq, yq, zq = m["xq"][:, 0], m["xq"][:, 1], m["xq"][:, 2]
    nq = len(xq)
    if f == 6:
        diff = 1 + x * y * z
        vel = np.stack([x, y, -2.0 * z], axis=1)
        fs = -5.0 + 2.0 * x ** 2 + 2.0 * y ** 2 - 12.0 * x * y * z - 4.0 * z ** 2
        anal = x ** 2 + y ** 2 + z ** 2 + t
        grad = (2.0 * xq, 2.0 * yq, 2.0 * zq)
    elif 7 <= f <= 9:
        nu = {7: 0.5, 8: 1.5, 9: 2.5}[f] * np.pi
        diff = 2.0 + np.sin(nu * x * y * z)
        vel = np.stack([x, y, -2.0 * z], axis=1)
        fs = (-11.0 + 2.0 * x ** 2 + 2.0 * y ** 2 - 4.0 * z ** 2 - 6.0 * x * y * z * nu * np.cos(x * y * z * nu)
              - 6.0 * np.sin(x * y * z * nu))
        anal = x ** 2 + y ** 2 + z ** 2 + t
        grad = (2.0 * xq, 2.0 * yq, 2.0 * zq)
    elif 10 <= f <= 12:
        nu = {10: 2.0, 11: 5.0, 12: 10.0}[f]
        diff = 1.0 + x * y * z
        vel = np.stack([np.full_like(x, nu), np.ones_like(x), np.ones_like(x)], axis=1)
        if f == 10:
            fs = 1 + nu * (-1.0 + nu - x * nu + x * y * z * nu)
        else:
            fs = 1 + nu * (-1.0 + nu - x * nu + x * y * z * nu) * x ** (nu - 2.0)
        anal = t - x ** nu
        grad = (-nu * xq ** (nu - 1.0), np.zeros(nq), np.zeros(nq))
    else:
        raise ValueError("manufactured problem f must be 6..12")
    rhs = mhtime - fs
    u_known = np.where(cd["kode"] > 0, anal, np.nan)
    sign = {1: (2, -1.0), 6: (2, 1.0), 5: (0, -1.0), 3: (0, 1.0), 2: (1, -1.0), 4: (1, 1.0)}
    q_known = np.full(len(cd["qkode"]), np.nan)
    for k in np.nonzero(cd["qkode"] > 0)[0]:
        d, sg = sign[int(cd["bewn"][k // 4])]
        q_known[k] = sg * grad[d][cd["ibcf"][k] - 1]
    return dict(diff=np.ascontiguousarray(diff, np.float64), vel=np.asfortranarray(vel), rhs=rhs, anal=anal,
                u_known=u_known, q_known=q_known)


_LC = np.array([[0, 0, 1, 1, 0, 0, 1, 1, 0, .5, 1, .5, 0, 0, 1, 1, 0, .5, 1, .5, .5, 0, .5, 1, .5, .5, .5],
                [1, 1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, .5, .5, .5, .5, 0, 0, 0, 0, 1, .5, .5, .5, .5, 0, .5],
                [1, 0, 0, 1, 1, 0, 0, 1, .5, 0, .5, 1, 1, 0, 0, 1, .5, 0, .5, 1, .5, .5, 0, .5, 1, .5, .5]])
_GEO8 = np.array([6, 7, 2, 3, 5, 8, 1, 4]) - 1


def _lag(a, t):
    """the quadratic Lagrange factor of local position a in {-1, 0, 1} at t"""
    return np.where(a < 0, 0.5 * t * (t - 1.0), np.where(a > 0, 0.5 * t * (t + 1.0), (1.0 - t) * (1.0 + t)))


def _dlag(a, t):
    return np.where(a < 0, t - 0.5, np.where(a > 0, t + 0.5, -2.0 * t))


def shape27(xi, eta, zeta):
    """the 27-node shape functions (cshape27, lib_interpol.f:352-405) as products of 1-D factors"""
    a = 2.0 * _LC - 1.0
    return _lag(a[0], xi) * _lag(a[1], eta) * _lag(a[2], zeta)


def int2macro(m, diff):
    """Int2MacroNodes (lib_util.f:1377-1475): diffusivity at the flux nodes by
    27-node interpolation in each cell (later cells overwrite shared nodes)."""
    tc = 0.75
    # flux-point local positions, face by face (the reference's xsi/eta/ceta tables)
    xsi = [-tc, -tc, tc, tc, -tc, tc, tc, -tc, 1, 1, 1, 1, tc, -tc, -tc, tc, -1, -1, -1, -1, -tc, tc, tc, -tc]
    eta = [-tc, tc, tc, -tc, -1, -1, -1, -1, -tc, tc, tc, -tc, 1, 1, 1, 1, tc, -tc, -tc, tc, -tc, -tc, tc, tc]
    ceta = [-1, -1, -1, -1, -tc, -tc, tc, tc, -tc, -tc, tc, tc, -tc, -tc, tc, tc, -tc, -tc, tc, tc, 1, 1, 1, 1]
    fi = np.array([shape27(xsi[k], eta[k], ceta[k]) for k in range(24)])  # (24, 27)
    vals = diff[m["idc"] - 1] @ fi.T  # (nc, 24): face ii, point jj at column 4 ii + jj
    qdiff = np.zeros(m["nq"])
    ibf = m["ibf"]  # (nc, 6, 4)
    for ic in range(m["nc"]):
        for ii in range(6):
            for jj in range(4):
                qdiff[ibf[ic, ii, jj] - 1] = vals[ic, ii * 4 + jj]
    return qdiff


def cell_volumes(m):
    """IntCellVolume8 (lib_int.f:1003-1084): 4-point Gauss rule per direction."""
    g, w = np.polynomial.legendre.leggauss(4)
    X = m["x"][m["idc"][:, _GEO8] - 1]  # (nc, 8, 3)
    vol = np.zeros(m["nc"])
    for k in range(4):
        for j in range(4):
            for i in range(4):
                e1m, e1p, e2m, e2p, e3m, e3p = 1 - g[i], 1 + g[i], 1 - g[j], 1 + g[j], 1 - g[k], 1 + g[k]
                d1 = np.array([-e2m * e3m, e2m * e3m, -e2p * e3m, e2p * e3m, -e2m * e3p, e2m * e3p, -e2p * e3p,
                               e2p * e3p])
                d2 = np.array([-e1m * e3m, -e1p * e3m, e1m * e3m, e1p * e3m, -e1m * e3p, -e1p * e3p, e1m * e3p,
                               e1p * e3p])
                d3 = np.array([-e1m * e2m, -e1p * e2m, -e1m * e2p, -e1p * e2p, e1m * e2m, e1p * e2m, e1m * e2p,
                               e1p * e2p])
                J = np.stack([np.einsum("k,ckd->cd", d1, X), np.einsum("k,ckd->cd", d2, X),
                              np.einsum("k,ckd->cd", d3, X)], axis=1)  # (nc, 3, 3)
                vol += np.linalg.det(J) * 1.953125e-3 * w[i] * w[j] * w[k]
    return vol


def set_grad(m, u):
    """setGRAD (lib_interpol.f:124-211): nodal gradient of u as the
    volume-weighted mean of the 27-node cell derivatives at the node."""
    a = 2.0 * _LC - 1.0
    X = m["x"][m["idc"][:, _GEO8] - 1]  # (nc, 8, 3)
    vol = cell_volumes(m)
    grad = np.zeros((m["nn"], 3))
    svol = np.zeros(m["nn"])
    uc = u[m["idc"] - 1]  # (nc, 27)
    for it in range(27):
        xi, eta, zeta = a[0, it], a[1, it], a[2, it]
        dxi = _dlag(a[0], xi) * _lag(a[1], eta) * _lag(a[2], zeta)
        deta = _lag(a[0], xi) * _dlag(a[1], eta) * _lag(a[2], zeta)
        dzeta = _lag(a[0], xi) * _lag(a[1], eta) * _dlag(a[2], zeta)
        e1m, e1p, e2m, e2p, e3m, e3p = 1 - xi, 1 + xi, 1 - eta, 1 + eta, 1 - zeta, 1 + zeta
        d1 = 0.125 * np.array([-e2m * e3m, e2m * e3m, -e2p * e3m, e2p * e3m, -e2m * e3p, e2m * e3p, -e2p * e3p,
                               e2p * e3p])
        d2 = 0.125 * np.array([-e1m * e3m, -e1p * e3m, e1m * e3m, e1p * e3m, -e1m * e3p, -e1p * e3p, e1m * e3p,
                               e1p * e3p])
        d3 = 0.125 * np.array([-e1m * e2m, -e1p * e2m, -e1m * e2p, -e1p * e2p, e1m * e2m, e1p * e2m, e1m * e2p,
                               e1p * e2p])
        J = np.stack([np.einsum("k,ckd->cd", d1, X), np.einsum("k,ckd->cd", d2, X),
                      np.einsum("k,ckd->cd", d3, X)], axis=1)  # rows d/dxi, d/deta, d/dzeta of (x, y, z)
        Jinv = np.linalg.inv(J)  # (nc, 3, 3): d(xi, eta, zeta)/d(x, y, z) by columns
        dloc = np.stack([uc @ dxi, uc @ deta, uc @ dzeta], axis=1)  # (nc, 3)
        g = np.einsum("cij,cj->ci", Jinv, dloc)
        dn = m["idc"][:, it] - 1
        np.add.at(grad, dn, g * vol[:, None])
        np.add.at(svol, dn, vol)
    return np.asfortranarray(grad / svol[:, None])


def rms_error(m, u, anal):
    """modHelm_calRMS2 (lib_mHelm.f:617-650): relative RMS error over the domain nodes."""
    inner = m["lbn"] == 0
    return float(np.sqrt(np.sum((anal[inner] - u[inner]) ** 2) / np.sum(anal[inner] ** 2)))


# ---------------------------------------------------------------------------
class Operator:
    """tt_mhsd: integrals on creation, then system(), compress(), solve()."""

    def __init__(self, view, q, cd, diff, qdiff, beta, matrices=None):
        """q: tt_quad of the integrals, or matrices = dict(U=[10], Q=[10]) of
        integrals already computed (e.g. a reference integral file)."""
        m = view.mesh if hasattr(view, "mesh") else view
        self.nn, self.nb, self.ne, self.nq = m.nnodes, m.nbnodes, m.nbelem, m.nq
        self.nf = 4 * self.ne
        self._ibcf = np.ascontiguousarray(cd["ibcf"], np.int32)
        self._diff = np.ascontiguousarray(diff, np.float64)
        self._qdiff = np.ascontiguousarray(qdiff, np.float64)
        self.beta = float(beta)
        self.ont = np.zeros(5)
        mumax = C.c_double()
        h = P()
        args = (self._ibcf.ctypes.data_as(PI), self._diff.ctypes.data_as(PD), self._qdiff.ctypes.data_as(PD),
                self.beta)
        if matrices is None:
            check(_create(C.byref(m), C.byref(q), *args, C.byref(h), self.ont.ctypes.data_as(PD), C.byref(mumax)),
                  "tt_mhsd_create")
            self.mumax = mumax.value
        else:
            U = [np.asfortranarray(a, np.float64) for a in matrices["U"]]
            Q = [np.asfortranarray(a, np.float64) for a in matrices["Q"]]
            check(_create_from(C.byref(m), *args, (PD * 10)(*[a.ctypes.data_as(PD) for a in U]),
                               (PD * 10)(*[a.ctypes.data_as(PD) for a in Q]), C.byref(h)), "tt_mhsd_create_from")
            self.mumax = float(max(np.sqrt(self.beta / self._diff).max(),
                                   np.sqrt(self.beta / self._qdiff[self._ibcf - 1]).max()))
        self._h = h
        self.cd = cd

    def cols(self, k):
        return {0: self.nb, 1: self.nf, 4: self.nb, 5: self.nb, 6: self.nb}.get(k, self.nn)

    def matrix(self, set_, name):
        k = NAMES.index(name)
        rows = self.nn if set_ in (0, "U") else self.nf
        a = np.zeros((rows, self.cols(k)), order="F")
        check(_matrix(self._h, 0 if set_ in (0, "U") else 1, k, a.ctypes.data_as(PD)), "tt_mhsd_matrix")
        return a

    def system(self):
        kode = np.ascontiguousarray(self.cd["kode"], np.int32)
        qkode = np.ascontiguousarray(self.cd["qkode"], np.int32)
        check(_system(self._h, kode.ctypes.data_as(PI), qkode.ctypes.data_as(PI), self.cd["nunk"],
                      self.cd["nknown"]), "tt_mhsd_system")
        return self

    def compress(self, mode="dense", aca_type=2, aca_eps=1e-6, wt_kappa=1e-5, ct_type=3, ct_eta=1.0, leaf=64):
        r = C.c_double()
        check(_compress(self._h, MODES.get(mode, mode), int(aca_type), float(aca_eps), float(wt_kappa), int(ct_type),
                        float(ct_eta), int(leaf), C.byref(r)), "tt_mhsd_compress")
        return r.value

    def solve(self, u, qu, vel, rhsv, dgrad, urmh=1.0):
        """One solve; u (nn) and qu (4 ne) are updated in place."""
        assert u.dtype == np.float64 and qu.dtype == np.float64 and u.flags.c_contiguous and qu.flags.c_contiguous
        vel = np.asfortranarray(vel, np.float64)
        dg = np.asfortranarray(dgrad, np.float64)
        rhsv = np.ascontiguousarray(rhsv, np.float64)
        check(_solve(self._h, u.ctypes.data_as(PD), qu.ctypes.data_as(PD), vel.ctypes.data_as(PD),
                     rhsv.ctypes.data_as(PD), dg.ctypes.data_as(PD), float(urmh)), "tt_mhsd_solve")

    def stats(self):
        ms = np.zeros(4)
        r = C.c_double()
        n = i32()
        check(_stats(self._h, ms.ctypes.data_as(PD), C.byref(r), C.byref(n)), "tt_mhsd_stats")
        return dict(ms_integrals=ms[0], ms_system=ms[1], ms_compress=ms[2], ms_solve=ms[3], ratio=r.value,
                    nunk=n.value)

    def write_int(self, path, version, sumdiff):
        """tri.modHelm.full.int of this operator's matrices."""
        U = [self.matrix(0, k) for k in NAMES]
        Q = [self.matrix(1, k) for k in NAMES]
        up = (PD * 10)(*[a.ctypes.data_as(PD) for a in U])
        qp = (PD * 10)(*[a.ctypes.data_as(PD) for a in Q])
        check(_write_int(path.encode(), int(version), self.nn, self.nb, self.nf, self.beta, float(sumdiff),
                         self.mumax, up, qp, self.ont.ctypes.data_as(PD)), "tt_write_mhsd_int")

    def close(self):
        if self._h:
            _free(self._h)
            self._h = P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prepare(host, f, tstep):
    """The host-side set-up at t = 0 (main.f90:532-554, 1124-1130): codes,
    modHelm_SetProblem, setGRAD of the diffusivity, Int2MacroNodes, and the
    initial u and q with the prescribed values."""
    m = host.mesh_arrays()
    cd = codes(host)
    beta, beta2, beta3 = time_scheme(tstep)
    u = np.ascontiguousarray(cd["u"], np.float64).copy()
    qu = np.ascontiguousarray(cd["q"], np.float64).copy()
    p = set_problem(m, cd, f, 0.0, np.zeros(m["nn"]))
    u[cd["kode"] > 0] = p["u_known"][cd["kode"] > 0]
    qu[cd["qkode"] > 0] = p["q_known"][cd["qkode"] > 0]
    diff = p["diff"]
    return dict(m=m, cd=cd, f=f, tstep=tstep, beta=beta, beta2=beta2, beta3=beta3, diff=diff,
                qdiff=int2macro(m, diff), dgrad=set_grad(m, diff), u0=u, q0=qu)


def time_loop(op, pr, nstep, nnlit=3, eps=1e-12, urmh=1.0, on_step=None):
    """The reference's time loop for imH >= 3 (main.f90:1618-1671): per step
    the time-derivative source, up to nnlit fixed-point iterations
    (SetProblem, solve, the change check of CheckConvergencevw's mH entry,
    lib_util.f:1048-1150) and the RMS error.  Returns per-step dicts (t, u, q,
    rms, changes, solve_ms: device time of each solve)."""
    m, cd, f = pr["m"], pr["cd"], pr["f"]
    u, qu = pr["u0"].copy(), pr["q0"].copy()
    pts, ppts, pnls = u.copy(), u.copy(), u.copy()
    t, out = 0.0, []
    for step in range(nstep):
        t = t + pr["tstep"]
        ppts, pts = pts, u.copy()
        mhtime = -pts * pr["beta2"] + ppts * pr["beta3"]
        changes, ms = [], []
        for it in range(nnlit):
            p = set_problem(m, cd, f, t, mhtime)
            u[cd["kode"] > 0] = p["u_known"][cd["kode"] > 0]
            qu[cd["qkode"] > 0] = p["q_known"][cd["qkode"] > 0]
            op.solve(u, qu, p["vel"], p["rhs"], pr["dgrad"], urmh)
            ms.append(op.stats()["ms_solve"])
            dif = np.sum((u - pnls) ** 2)
            nrm = np.sum(u ** 2)
            ch = float(np.sqrt(dif / nrm)) if nrm > eps else 0.0
            pnls = u.copy()
            changes.append(ch)
            if ch < eps:
                break
        rec = dict(t=t, u=u.copy(), q=qu.copy(), rms=rms_error(m, u, p["anal"]), changes=changes, solve_ms=ms)
        out.append(rec)
        if on_step:
            on_step(rec)
    return out


def run(host, q, f, tstep, nstep, nnlit=3, eps=1e-12, urmh=1.0, mode="dense", on_step=None, matrices=None,
        **comp):
    """prepare(), the operator (integrals or given matrices, system,
    compression by mode) and time_loop().  Returns the per-step records and
    the operator's timings."""
    pr = prepare(host, f, tstep)
    with Operator(host.view(), q, pr["cd"], pr["diff"], pr["qdiff"], pr["beta"], matrices=matrices) as op:
        op.system()
        ratio = op.compress(mode, **comp)
        out = time_loop(op, pr, nstep, nnlit, eps, urmh, on_step)
        st = op.stats()
    st["ratio"] = ratio
    return out, st
