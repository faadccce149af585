"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("oracle/liboracle.so not built (make -C oracle)")
lib = C.CDLL(LIB_PATH)

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)
_keep = []  # tables referenced by quadrature structs


class OrQuad(C.Structure):
    _fields_ = [("kmBr", C.c_int), ("kmBs", C.c_int), ("kmDr", C.c_int), ("kmDs", C.c_int), ("iiDiv", C.c_int),
                ("gi", PD), ("ome", PD), ("ng1", PI), ("ng2", PI)]


class OrCase(C.Structure):
    _fields_ = [("nn", C.c_int), ("nb", C.c_int), ("ne", C.c_int), ("nc", C.c_int), ("nq", C.c_int),
                ("x", PD), ("idc", PI), ("ibc", PI), ("ibf", PI), ("flpm", PI), ("gbn", PI), ("lbn", PI),
                ("xq", PD), ("nx", PD), ("ny", PD), ("nz", PD),
                ("kode", PI), ("kobc", PI), ("wkode", PI), ("wqkode", PI), ("Tkode", PI), ("Tqkode", PI),
                ("nunk", C.c_int * 3), ("nbk", C.c_int * 3), ("Tnunk", C.c_int), ("Tnb", C.c_int),
                ("kmu", PI), ("nsmcol", C.c_int), ("nrhscol", C.c_int),
                ("iDT", C.c_int), ("ReN", C.c_double), ("RaN", C.c_double), ("PrN", C.c_double),
                ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double), ("tstep", C.c_double),
                ("urBw", C.c_double * 3), ("urDv", C.c_double * 3), ("urDw", C.c_double * 3), ("urDT", C.c_double),
                ("eps", C.c_double), ("nnlit", C.c_int), ("lsqs_maxit", C.c_int), ("iDLSE", C.c_int),
                ("lsqs_eps", C.c_float), ("dlse_r", C.c_float), ("dlse_epsth", C.c_float),
                ("quad", OrQuad), ("copy", C.c_int), ("iBw", C.c_int)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_gauss = _sig("or_gauss", C.c_int, C.c_int, PD, PD)
_shape27 = _sig("or_shape27", None, C.c_double, C.c_double, C.c_double, PD)
_shape9 = _sig("or_shape9", None, C.c_double, C.c_double, PD)
_dl34 = _sig("or_dl34", None, C.c_double, C.c_double, PD)
_inted27 = _sig("or_inted27", C.c_long, C.POINTER(OrQuad), PD, PD, C.c_int, PD, PD, PD, PD)
_inteb9 = _sig("or_inteb9", C.c_long, C.POINTER(OrQuad), C.c_int, PD, PD, C.c_int, C.c_double, PD, PD, PD, PD, PD,
               PD, PD, PD)
_macro = _sig("or_macro_integrals", None, C.POINTER(OrQuad), C.c_int, C.c_int, C.c_int, PD, PI, PI, PD, PI, C.c_int,
              *([PD] * 12), PD, C.c_int)
_macro_mh = _sig("or_macro_integrals_mh", None, C.POINTER(OrQuad), C.c_int, C.c_int, C.c_int, PD, PI, PI, PD, PI,
                 C.c_int, PD, PD, C.c_double, *([PD] * 12), PD, PD, C.c_int)
_inted27_mh = _sig("or_inted27_mh", C.c_long, C.POINTER(OrQuad), PD, PD, C.c_int, C.c_double, PD, PD, PD, PD)
_inteb9_mh = _sig("or_inteb9_mh", C.c_long, C.POINTER(OrQuad), PD, PD, C.c_int, C.c_double, C.c_double, PD, PD, PD,
                  PD, PD, PD, PD, PD)
_lastpts = _sig("or_last_points", None, C.POINTER(C.c_long), C.POINTER(C.c_long))
_sd = _sig("or_sd_integrals", None, C.POINTER(OrQuad), C.c_int, C.c_int, C.c_int, C.c_int, PD, PI, PI, PI, PI,
           C.c_int, C.c_int, C.c_int, *([PD] * 7), PD, C.c_int)
_dnrm2 = _sig("or_dnrm2", C.c_double, C.c_int, PD)
_prelsqr2 = _sig("or_prelsqr2", None, C.c_int, C.c_int, PI, PI, PD, PD)
_slvlsqr2 = _sig("or_slvlsqr2", C.c_int, C.c_int, C.c_int, PI, PI, PD, PD, PD, PD, C.c_int, C.c_float,
                 C.POINTER(C.c_int))
_dgefa = _sig("or_dgefa", C.c_int, PD, C.c_int, C.c_int, PD)
_dgesl = _sig("or_dgesl", None, PD, C.c_int, C.c_int, PD, PD)
_picard = _sig("or_picard", C.c_int, C.POINTER(OrCase), PD, PD, PD, PD, PD, C.c_int, PD, PI, C.c_int, PD)
_sdkm_rhs = _sig("or_sdkm_rhs", None, C.c_int, C.c_int, C.c_int, PI, PI, PI, PD, PD, PD, *([PD] * 7), C.c_int, PD, PD,
                 PD, C.c_int)


def _d(a):
    return a.ctypes.data_as(PD)


def _i(a):
    return a.ctypes.data_as(PI)


def quad(kmBr=4, kmBs=4, kmDr=4, kmDs=4, iiDiv=2, table=None):
    """table: dict(gi, ome, ng1, ng2) of the reference's GaussType, or None."""
    q = OrQuad(kmBr, kmBs, kmDr, kmDs, iiDiv)
    if table is not None:
        t = [np.ascontiguousarray(table["gi"], np.float64), np.ascontiguousarray(table["ome"], np.float64),
             np.ascontiguousarray(table["ng1"], np.int32), np.ascontiguousarray(table["ng2"], np.int32)]
        _keep.append(t)
        q.gi, q.ome, q.ng1, q.ng2 = _d(t[0]), _d(t[1]), _i(t[2]), _i(t[3])
    return q


def gauss(rule):
    gi, om = np.zeros(64), np.zeros(64)
    n = _gauss(rule, _d(gi), _d(om))
    return gi[:n].copy(), om[:n].copy()


def shape27(xi, eta, zeta):
    f = np.zeros(27)
    _shape27(xi, eta, zeta, _d(f))
    return f


def shape9(xi, eta):
    f = np.zeros(9)
    _shape9(xi, eta, _d(f))
    return f


def dl34(xi, eta):
    f = np.zeros(4)
    _dl34(xi, eta, _d(f))
    return f


def inted27(q, p, corners, isrc):
    p = np.ascontiguousarray(p, np.float64)
    c = np.ascontiguousarray(corners, np.float64)
    out = [np.zeros(27) for _ in range(4)]
    n = _inted27(C.byref(q), _d(p), _d(c), int(isrc), *[_d(o) for o in out])
    return (*out, n)


def inteb9(q, macro, p, corners, isrc, minedge=1.0):
    p = np.ascontiguousarray(p, np.float64)
    c = np.ascontiguousarray(corners, np.float64)
    he, ge, ax, ay, az, hyz, hzx, hxy = [np.zeros(9) for _ in range(8)]
    n = _inteb9(C.byref(q), int(macro), _d(p), _d(c), int(isrc), float(minedge), _d(he), _d(ge), _d(ax), _d(ay),
                _d(az), _d(hyz), _d(hzx), _d(hxy))
    return dict(he=he, ge=ge[:4], adx=ax, ady=ay, adz=az, heyz=hyz, hezx=hzx, hexy=hxy, n=n)


def macro_integrals(q, m, copy=0, nthreads=8):
    """m: dict from Host.mesh_arrays().  Returns dict of (rows, cols) arrays."""
    nn, nc, nq = m["nn"], m["nc"], m["nq"]
    x = np.asfortranarray(m["x"])
    idc = np.asfortranarray(m["idc"])
    ibf = np.asfortranarray(m["ibf"])
    xq = np.asfortranarray(m["xq"])
    flpm = np.asfortranarray(m["flpm"])
    m1, m2 = nc * 51, nc * 27
    names = ["H", "G", "B", "Abdx", "Abdy", "Abdz", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz"]
    shp = {"G": (m1, 24), "Htx": (m2, 27), "Hty": (m2, 27), "Htz": (m2, 27)}
    out = {k: np.zeros(shp.get(k, (m1, 27)), order="F") for k in names}
    ont = np.zeros(11)
    _macro(C.byref(q), nn, nc, nq, _d(x), _i(idc), _i(ibf), _d(xq), _i(flpm), int(copy), *[_d(out[k]) for k in names],
           _d(ont), nthreads)
    out["ont"] = ont
    return out


def macro_integrals_mh(q, m, diff, qdiff, beta, copy=0, nthreads=8):
    """FMATmodHelm: diff(nn), qdiff(nq) per-node diffusivities, mu = sqrt(beta/diff).
    Returns the dict of macro_integrals plus 'mumax'."""
    nn, nc, nq = m["nn"], m["nc"], m["nq"]
    x = np.asfortranarray(m["x"])
    idc = np.asfortranarray(m["idc"])
    ibf = np.asfortranarray(m["ibf"])
    xq = np.asfortranarray(m["xq"])
    flpm = np.asfortranarray(m["flpm"])
    diff = np.ascontiguousarray(diff, np.float64)
    qdiff = np.ascontiguousarray(qdiff, np.float64)
    assert diff.shape == (nn,) and qdiff.shape == (nq,)
    m1, m2 = nc * 51, nc * 27
    names = ["H", "G", "B", "Abdx", "Abdy", "Abdz", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz"]
    shp = {"G": (m1, 24), "Htx": (m2, 27), "Hty": (m2, 27), "Htz": (m2, 27)}
    out = {k: np.zeros(shp.get(k, (m1, 27)), order="F") for k in names}
    ont = np.zeros(11)
    mumax = C.c_double()
    _macro_mh(C.byref(q), nn, nc, nq, _d(x), _i(idc), _i(ibf), _d(xq), _i(flpm), int(copy), _d(diff), _d(qdiff),
              float(beta), *[_d(out[k]) for k in names], _d(ont), C.byref(mumax), nthreads)
    out["ont"] = ont
    out["mumax"] = mumax.value
    return out


def inted27_mh(q, p, corners, isrc, mu):
    p = np.ascontiguousarray(p, np.float64)
    c = np.ascontiguousarray(corners, np.float64)
    out = [np.zeros(27) for _ in range(4)]
    n = _inted27_mh(C.byref(q), _d(p), _d(c), int(isrc), float(mu), *[_d(o) for o in out])
    return (*out, n)


def inteb9_mh(q, p, corners, isrc, mu, minedge=1.0):
    p = np.ascontiguousarray(p, np.float64)
    c = np.ascontiguousarray(corners, np.float64)
    he, ge, ax, ay, az, hyz, hzx, hxy = [np.zeros(9) for _ in range(8)]
    n = _inteb9_mh(C.byref(q), _d(p), _d(c), int(isrc), float(minedge), float(mu), _d(he), _d(ge), _d(ax), _d(ay),
                   _d(az), _d(hyz), _d(hzx), _d(hxy))
    return dict(he=he, ge=ge[:4], adx=ax, ady=ay, adz=az, heyz=hyz, hezx=hzx, hexy=hxy, n=n)


def sd_integrals(q, m, row0=0, row1=None, nthreads=8):
    nn, nb, ne, nc = m["nn"], m["nb"], m["ne"], m["nc"]
    row1 = nb if row1 is None else row1
    r = row1 - row0
    x = np.asfortranarray(m["x"])
    ibc = np.asfortranarray(m["ibc"])
    idc = np.asfortranarray(m["idc"])
    gbn = np.ascontiguousarray(m["gbn"])
    lbn = np.ascontiguousarray(m["lbn"])
    names = ["H", "Htx", "Hty", "Htz", "Dx", "Dy", "Dz"]
    out = {k: np.zeros((r, nb if k.startswith("H") else nn), order="F") for k in names}
    ont = np.zeros(10)
    _sd(C.byref(q), nn, nb, ne, nc, _d(x), _i(ibc), _i(idc), _i(gbn), _i(lbn), row0, row1, r,
        *[_d(out[k]) for k in names], _d(ont), nthreads)
    out["ont"] = ont
    return out


def dnrm2(x):
    x = np.ascontiguousarray(x, np.float64)
    return _dnrm2(len(x), _d(x))


def prelsqr2(m, n, ia, ja, a):
    """In-place column scaling of a 1-based CSR; returns q."""
    q = np.zeros(n)
    _prelsqr2(m, n, _i(ia), _i(ja), _d(a), _d(q))
    return q


def slvlsqr2(m, n, ia, ja, a, q, b, x0, maxit, eps4):
    x = np.array(x0, np.float64)
    nit = C.c_int()
    istop = _slvlsqr2(m, n, _i(ia), _i(ja), _d(a), _d(q), _d(np.ascontiguousarray(b, np.float64)), _d(x), maxit,
                      float(eps4), C.byref(nit))
    return x, istop, nit.value


def dgefa(a):
    a = np.array(a, np.float64, order="F")
    n = a.shape[0]
    piv = np.zeros(n)
    info = _dgefa(_d(a), n, n, _d(piv))
    return a, piv, info


def dgesl(lu, piv, b):
    b = np.array(b, np.float64)
    _dgesl(_d(lu), lu.shape[0], lu.shape[0], _d(piv), _d(b))
    return b


def case_from_view(v):
    """OrCase from a tritok_amd CaseView (same arrays, oracle field order)."""
    m = v.mesh
    c = OrCase()
    c.nn, c.nb, c.ne, c.nc, c.nq = m.nnodes, m.nbnodes, m.nbelem, m.nicell, m.nq
    c.x, c.idc, c.ibc, c.ibf, c.flpm, c.gbn, c.lbn = (C.cast(m.x, PD), C.cast(m.idc, PI), C.cast(m.ibc, PI),
                                                      C.cast(m.ibf, PI), C.cast(m.flpm, PI), C.cast(m.gbn, PI),
                                                      C.cast(m.lbn, PI))
    c.xq, c.nx, c.ny, c.nz = C.cast(m.xq, PD), C.cast(m.nx, PD), C.cast(m.ny, PD), C.cast(m.nz, PD)
    for f in ["kode", "kobc", "wkode", "wqkode", "Tkode", "Tqkode", "kmu"]:
        setattr(c, f, C.cast(getattr(v, f), PI))
    for k in range(3):
        c.nunk[k], c.nbk[k] = v.nunk[k], v.nbk[k]
        c.urBw[k], c.urDv[k], c.urDw[k] = v.urBw[k], v.urDv[k], v.urDw[k]
    for f in ["Tnunk", "Tnb", "nsmcol", "nrhscol", "iDT", "ReN", "RaN", "PrN", "gx", "gy", "gz", "tstep", "urDT",
              "eps", "nnlit", "lsqs_maxit", "iDLSE", "lsqs_eps", "dlse_r", "dlse_epsth", "copy", "iBw"]:
        setattr(c, f, getattr(v, f))
    c.quad = OrQuad(v.quad.kmBr, v.quad.kmBs, v.quad.kmDr, v.quad.kmDs, v.quad.iiDiv,
                    C.cast(v.quad.gi, PD), C.cast(v.quad.ome, PD), C.cast(v.quad.ng1, PI), C.cast(v.quad.ng2, PI))
    return c


def picard(case, fields, maxhist=2000, nthreads=8):
    """Runs the oracle's Picard loop on copies of the fields; returns
    (fields_out, sta[its,9], ite[its,12], seconds_per_iteration)."""
    f = {k: np.asfortranarray(np.array(v, np.float64)) for k, v in fields.items()}
    sta = np.zeros((maxhist, 9))
    ite = np.zeros((maxhist, 12), np.int32)
    spi = C.c_double()
    its = _picard(C.byref(case), _d(f["velocity"]), _d(f["vorticity"]), _d(f["qvorticity"]), _d(f["temp"]),
                  _d(f["qtemp"]), maxhist, _d(sta), _i(ite), nthreads, C.byref(spi))
    n = min(its, maxhist)
    return f, sta[:n], ite[:n], spi.value


# ---------------------------------------------------------------- compressed
_aca = _sig("or_aca", C.c_int, C.c_int, C.c_int, C.c_int, PD, C.c_double, PD, PD, PD, PD)
_aca_abxv = _sig("or_aca_abxv", None, C.c_int, C.c_int, C.c_int, PD, PD, PD, PD)
_dvanak = _sig("or_dvanak", C.c_int, C.c_int)
_wxv = _sig("or_wt_wxv", None, C.c_int, PD)
_wtxv = _sig("or_wt_wtxv", None, C.c_int, PD)
_wawt = _sig("or_wt_wawt", None, C.c_int, C.c_int, PD)
_wthr = _sig("or_wt_threshold", C.c_double, C.c_int, C.c_int, PD, C.c_double, PI)
_wcsr = _sig("or_wt_form_csr", C.c_int, C.c_int, C.c_int, PD, PI, PI, PD)
_wtawx = _sig("or_wt_wtawx", None, C.c_int, C.c_int, PI, PI, PD, PD, PD)
PL = C.POINTER(C.c_long)
_hmxv = _sig("or_hmat_mxv", None, C.c_int, C.c_int, PI, PI, PI, PI, PI, PI, PL, PD, PL, PI, PD, PD)


def aca(type_, A, eps, vec=None):
    """Cross approximation of A (copied).  Returns (a, b, rank, norm, residual)."""
    R = np.array(A, np.float64, order="F")
    m, n = R.shape
    mn = min(m, n)
    a = np.zeros((m, mn), order="F")
    b = np.zeros(mn * n)
    norm = C.c_double(np.nan)
    v = None if vec is None else np.ascontiguousarray(vec, np.float64)
    r = _aca(int(type_), m, n, _d(R), float(eps), None if v is None else _d(v), _d(a), _d(b), C.byref(norm))
    return (np.asfortranarray(a[:, :r]), b[:r * n].reshape((n, r)).T.copy(order="F"), r, norm.value, R)


def aca_abxv(a, b, x):
    a = np.asfortranarray(a, np.float64)
    b = np.asfortranarray(b, np.float64)
    y = np.zeros(a.shape[0])
    _aca_abxv(a.shape[0], b.shape[1], a.shape[1], _d(a), _d(b), _d(np.ascontiguousarray(x, np.float64)), _d(y))
    return y


def dvanak(n):
    return _dvanak(int(n))


def wt_wxv(x):
    r = np.array(x, np.float64)
    _wxv(len(r), _d(r))
    return r


def wt_wtxv(x):
    r = np.array(x, np.float64)
    _wtxv(len(r), _d(r))
    return r


def wt_wawt(A):
    W = np.array(A, np.float64, order="F")
    _wawt(W.shape[0], W.shape[1], _d(W))
    return W


def wt_compress(A, kappa):
    """Che_WcomMatrix without the random check: returns dict(zvr, c, v, nnz_ref,
    tresh, T) with T the thresholded transformed matrix."""
    T = wt_wawt(A)
    m, n = T.shape
    nnz = C.c_int()
    tresh = _wthr(m, n, _d(T), float(kappa), C.byref(nnz))
    cnt = int(np.count_nonzero(T))
    zvr = np.zeros(m + 1, np.int32)
    c = np.zeros(max(cnt, 1), np.int32)
    v = np.zeros(max(cnt, 1))
    st = _wcsr(m, n, _d(T), _i(zvr), _i(c), _d(v))
    return dict(zvr=zvr, c=c[:st].copy(), v=v[:st].copy(), nnz_ref=nnz.value, tresh=tresh, T=T)


def wt_wtawx(w, x, nrow):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(nrow)
    _wtawx(nrow, len(x), _i(w["zvr"]), _i(w["c"]), _d(w["v"]), _d(x), _d(y))
    return y


def hmat_mxv(nrow, parts, x):
    """parts: list of dict(rows, cols (1-based arrays), kind in 'dense'/'wt'/'aca',
    dense: (nr, nc) array; wt: wt_compress dict; aca: (a, b))."""
    rptr, cptr, rn, cn, ct, rk, doff, ioff, dp, ip = [0], [0], [], [], [], [], [], [], [], []
    dpos = ipos = 0
    for p in parts:
        rn.append(np.asarray(p["rows"], np.int32))
        cn.append(np.asarray(p["cols"], np.int32))
        rptr.append(rptr[-1] + len(p["rows"]))
        cptr.append(cptr[-1] + len(p["cols"]))
        doff.append(dpos)
        ioff.append(ipos)
        if p["kind"] == "dense":
            d = np.asfortranarray(p["dense"], np.float64).ravel(order="F")
            ct.append(1), rk.append(0), dp.append(d)
            dpos += len(d)
        elif p["kind"] == "wt":
            w = p["wt"]
            ct.append(2), rk.append(0), dp.append(w["v"]), ip.append(w["zvr"]), ip.append(w["c"])
            dpos += len(w["v"])
            ipos += len(w["zvr"]) + len(w["c"])
        else:
            a, b = p["aca"]
            d = np.concatenate([np.asfortranarray(a).ravel(order="F"), np.asfortranarray(b).ravel(order="F")])
            ct.append(3), rk.append(a.shape[1]), dp.append(d)
            dpos += len(d)
    arr = lambda v, t: np.ascontiguousarray(np.concatenate(v) if v else np.zeros(1), t)
    rptr, cptr = np.asarray(rptr, np.int32), np.asarray(cptr, np.int32)
    rn, cn, ct, rk = arr(rn, np.int32), arr(cn, np.int32), np.asarray(ct, np.int32), np.asarray(rk, np.int32)
    doff, ioff = np.asarray(doff, np.int64), np.asarray(ioff, np.int64)
    dp, ip = arr(dp, np.float64), arr(ip, np.int32)
    y = np.zeros(nrow)
    x = np.ascontiguousarray(x, np.float64)
    _hmxv(nrow, len(parts), _i(rptr), _i(rn), _i(cptr), _i(cn), _i(ct), _i(rk), doff.ctypes.data_as(PL), _d(dp),
          ioff.ctypes.data_as(PL), _i(ip), _d(x), _d(y))
    return y


def last_points():
    """(domain, boundary) kernel evaluations of the last macro assembly."""
    d, b = C.c_long(), C.c_long()
    _lastpts(C.byref(d), C.byref(b))
    return d.value, b.value


# ---- single-domain modified Helmholtz (lib_fm.f) ----------------------------
MHSD_NAMES = ["H", "G", "B", "Bp", "Abdx", "Abdy", "Abdz", "Dx", "Dy", "Dz"]
_mhsd_int = _sig("or_mhsd_integrals", None, C.POINTER(OrQuad), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, PD, PI,
                 PI, PI, PD, PI, PI, PI, PD, PD, C.c_double, C.POINTER(PD), C.POINTER(PD), PD, PD, C.c_int)
_mhsd_bprime = _sig("or_mhsd_bprime", None, C.c_int, C.c_int, PI, PD, PD, C.c_double, PD, PD, PD, PD)
_mhsd_sysm = _sig("or_mhsd_sysm", None, C.c_int, C.c_int, C.c_int, PI, PI, PI, PI, C.c_int, C.c_int, PD, PD, PD, PD,
                  PD, PD, PD, PD)
_mhsd_rhs = _sig("or_mhsd_rhs", None, C.c_int, C.c_int, C.c_int, PI, PI, PI, PI, C.c_int, C.c_int, PD, PD,
                 C.c_double, PD, PD, PD, PD, PD, C.POINTER(PD), C.POINTER(PD), PD, PD)
_mhsd_vectors = _sig("or_mhsd_vectors", None, C.c_int, C.c_int, PI, PD, PD, PD, PD, C.POINTER(PD))
_mhsd_rhs_mv = _sig("or_mhsd_rhs_mv", None, C.c_int, C.c_int, PI, PI, C.c_int, C.c_int, PD, PD, PD, C.POINTER(PD),
                    C.POINTER(PD), PD)


def _ibcf_cm(ibcf_em, ne):
    """element-major flux-node ids (4 ne) -> the reference's (ne, 4) column-major"""
    return np.ascontiguousarray(np.asarray(ibcf_em, np.int32).reshape(ne, 4).T.reshape(-1))


def mhsd_integrals(q, m, ibcf_em, diff, qdiff, beta, nthreads=8):
    """FMATsdmHfs + MakeBprime.  m: Host.mesh_arrays(); ibcf_em: element-major
    flux-node ids.  Returns dict U, Q (name -> array), ont, mumax."""
    nn, nb, ne, nc, nq = m["nn"], m["nb"], m["ne"], m["nc"], m["nq"]
    nf = 4 * ne
    x, ibc, idc, ibf, xq = (np.asfortranarray(m[k]) for k in ("x", "ibc", "idc", "ibf", "xq"))
    gbn, lbn = np.ascontiguousarray(m["gbn"]), np.ascontiguousarray(m["lbn"])
    ibcf = _ibcf_cm(ibcf_em, ne)
    cols = [nb, nf, nn, nn, nb, nb, nb, nn, nn, nn]
    U = [np.zeros((nn, c), order="F") for c in cols]
    Q = [np.zeros((nf, c), order="F") for c in cols]
    diff = np.ascontiguousarray(diff, np.float64)
    qdiff = np.ascontiguousarray(qdiff, np.float64)
    ont = np.zeros(5)
    mumax = C.c_double()
    up = (PD * 10)(*[_d(a) for a in U])
    qp = (PD * 10)(*[_d(a) for a in Q])
    _mhsd_int(C.byref(q), nn, nb, ne, nc, nq, _d(x), _i(ibc), _i(idc), _i(ibf), _d(xq), _i(ibcf), _i(gbn), _i(lbn),
              _d(diff), _d(qdiff), float(beta), up, qp, _d(ont), C.byref(mumax), nthreads)
    _mhsd_bprime(nn, ne, _i(ibcf), _d(diff), _d(qdiff), float(beta), _d(U[2]), _d(U[3]), _d(Q[2]), _d(Q[3]))
    return dict(U=dict(zip(MHSD_NAMES, U)), Q=dict(zip(MHSD_NAMES, Q)), ont=ont, mumax=mumax.value)


def mhsd_sysm(m, ibcf_em, kode, qkode, nunk, nknown, diff, qdiff, mats):
    nn, nb, ne = m["nn"], m["nb"], m["ne"]
    ibcf = _ibcf_cm(ibcf_em, ne)
    A = np.zeros((nunk, nunk), order="F")
    R = np.zeros((nunk, max(nknown, 1)), order="F")
    kode, qkode = np.ascontiguousarray(kode, np.int32), np.ascontiguousarray(qkode, np.int32)
    _mhsd_sysm(nn, nb, ne, _i(ibcf), _i(np.ascontiguousarray(m["lbn"])), _i(kode), _i(qkode), nunk, nknown,
               _d(np.ascontiguousarray(diff)), _d(np.ascontiguousarray(qdiff)), _d(mats["U"]["H"]), _d(mats["U"]["G"]),
               _d(mats["Q"]["H"]), _d(mats["Q"]["G"]), _d(A), _d(R))
    return A, R


def mhsd_rhs(m, ibcf_em, kode, qkode, nunk, nknown, diff, qdiff, beta, u, qu, vel, rhsv, dgrad, mats, RHS):
    """fmSolveModHelmDC's right-hand side (dense, inline form)."""
    nn, nb, ne = m["nn"], m["nb"], m["ne"]
    ibcf = _ibcf_cm(ibcf_em, ne)
    r = np.zeros(nunk)
    up = (PD * 10)(*[_d(mats["U"][k]) for k in MHSD_NAMES])
    qp = (PD * 10)(*[_d(mats["Q"][k]) for k in MHSD_NAMES])
    vel, dgrad = np.asfortranarray(vel, np.float64), np.asfortranarray(dgrad, np.float64)
    _mhsd_rhs(nn, nb, ne, _i(ibcf), _i(np.ascontiguousarray(m["gbn"])), _i(np.ascontiguousarray(kode, np.int32)),
              _i(np.ascontiguousarray(qkode, np.int32)), nunk, nknown, _d(np.ascontiguousarray(diff)),
              _d(np.ascontiguousarray(qdiff)), float(beta), _d(np.ascontiguousarray(u)), _d(np.ascontiguousarray(qu)),
              _d(vel), _d(np.ascontiguousarray(rhsv, np.float64)), _d(dgrad), up, qp, _d(RHS), _d(r))
    return r


def mhsd_vectors(m, u, vel, rhsv, dgrad):
    """the eight product vectors of the matrix-vector variants"""
    nn, nb = m["nn"], m["nb"]
    v = [np.zeros(nn) for _ in range(5)] + [np.zeros(nb) for _ in range(3)]
    vp = (PD * 8)(*[_d(a) for a in v])
    _mhsd_vectors(nn, nb, _i(np.ascontiguousarray(m["gbn"])), _d(np.ascontiguousarray(u)),
                  _d(np.asfortranarray(vel, np.float64)), _d(np.ascontiguousarray(rhsv, np.float64)),
                  _d(np.asfortranarray(dgrad, np.float64)), vp)
    return v


def mhsd_rhs_mv(m, kode, qkode, nunk, nknown, u, qu, RHS, yu, yq):
    r = np.zeros(nunk)
    yu = [np.ascontiguousarray(a, np.float64) for a in yu]
    yq = [np.ascontiguousarray(a, np.float64) for a in yq]
    _mhsd_rhs_mv(m["nn"], m["ne"], _i(np.ascontiguousarray(kode, np.int32)), _i(np.ascontiguousarray(qkode, np.int32)),
                 nunk, nknown, _d(np.ascontiguousarray(u)), _d(np.ascontiguousarray(qu)), _d(RHS),
                 (PD * 8)(*[_d(a) for a in yu]), (PD * 8)(*[_d(a) for a in yq]), _d(r))
    return r
